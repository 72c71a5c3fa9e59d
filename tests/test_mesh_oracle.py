"""Pins of the FSI-coupling oracle (oracle/mesh_oracle.c): signed distance,
dirty-voxel marking and contact forces against closed forms and textbook
geometry, and the host SDF build (fsi_sdf_build) against the oracle."""
import numpy as np
import pytest

import pyoracle

TRI = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
TRI_F = np.array([[0, 1, 2]], dtype=np.int32)


def point_sdf(verts, faces, p):
    return float(pyoracle.mesh_sdf(verts, faces, p, 1.0, (1, 1, 1))[0, 0, 0])


@pytest.mark.parametrize("p, dist, sign", [
    ((0.25, 0.25, 0.5), 0.5, 1.0),                  # face region, above (normal +z)
    ((0.25, 0.25, -0.5), 0.5, -1.0),                # face region, below
    ((-1.0, -1.0, 0.0), np.sqrt(2.0), 1.0),         # vertex A region, in the plane
    ((0.5, -1.0, 0.0), 1.0, 1.0),                   # edge AB region
    ((1.0, 1.0, 0.3), np.sqrt(0.59), 1.0),          # edge BC region, closest point (0.5, 0.5, 0)
    ((0.2, 0.3, -0.25), 0.25, -1.0),                # face, below
])
def test_single_triangle_feature_regions(p, dist, sign):
    d = point_sdf(TRI, TRI_F, p)
    assert abs(abs(d) - dist) <= 1e-6 * max(1.0, dist)
    assert np.sign(d) == sign


def test_icosphere_distance_is_the_sphere_within_the_chord_error():
    import fsi_pbf

    verts, faces = fsi_pbf.icosphere(3, (0.0, 0.0, 0.0), 1.0)
    cell = 0.2
    origin = np.array([-1.5, -1.5, -1.5])
    dims = (16, 16, 16)
    grid = pyoracle.mesh_sdf(verts, faces, origin, cell, dims)
    z, y, x = np.meshgrid(np.arange(16), np.arange(16), np.arange(16), indexing="ij")
    p = np.stack([origin[0] + x * cell, origin[1] + y * cell, origin[2] + z * cell], axis=-1)
    exact = np.linalg.norm(p, axis=-1) - 1.0
    # chord error of a level-3 icosphere: 1 - cos(half the edge angle) < 0.01
    assert np.abs(grid - exact).max() < 0.01
    assert (grid[exact < -0.05] < 0).all() and (grid[exact > 0.05] > 0).all()


def test_host_sdf_build_matches_the_oracle():
    import fsi_pbf

    verts, faces = fsi_pbf.icosphere(2, (0.1, -0.2, 0.05), 0.7)
    sdf = fsi_pbf.sdf_build(verts, faces, 0.1)
    ref = pyoracle.mesh_sdf(verts, faces, sdf.origin, sdf.cell, sdf.dims)
    diff = np.abs(sdf.grid - ref)
    # equidistant triangles (a point nearest a shared edge or vertex) can
    # give distances one rounding apart
    assert diff.max() <= 1e-6
    assert (np.sign(sdf.grid) == np.sign(ref))[np.abs(ref) > 1e-6].all()


def test_dirty_mask_of_one_moved_vertex():
    old = np.array([[0.5, 0.5, 0.5], [0.2, 0.2, 0.2]])
    new = np.array([[0.55, 0.5, 0.5], [0.2, 0.2, 0.2 + 1e-9]])   # the second moves less than eps
    n, mask, box = pyoracle.dirty_mask(old, new, 0.12, 1e-6, (0.0, 0.0, 0.0), 0.1, (10, 10, 10))
    # old band: floor(3.8)..ceil(6.2) = 3..7 per axis; new band x: 4..7 (inside)
    assert n == 125
    assert list(box) == [3, 3, 3, 7, 7, 7]
    assert mask[3:8, 3:8, 3:8].all() and mask.sum() == 125


def test_dirty_mask_clamps_to_the_grid():
    old = np.array([[0.0, 0.0, 0.0]])
    new = np.array([[0.05, 0.0, 0.0]])
    n, mask, box = pyoracle.dirty_mask(old, new, 0.15, 0.0, (0.0, 0.0, 0.0), 0.1, (4, 4, 4))
    # old: floor(-1.5) = -2 -> 0 .. ceil(1.5) = 2; new x: floor(-1) -> 0 .. ceil(2) = 2
    assert list(box) == [0, 0, 0, 2, 2, 2] and n == 27


def plane_candidates():
    # a large triangle in z = 0 with normal +z
    verts = np.array([[-10.0, -10.0, 0.0], [10.0, -10.0, 0.0], [0.0, 10.0, 0.0]])
    faces = np.array([[0, 1, 2]], dtype=np.int32)
    cand = np.zeros((4, 8), dtype=np.float32)
    cand[0, 0:3] = (0.5, 0.25, 0.05)
    cand[0, 4:7] = (0.1, 0.0, -2.0)       # moving into the surface
    cand[1, 0:3] = (-1.0, 2.0, 0.08)
    cand[1, 4:7] = (0.0, 0.0, 1.0)        # moving away: no contact by velocity
    cand[2, 0:3] = (0.0, 0.0, 0.5)        # outside the band
    cand[3, 0:3] = (3.0, -4.0, -0.02)     # below: the push points up
    cand[3, 4:7] = (0.0, 0.0, 1.5)
    return verts, faces, cand


def bary(p):
    # p = A + s (B - A) + t (C - A) for A = (-10, -10), B - A = (20, 0), C - A = (10, 20)
    t = (p[1] + 10.0) / 20.0
    s = (p[0] + 10.0 - 10.0 * t) / 20.0
    return np.array([1.0 - s - t, s, t])


@pytest.mark.parametrize("use_vel", [True, False])
def test_contact_forces_closed_form_on_a_plane(use_vel):
    verts, faces, cand = plane_candidates()
    forces, (contacts, controls) = pyoracle.contact_forces(verts, faces, cand, 0.1, use_vel, 0.5)
    expect = np.zeros((3, 3))
    n = 0
    for c in cand:
        p, v = c[0:3].astype(np.float64), c[4:7].astype(np.float64)
        if abs(p[2]) > 0.1:
            continue
        d = np.array([0.0, 0.0, -np.sign(p[2])])
        mag = max(0.0, float(v @ d)) if use_vel else 1.0
        if mag <= 0.0:
            continue
        expect += np.outer(bary(p), d * mag)
        n += 1
    assert contacts == n == (2 if use_vel else 3)
    assert np.allclose(forces, expect, rtol=1e-5, atol=1e-6)
    # the split conserves the push: the weights of a point sum to one
    assert abs(forces.sum(axis=0) - expect.sum(axis=0)).max() < 1e-5
    assert controls == int((np.linalg.norm(expect, axis=1) >= 0.5).sum())


@pytest.mark.parametrize("p, q", [
    ((0.25, 0.25, 0.5), (0.25, 0.25, 0.0)),         # face region: the foot of the perpendicular
    ((-1.0, -1.0, 0.0), (0.0, 0.0, 0.0)),           # vertex A region
    ((2.0, -0.5, 0.25), (1.0, 0.0, 0.0)),           # vertex B region
    ((0.5, -1.0, 0.0), (0.5, 0.0, 0.0)),            # edge AB region
    ((1.0, 1.0, 0.3), (0.5, 0.5, 0.0)),             # edge BC region
    ((-0.5, 0.5, -2.0), (0.0, 0.5, 0.0)),           # edge CA region
])
def test_closest_triangle_feature_regions(p, q):
    """closestTriangle on one triangle: the closest point of each Voronoi
    region (Ericson 5.1.5) in closed form, its squared distance, face 0."""
    face, pt, d2 = pyoracle.closest_triangles(TRI, TRI_F, np.array([p], dtype=np.float32))
    assert face[0] == 0
    assert np.allclose(pt[0], q, atol=1e-7)
    exp = float(np.sum((np.asarray(p, dtype=np.float32).astype(np.float64) - np.asarray(q)) ** 2))
    assert abs(d2[0] - exp) <= 1e-12 * max(1.0, exp)


def test_closest_triangle_of_a_cube_is_the_nearest_face():
    """Points inside and outside a unit cube (12 triangles): the closest point
    is the projection on the nearest face plane (clamped to the square), and
    the reported face lies in that plane."""
    v = np.array([[x, y, z] for z in (0.0, 1.0) for y in (0.0, 1.0) for x in (0.0, 1.0)])
    quads = [(0, 2, 3, 1), (4, 5, 7, 6), (0, 1, 5, 4), (2, 6, 7, 3), (0, 4, 6, 2), (1, 3, 7, 5)]
    f = np.array([t for a, b, c, d in quads for t in ((a, b, c), (a, c, d))], dtype=np.int32)
    rng = np.random.default_rng(4)
    pts = rng.uniform(-0.5, 1.5, size=(400, 3)).astype(np.float32)
    face, got, d2 = pyoracle.closest_triangles(v, f, pts)
    p64 = pts.astype(np.float64)
    inside = np.all((p64 > 0) & (p64 < 1), axis=1)
    exp = np.clip(p64, 0.0, 1.0)
    # inside: move the nearest coordinate onto its face
    dist = np.concatenate([p64, 1.0 - p64], axis=1)
    k = np.argmin(dist, axis=1)
    for i in np.nonzero(inside)[0]:
        exp[i, k[i] % 3] = 0.0 if k[i] < 3 else 1.0
    assert np.allclose(got, exp, atol=2e-7)
    assert np.allclose(d2, np.sum((p64 - exp) ** 2, axis=1), atol=1e-12)
    for i in range(len(pts)):
        tri = v[f[face[i]]]
        normal = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        assert abs(np.dot(normal, exp[i] - tri[0])) <= 1e-9
