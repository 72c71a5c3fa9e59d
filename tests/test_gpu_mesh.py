"""The FSI coupling on the device (fsi_mesh_*, fsi_pbf_set_sdf_from_mesh,
fsi_pbf_contact_forces) against the host SDF build and the CPU oracle."""
import numpy as np
import pytest

import pyoracle

pytestmark = pytest.mark.gpu


def plate(nx=32, ny=8, size=(4.0, 1.0), origin=(0.0, 0.0, 0.0)):
    """A flat nx x ny quad grid in the plane z = origin[2], two triangles per
    quad, normals +z."""
    xs = np.linspace(0.0, size[0], nx + 1)
    ys = np.linspace(0.0, size[1], ny + 1)
    X, Y = np.meshgrid(xs, ys, indexing="xy")
    verts = np.stack([X.ravel() + origin[0], Y.ravel() + origin[1], np.full(X.size, float(origin[2]))], axis=1)
    faces = []
    for j in range(ny):
        for i in range(nx):
            a = j * (nx + 1) + i
            b, c, d = a + 1, a + nx + 1, a + nx + 2
            faces += [[a, b, d], [a, d, c]]
    return verts, np.array(faces, dtype=np.int32)


def meshes():
    import fsi_pbf

    yield "icosphere", fsi_pbf.icosphere(3, (0.3, -0.1, 0.2), 0.8), 0.08
    yield "plate", plate(), 0.1


@pytest.mark.parametrize("which", ["icosphere", "plate"])
def test_device_grid_is_the_host_build(gpu, which):
    import fsi_pbf

    (verts, faces), cell = [(m, c) for n, m, c in meshes() if n == which][0]
    host = fsi_pbf.sdf_build(verts, faces, cell)
    with fsi_pbf.Mesh(verts, faces, cell) as m:
        dev = m.read_sdf()
    assert dev.dims == host.dims and np.array_equal(dev.origin, host.origin)
    # same tree, same FP64 operations: identical
    assert np.array_equal(dev.grid, host.grid)


@pytest.mark.parametrize("which", ["icosphere", "plate"])
def test_closest_triangle_matches_the_oracle(gpu, which):
    """SDFBoundary::closestTriangle on the device (BVH walk) against the
    brute-force oracle, on the built and on a refit (bent) mesh: the closest
    points agree to float rounding, the distances to 1e-6, and the faces
    wherever no other triangle is as close."""
    import fsi_pbf

    (verts, faces), cell = [(m, c) for n, m, c in meshes() if n == which][0]
    rng = np.random.default_rng(17)
    lo, hi = verts.min(0) - 0.5, verts.max(0) + 0.5
    pts = rng.uniform(lo, hi, size=(5000, 3)).astype(np.float32)
    with fsi_pbf.Mesh(verts, faces, cell) as m:
        for stage in ("built", "refit"):
            cur = verts
            if stage == "refit":
                cur = verts.copy()
                cur[:, 2] += 0.1 * np.sin(np.pi * cur[:, 0]) * cur[:, 1]
                m.refit(cur)
            face, got = m.closest(pts)
            of, op, od2 = pyoracle.closest_triangles(cur, faces, pts)
            d2 = np.sum((pts.astype(np.float64) - got.astype(np.float64)) ** 2, axis=1)
            print(f"{which} {stage}: max |p - p_oracle| = {np.abs(got - op).max():.2e}, "
                  f"faces equal {np.mean(face == of) * 100:.2f}%", flush=True)
            assert face.min() >= 0 and face.max() < len(faces)
            assert np.abs(got - op).max() <= 2e-6 * max(1.0, float(np.abs(op).max()))
            assert np.abs(np.sqrt(d2) - np.sqrt(od2)).max() <= 1e-6
            # a different face only where it is as close as the oracle's (a shared edge or vertex)
            for i in np.nonzero(face != of)[0]:
                _, _, d2f = pyoracle.closest_triangles(cur, faces[face[i]][None], pts[i][None])
                assert abs(np.sqrt(d2f[0]) - np.sqrt(od2[i])) <= 1e-12, (i, face[i], of[i])
        face, got = m.closest(np.zeros((0, 3), dtype=np.float32))
        assert len(face) == 0


def test_refit_matches_the_oracle_with_build_normals(gpu):
    import fsi_pbf

    verts, faces = plate()
    cell = 0.1
    with fsi_pbf.Mesh(verts, faces, cell) as m:
        # bend the plate: a smooth vertical bump, as a deformation does
        moved = verts.copy()
        moved[:, 2] += 0.15 * np.sin(np.pi * moved[:, 0] / 4.0) * moved[:, 1]
        m.refit(moved)
        dev = m.read_sdf()
    ref = pyoracle.mesh_sdf(moved, faces, dev.origin, dev.cell, dev.dims, verts_build=verts)
    assert np.abs(dev.grid - ref).max() <= 1e-6
    assert (np.sign(dev.grid) == np.sign(ref))[np.abs(ref) > 1e-6].all()


def test_dirty_refit_recomputes_exactly_the_marked_voxels(gpu):
    import fsi_pbf

    verts, faces = plate()
    cell, band, eps = 0.1, 0.25, 1e-4
    moved = verts.copy()
    sel = (np.abs(verts[:, 0] - 3.0) < 0.3) & (verts[:, 1] > 0.4)
    moved[sel, 2] -= 0.08
    with fsi_pbf.Mesh(verts, faces, cell) as m, fsi_pbf.Mesh(verts, faces, cell) as full:
        before = m.read_sdf().grid
        n, lo, hi = m.refit_dirty(moved, band, eps)
        after = m.read_sdf()
        full.refit(moved)
        ref = full.read_sdf().grid
    cnt, mask, box = pyoracle.dirty_mask(verts, moved, band, eps, after.origin, after.cell, after.dims)
    assert n == cnt > 0
    assert list(lo) == list(box[0:3]) and list(hi) == list(box[3:6])
    marked = mask.astype(bool)
    assert np.array_equal(after.grid[marked], ref[marked])
    assert np.array_equal(after.grid[~marked], before[~marked])
    # a second call with no motion recomputes nothing
    with fsi_pbf.Mesh(verts, faces, cell) as m2:
        n2, lo2, hi2 = m2.refit_dirty(verts, band, eps)
    assert n2 == 0 and (lo2 > hi2).all()


def fluid_near_sphere(centre, radius, n_side=24, seed=7):
    import fsi_pbf
    from conftest import dam_break_config

    cfg = dam_break_config(n_iter=2, substeps=1, box=(4.0, 4.0, 4.0), hash_size=1 << 14)
    spacing = 2.2 * radius / n_side
    lo = np.asarray(centre) - 1.1 * radius
    parts = fsi_pbf.lattice(n_side, n_side, n_side, lo, spacing, 0.1 * spacing, seed)
    rng = np.random.default_rng(seed)
    parts[:, 8:11] = rng.normal(0.0, 0.5, size=(len(parts), 3)).astype(np.float32)
    return cfg, parts


@pytest.mark.parametrize("use_vel", [True, False])
def test_contact_forces_match_the_oracle(gpu, use_vel):
    import fsi_pbf

    centre, radius = (2.0, 2.0, 2.0), 0.6
    verts, faces = fsi_pbf.icosphere(3, centre, radius)
    cfg, parts = fluid_near_sphere(centre, radius)
    band = 0.05
    with fsi_pbf.Mesh(verts, faces, 0.05) as m, fsi_pbf.Solver(cfg) as s:
        s.set_particles(parts)
        s.set_sdf_from_mesh(m)
        s.filter_contacts(band)
        cand = s.read_contacts()
        forces, (contacts, controls) = s.contact_forces(m, band, use_vel, 0.05)
        again, counts2 = s.contact_forces(m, band, use_vel, 0.05)
        stored = m.read_forces()
    assert len(cand) > 100
    ref, (rc, rctl) = pyoracle.contact_forces(verts, faces, cand, band, use_vel, 0.05)
    assert contacts == rc and controls == rctl
    scale = np.abs(ref).max()
    assert np.abs(forces - ref).max() <= 1e-5 * scale
    assert np.array_equal(forces, again) and counts2 == (contacts, controls)  # a fixed order
    assert np.array_equal(forces, stored)


def test_solver_steps_identically_on_the_mesh_grid(gpu):
    import fsi_pbf

    centre, radius = (2.0, 2.0, 2.0), 0.6
    verts, faces = fsi_pbf.icosphere(2, centre, radius)
    cfg, parts = fluid_near_sphere(centre, radius, n_side=12)
    host = fsi_pbf.sdf_build(verts, faces, 0.05)
    out = []
    with fsi_pbf.Mesh(verts, faces, 0.05) as m:
        for use_mesh in (True, False):
            with fsi_pbf.Solver(cfg) as s:
                s.set_particles(parts)
                if use_mesh:
                    s.set_sdf_from_mesh(m, padding=0.01)
                else:
                    s.set_sdf(host, padding=0.01)
                s.step(3)
                out.append(s.read_particles())
    assert np.array_equal(out[0], out[1])


def test_contact_forces_need_candidates(gpu):
    import fsi_pbf

    verts, faces = plate(4, 2)
    cfg, parts = fluid_near_sphere((1.0, 0.5, 0.0), 0.5, n_side=6)
    with fsi_pbf.Mesh(verts, faces, 0.1) as m, fsi_pbf.Solver(cfg) as s:
        s.set_particles(parts)
        with pytest.raises(fsi_pbf.FsiError):
            s.contact_forces(m, 0.05)


def test_two_way_coupling_frames(gpu):
    """Contact forces from the device drive the host deformation
    (deformStep), whose mesh refits the device SDF: anchors stay, the plate
    is pushed along the fluid's impact, and the grid follows the mesh."""
    import fsi_pbf
    from conftest import dam_break_config

    verts, faces = plate(32, 8, (3.2, 0.8), (0.8, 0.8, 0.8))
    verts = verts[:, [0, 2, 1]].copy()  # the plate in y = 0.8 (x 0.8..4.0, z 0.8..1.6), normals +y
    faces = faces[:, [0, 2, 1]].copy()
    anchors = np.nonzero(verts[:, 0] == verts[:, 0].min())[0]
    cfg = dam_break_config(n_iter=3, substeps=2, box=(4.8, 3.2, 3.2), hash_size=1 << 16)
    parts = fsi_pbf.lattice(32, 16, 12, (1.2, 1.0, 1.0), 0.05, 0.001, 3)
    parts[:, 8:11] = (0.0, -4.0, 0.0)  # falling onto the plate
    params = fsi_pbf.deform_params()
    with fsi_pbf.Mesh(verts, faces, 0.05) as m, fsi_pbf.Solver(cfg) as s, fsi_pbf.Deform(verts, faces) as d:
        for a in anchors:
            d.add_anchor(int(a))
        s.set_particles(parts)
        s.set_sdf_from_mesh(m, padding=0.025)
        s.filter_contacts(0.12)
        moved = 0
        for _ in range(12):
            forces, (nc, _) = s.contact_forces(m, 0.12, True, params.control_threshold)
            nh, solved = d.step(forces, cfg.dt, params)
            if solved:
                m.refit(d.vertices())
                s.set_sdf_from_mesh(m, padding=0.025)
                moved += 1
            s.step()
            s.filter_contacts(0.12)
        cur = d.vertices()
        grid = m.read_sdf()
    assert moved > 0
    # anchors stay where the deformation's float mesh put them
    assert np.array_equal(cur[anchors], verts[anchors].astype(np.float32).astype(np.float64))
    disp = cur - verts
    assert disp[:, 1].min() < -1e-3  # pushed down where the fluid lands
    assert np.abs(disp).max() <= 0.4 + 1e-6  # the total-displacement clamp (float vertices)
    ref = pyoracle.mesh_sdf(cur, faces, grid.origin, grid.cell, grid.dims, verts_build=verts)
    assert np.abs(grid.grid - ref).max() <= 1e-6
