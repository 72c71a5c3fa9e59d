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
