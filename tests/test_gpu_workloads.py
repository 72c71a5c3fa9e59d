"""The benchmark workloads at their full sizes, on the GPU, against the oracle.

  * 1,048,576 particles with the icosphere obstacle (subdivision 5, r 1.5,
    SDF cell 0.05, padding 0.025), I = 6, S = 2, H = 2^20: one frame with the
    taps on (keys, permutation, cell table, candidate and neighbour counts
    bit-exact), then two more frames (positions within 1e-5 h).
  * The FSI coupling at 262,144 particles on the 128 x 32 quad plate (4,257
    vertices, 8,192 triangles, first column anchored): contact forces from the
    device drive the host deformation, and the SDF is refit from the deformed
    mesh every frame.  Every frame the fluid matches the oracle stepped on the
    same grid, the contact candidates are the oracle's, the contact forces
    match the oracle's; at the end the grid matches the brute-force distance
    of the deformed mesh and the anchors have not moved.
  * 256 scenes of 4,096 .. 131,072 particles (Sum 9,185,966): the batched
    handle equals the loop of single-scene handles bit for bit, and three
    scenes (the smallest, a middle one, the largest) equal the oracle.
"""
import os
import sys

import numpy as np
import pytest

import pyoracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
import bench  # noqa: E402  (the workload builders)


def _fsi():
    import fsi_pbf

    return fsi_pbf


def _say(msg):
    print(msg, flush=True)


def test_full_size_frame_with_icosphere_against_the_oracle(gpu):
    fsi = _fsi()
    cfg = bench.make_config(fsi)
    parts, sdf = bench.make_scene(fsi)
    assert len(parts) == 1 << 20 and sdf.dims[0] > 0
    _say(f"1M frame: n={len(parts)} sdf dims={sdf.dims}")
    ref = pyoracle.OracleSim(cfg, parts, taps=True, sdf=sdf, sdf_padding=0.025)
    with fsi.Solver(cfg) as s:
        s.set_taps(True)
        s.set_particles(parts)
        s.set_sdf(sdf, 0.025, False)
        s.step(1)
        keys, ids, cells = s.read_grid()
        cand, nbr = s.read_counts()
        got1 = s.read_particles()
        s.set_taps(False)
        s.step(2)
        got3 = s.read_particles()
    ref.step(1)
    assert np.array_equal(keys, ref.keys), "sorted keys"
    assert np.array_equal(ids, ref.ids), "sorted permutation"
    assert np.array_equal(cells, ref.cells), "cell table"
    assert np.array_equal(cand, ref.cand), "candidate counts"
    assert np.array_equal(nbr, ref.nbr), "neighbour counts"
    tol = 1e-5 * cfg.h
    exp1 = ref.particles()
    d1 = float(np.abs(got1[:, 0:3] - exp1[:, 0:3]).max())
    # the obstacle is in play: particles of the block start inside the sphere
    inside = np.linalg.norm(parts[:, 0:3] - np.array([6.4, 3.2, 6.4], dtype=np.float32), axis=1) < 1.5
    _say(f"1M frame 1: max|dx| {d1:.3e} (tol {tol:.1e}), {int(inside.sum())} particles start inside the sphere, "
         f"neighbours mean {nbr.mean():.1f} max {nbr.max()}")
    assert inside.sum() > 1000
    assert d1 <= tol
    ref.step(2)
    exp3 = ref.particles()
    d3 = float(np.abs(got3[:, 0:3] - exp3[:, 0:3]).max())
    _say(f"1M frame 3: max|dx| {d3:.3e}, exact {np.array_equal(got3[:, 0:3], exp3[:, 0:3])}")
    assert d3 <= tol
    # no particle is left inside the obstacle beyond the padding's reach
    r = np.linalg.norm(got3[:, 0:3] - np.array([6.4, 3.2, 6.4], dtype=np.float32), axis=1)
    assert r.min() > 1.5 - 0.1


def test_coupling_at_full_size_refit_every_frame(gpu):
    fsi = _fsi()
    verts, faces = bench.plate_mesh()
    assert len(verts) == 4257 and len(faces) == 8192
    cell, band, padding = 0.05, 0.12, 0.025
    params = fsi.deform_params()
    cfg = fsi.FluidConfig(dt=1.0 / 120.0, gravity=(0.0, -9.81, 0.0), h=0.1, rho0=6378.0, solver_iterations=6,
                          substep_iterations=2, eps=600.0, bounds_min=(0.0, 0.0, 0.0), bounds_max=(6.4, 6.4, 6.4),
                          hash_size=1 << 19)
    # the 64^3 block starts just above the plate (y = 1.0), falling at 3 m/s,
    # so that it lands within the first frames
    parts = fsi.lattice(64, 64, 64, (1.6, 1.06, 1.6), 0.05, 0.001, 1337)
    parts[:, 9] = -3.0
    anchors = np.nonzero(verts[:, 0] == verts[:, 0].min())[0]
    assert len(parts) == 262144 and len(anchors) == 33
    frames, solves, contacts = 8, 0, 0
    with fsi.Mesh(verts, faces, cell) as m, fsi.Solver(cfg) as s, fsi.Deform(verts, faces) as d:
        for a in anchors:
            d.add_anchor(int(a))
        s.set_particles(parts)
        s.set_sdf_from_mesh(m, padding=padding)
        grid = m.read_sdf()
        ref = pyoracle.OracleSim(cfg, parts, sdf=grid, sdf_padding=padding)
        s.filter_contacts(band)
        ref_c = ref.contacts(band)
        assert np.array_equal(s.read_contacts(), ref_c)
        for f in range(frames):
            cand = s.read_contacts()
            forces, (nc, nctl) = s.contact_forces(m, band, True, params.control_threshold)
            rf, (rc, rctl) = pyoracle.contact_forces(m.verts, faces, cand, band, True, params.control_threshold)
            assert (nc, nctl) == (rc, rctl), f"frame {f}: contact counts"
            if nc:
                assert np.abs(forces - rf).max() <= 1e-5 * np.abs(rf).max(), f"frame {f}: forces"
            contacts += nc
            nh, solved = d.step(forces, cfg.dt, params)
            if solved:
                solves += 1
            m.refit(d.vertices())  # SDFBoundary::refitFromMesh, every frame
            s.set_sdf_from_mesh(m, padding=padding)
            grid = m.read_sdf()
            ref.set_sdf(grid, padding)
            s.step()
            ref.step(1)
            got, exp = s.read_particles(), ref.particles()
            dx = float(np.abs(got[:, 0:3] - exp[:, 0:3]).max())
            s.filter_contacts(band)
            c_got, c_exp = s.read_contacts(), ref.contacts(band)
            _say(f"coupling frame {f}: contacts {nc} handles {nh} solved {solved} max|dx| {dx:.3e} "
                 f"candidates {len(c_got)}")
            assert dx <= 1e-5 * cfg.h, f"frame {f}: positions"
            assert np.array_equal(c_got, c_exp), f"frame {f}: contact candidates"
        cur = d.vertices()
    assert contacts > 0 and solves > 0
    # anchors stay where the deformation's float mesh put them
    assert np.array_equal(cur[anchors], verts[anchors].astype(np.float32).astype(np.float64))
    assert np.abs(cur - verts).max() > 0.0
    brute = pyoracle.mesh_sdf(cur, faces, grid.origin, grid.cell, grid.dims, verts_build=verts)
    assert np.abs(grid.grid - brute).max() <= 1e-6


def test_256_scene_batch_equals_the_loop_and_the_oracle(gpu):
    fsi = _fsi()
    scenes = bench.make_batch_scenes(fsi)
    sizes = np.array([len(sc.particles) for sc in scenes])
    assert len(scenes) == 256 and sizes.min() >= 4096 and sizes.max() <= 131072
    _say(f"batch: {len(scenes)} scenes, {int(sizes.sum())} particles")
    frames = 2
    with fsi.Batch(bench.batch_config(fsi)) as b:
        b.set_scenes(scenes)
        b.set_graph(True)
        b.step(frames)
        got = b.read_particles()
        picks = [int(np.argmin(sizes)), int(np.argsort(sizes)[len(sizes) // 2]), int(np.argmax(sizes))]
        grids = {i: b.read_grid(i) for i in picks}
    offs = np.concatenate([[0], np.cumsum(sizes)])
    for i, sc in enumerate(scenes):
        with fsi.Solver(bench.batch_config(fsi, sc.bounds_max, sc.hash_size)) as sv:
            sv.set_particles(sc.particles)
            sv.set_sdf(sc.sdf, sc.sdf_padding, False)
            sv.step(frames)
            one = sv.read_particles()
        assert np.array_equal(one, got[offs[i]:offs[i + 1]]), f"scene {i} (n={sizes[i]}) differs from its solver"
    for i in picks:
        sc = scenes[i]
        cfg = bench.batch_config(fsi, sc.bounds_max, sc.hash_size)
        ref = pyoracle.OracleSim(cfg, sc.particles, sdf=sc.sdf, sdf_padding=sc.sdf_padding)
        ref.step(frames)
        keys, ids, cells = grids[i]
        assert np.array_equal(keys, ref.keys) and np.array_equal(ids, ref.ids), f"scene {i}: sort"
        assert np.array_equal(cells, ref.cells), f"scene {i}: cell table"
        dx = float(np.abs(got[offs[i]:offs[i + 1], 0:3] - ref.particles()[:, 0:3]).max())
        _say(f"batch scene {i} (n={sizes[i]}): max|dx| vs oracle {dx:.3e}")
        assert dx <= 1e-5 * cfg.h
