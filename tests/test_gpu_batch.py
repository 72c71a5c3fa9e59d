"""Parity of the multi-scene solver (fsi_batch_*): every scene of a batch must
come out exactly as the same scene run alone by a single-scene handle (which is
itself checked against the oracle), and the first scenes are also checked
against the oracle directly.  Bars as in test_gpu_parity.py: keys, permutation
and cell table bit-exact; positions within 1e-5*h (expected equal)."""
import numpy as np
import pytest

from conftest import dam_break_config

pytestmark = pytest.mark.gpu


def _fsi():
    import fsi_pbf

    return fsi_pbf


def make_scenes(seed=7, count=6, with_sdf=True):
    """Ragged scenes: different lattice sizes, hash sizes, boxes, with and without an SDF sphere."""
    fsi = _fsi()
    rng = np.random.default_rng(seed)
    scenes = []
    for s in range(count):
        nx, ny, nz = (int(v) for v in rng.integers(3, 19, size=3))
        box = (float(rng.uniform(1.2, 2.4)), float(rng.uniform(1.0, 2.0)), float(rng.uniform(1.0, 2.0)))
        H = 1 << int(rng.integers(4, 16))
        parts = fsi.lattice(nx, ny, nz, (0.05, 0.05, 0.05), 0.05, 0.001, 1000 + s)
        sdf, pad, cont = None, 0.0, False
        if with_sdf and s % 2 == 1:
            verts, faces = fsi.icosphere(2, (box[0] * 0.6, 0.3, box[2] * 0.4), float(rng.uniform(0.2, 0.4)))
            sdf = fsi.sdf_build(verts, faces, 0.05)
            pad = 0.02
        scenes.append(fsi.Scene(particles=parts, hash_size=H, bounds_min=(0.0, 0.0, 0.0), bounds_max=box, sdf=sdf,
                                sdf_padding=pad, sdf_containment=cont))
    return scenes


def scene_cfg(base, sc):
    cfg = dam_break_config(n_iter=base.solver_iterations, substeps=base.substep_iterations, box=sc.bounds_max,
                           hash_size=sc.hash_size)
    cfg.enable_vorticity = base.enable_vorticity
    return cfg


def run_single(base, sc, frames):
    fsi = _fsi()
    s = fsi.Solver(scene_cfg(base, sc))
    s.set_particles(sc.particles)
    if sc.sdf is not None:
        s.set_sdf(sc.sdf, sc.sdf_padding, sc.sdf_containment)
    s.step(frames)
    out = s.read_particles()
    grid = s.read_grid()
    s.close()
    return out, grid


@pytest.mark.parametrize("frames,graph,gather,block,tile", [(1, False, 1, 256, 0), (3, True, 1, 256, 0),
                                                             (3, True, 2, 256, 0), (2, False, 0, 128, 0),
                                                             (2, True, 1, 128, 0), (2, False, 2, 256, 1),
                                                             (3, True, 2, 256, 2), (3, True, 1, 256, -2),
                                                             (2, False, 2, 256, -1)])
def test_batch_scenes_equal_single_scene_runs(gpu, frames, graph, gather, block, tile):
    fsi = _fsi()
    base = dam_break_config(n_iter=3, substeps=2)
    scenes = make_scenes()
    b = fsi.Batch(base)
    b.set_tuning(fsi.TUNE_GATHER, gather)
    b.set_tuning(fsi.TUNE_BLOCK, block)
    if tile < 0:  # relabelled storage (processing order 2), tile shape -tile - 1
        b.set_tuning(fsi.TUNE_ORDER, 2)
        tile = -tile - 1
    b.set_tuning(fsi.TUNE_TILE, tile)
    b.set_graph(graph)
    b.set_scenes(scenes)
    b.step(frames)
    allp = b.read_particles()
    for i, sc in enumerate(scenes):
        exp, (k, ids, cells) = run_single(base, sc, frames)
        got = b.scene_particles(allp, i)
        gk, gi, gc = b.read_grid(i)
        print(f"scene {i}: n={len(got)} H={sc.hash_size} sdf={sc.sdf is not None} "
              f"max|dx|={np.abs(got[:, 0:3] - exp[:, 0:3]).max():.3e}")
        assert np.array_equal(gk, k), f"scene {i} keys"
        assert np.array_equal(gi, ids), f"scene {i} permutation"
        assert np.array_equal(gc, cells), f"scene {i} cell table"
        assert np.abs(got[:, 0:3] - exp[:, 0:3]).max() <= 1e-5 * base.h
        assert np.array_equal(got, exp), f"scene {i} particle records"
    b.close()


def test_batch_against_oracle(gpu):
    fsi = _fsi()
    import pyoracle

    base = dam_break_config(n_iter=3, substeps=1)
    base.enable_vorticity = False
    scenes = make_scenes(seed=11, count=3)
    b = fsi.Batch(base)
    b.set_scenes(scenes)
    b.step(1)
    allp = b.read_particles()
    for i, sc in enumerate(scenes):
        cfg = scene_cfg(base, sc)
        ref = pyoracle.OracleSim(cfg, sc.particles, taps=False, sdf=sc.sdf, sdf_padding=sc.sdf_padding,
                                 containment=sc.sdf_containment)
        ref.step(1)
        keys, ids, cells = b.read_grid(i)
        assert np.array_equal(keys, ref.keys) and np.array_equal(ids, ref.ids) and np.array_equal(cells, ref.cells)
        got = b.scene_particles(allp, i)
        assert np.abs(got[:, 0:3] - ref.particles()[:, 0:3]).max() <= 1e-5 * cfg.h
    b.close()


def test_batch_autotune_restores_state_and_graph_matches(gpu):
    fsi = _fsi()
    base = dam_break_config(n_iter=2, substeps=2)
    scenes = make_scenes(seed=3, count=4)
    ref = fsi.Batch(base)
    ref.set_scenes(scenes)
    ref.step(2)
    want = ref.read_particles()
    ref.close()
    b = fsi.Batch(base)
    b.set_scenes(scenes)
    gather, bits = b.autotune(1)
    assert gather in (0, 1, 2) and bits in (6, 8, 10)
    k = b.knobs()
    assert k["gather"] == gather and k["sort_digit_bits"] == bits
    assert k["block"] in (128, 256) and k["tile"] in (0, 1, 2), k
    b.set_graph(True)
    b.set_timing(True)
    b.step(2)
    assert np.array_equal(b.read_particles(), want)
    assert b.read_timing()["frames"] == 2  # graph replays: frames counted, passes not timed
    b.set_graph(False)
    b.timing_reset()
    b.step(1)
    t = b.read_timing()
    assert t["frames"] == 1 and t["launches"]["lambda"] == 2 * 2 and t["ms"]["lambda"] > 0.0
    b.close()


def test_batch_error_paths(gpu):
    fsi = _fsi()
    base = dam_break_config()
    b = fsi.Batch(base)
    parts = fsi.lattice(4, 4, 4, (0.1, 0.1, 0.1), 0.05, 0.0)
    bad_hash = fsi.Scene(particles=parts, hash_size=1000, bounds_min=(0, 0, 0), bounds_max=(1, 1, 1))
    with pytest.raises(fsi.FsiError):
        b.set_scenes([bad_hash])
    assert b.lib.fsi_batch_read_grid(b.h, 0, None, None, None) != 0  # no scenes set
    ok = fsi.Scene(particles=parts, hash_size=64, bounds_min=(0, 0, 0), bounds_max=(1, 1, 1))
    b.set_scenes([ok, ok])
    with pytest.raises(fsi.FsiError):
        b.read_grid(1)  # no step yet
    b.step(1)
    k0, _, _ = b.read_grid(0)
    k1, _, _ = b.read_grid(1)
    assert np.array_equal(k0, k1)
    assert b.lib.fsi_batch_read_grid(b.h, 2, None, None, None) != 0  # scene out of range
    b.close()


def test_single_scene_handles_on_one_stream(gpu):
    """fsi_pbf_set_stream: handles that share one stream (the reference's
    single-queue order) step exactly as on their own streams; NULL restores
    the handle's own stream."""
    fsi = _fsi()
    base = dam_break_config()
    scenes = make_scenes(seed=11, count=3)
    expect = [run_single(base, sc, 2)[0] for sc in scenes]
    solvers = []
    for sc in scenes:
        s = fsi.Solver(scene_cfg(base, sc))
        s.set_particles(sc.particles)
        if sc.sdf is not None:
            s.set_sdf(sc.sdf, sc.sdf_padding, sc.sdf_containment)
        solvers.append(s)
    own = [s.get_stream() for s in solvers]
    assert len(set(own)) == len(own) and all(own)
    shared = own[0]
    for s in solvers[1:]:
        s.set_stream(shared)
        assert s.get_stream() == shared
    for _ in range(2):
        for s in solvers:
            s.step()
    for s, e in zip(solvers, expect):
        assert np.array_equal(s.read_particles(), e)
    for s, o in zip(solvers[1:], own[1:]):
        s.set_stream(0)
        assert s.get_stream() == o
    solvers[1].step()  # on its own stream again, while solver 0 still owns the shared one
    solvers[1].sync()
    for s in solvers:
        s.close()
