"""The hand-written adjoint (fsi_adj_*) through its torch.autograd.Function.

  * its forward is the ordinary step: bit-identical to a Solver run;
  * the backward replays every frame bit for bit (the library checks this and
    fails otherwise);
  * directional derivatives of a density-matching loss agree with FP64
    central differences of the oracle (step 1e-4) to 1e-3 relative.
"""
import numpy as np
import pytest

torch = pytest.importorskip("torch")  # before the library is loaded: one HIP runtime

from adjoint_util import bspline2_splat_np, c4_scene, central_difference, unit_direction  # noqa: E402
from conftest import dam_break_config  # noqa: E402

pytestmark = pytest.mark.gpu

RES = 8


def scene(seed=5, dims=(10, 8, 8), frames=3, n_iter=2):
    import fsi_pbf

    cfg = dam_break_config(n_iter=n_iter, substeps=2, box=(1.2, 1.2, 1.2), hash_size=1 << 12, viscosity=0.01)
    parts = fsi_pbf.lattice(*dims, (0.0, 0.0, 0.0), 0.05, 0.001, seed)
    rng = np.random.default_rng(seed)
    v0 = rng.normal(0.0, 0.1, size=(len(parts), 3)).astype(np.float32)
    parts[:, 8:11] = v0
    return cfg, parts, v0, frames


def run(frames_h, parts, v0, visc, eps, need_grad=False, dtype=torch.float32):
    import fsi_torch

    dev = torch.device("cuda:0")
    pos0 = torch.tensor(parts[:, 0:3], device=dev, dtype=dtype)
    vel0 = torch.tensor(v0, device=dev, dtype=dtype, requires_grad=need_grad)
    vt = torch.tensor(visc, dtype=torch.float64, requires_grad=need_grad)
    et = torch.tensor(eps, dtype=torch.float64, requires_grad=need_grad)
    pos, vel = fsi_torch.simulate(frames_h, pos0, vel0, vt, et)
    return pos, vel, vel0, vt, et


def test_forward_is_the_plain_step(gpu):
    import fsi_pbf
    import fsi_torch

    cfg, parts, v0, frames = scene()
    fh = fsi_torch.Frames(cfg, len(parts), frames)
    pos, vel, *_ = run(fh, parts, v0, np.float32(cfg.viscosity), np.float32(cfg.eps))
    with fsi_pbf.Solver(cfg) as s:
        s.set_particles(parts)
        s.step(frames)
        ref = s.read_particles()
    assert np.array_equal(pos.cpu().numpy(), ref[:, 0:3])
    assert np.array_equal(vel.cpu().numpy(), ref[:, 8:11])
    fh.close()


def _single_threaded():
    import os

    os.environ["OMP_NUM_THREADS"] = "1"  # before the worker loads the oracle


def _central_differences(jobs, workers, say=print):
    """central_difference over jobs (argument tuples), in `workers` processes
    (spawned, single-threaded: they load only numpy and the oracle); a line
    per finished job, through say."""
    if workers <= 1:
        return [central_difference(*j) for j in jobs]
    import multiprocessing
    from concurrent.futures import ProcessPoolExecutor, as_completed

    out = [None] * len(jobs)
    with ProcessPoolExecutor(max_workers=workers, mp_context=multiprocessing.get_context("spawn"),
                             initializer=_single_threaded) as ex:
        futs = {ex.submit(central_difference, *j): k for k, j in enumerate(jobs)}
        for done, fut in enumerate(as_completed(futs), 1):
            out[futs[fut]] = fut.result()
            say(f"central difference {done}/{len(jobs)} done")
    return out


def gradient_check(cfg, parts, v0, frames, lo, hi, res, ndirs, seed=99, precision=32, workers=1, say=print):
    """Adjoint directional derivatives of L = sum (splat(x_final) - target)^2
    against FP64 central differences (step 1e-4) along `ndirs` seeded unit
    directions in (v0, viscosity, eps), then along viscosity alone (step 1e-4
    of its value) and eps alone (step 1e-6 of its value: on the 640-particle
    scene the loss has a kink in eps within 1e-4 of its value, where some
    branch of the step -- a clamp or floor -- switches; differences at 1e-5
    and below agree to 2e-8).  Returns the relative errors."""
    import fsi_torch

    visc, eps = float(np.float32(cfg.viscosity)), float(np.float32(cfg.eps))
    fh = fsi_torch.Frames(cfg, len(parts), frames, precision=precision)
    # the FP64 trajectory keeps states and loss in FP64 too
    dtype = torch.float64 if precision == 64 else torch.float32
    rng = np.random.default_rng(seed)
    v_t = (v0 + rng.normal(0.0, 0.05, size=v0.shape)).astype(np.float32)
    with torch.no_grad():
        pos_t, *_ = run(fh, parts, v_t, visc, eps, dtype=dtype)
    target = bspline2_splat_np(pos_t.cpu().numpy(), lo, hi, res)
    pos, vel, vel0, vt, et = run(fh, parts, v0, visc, eps, need_grad=True, dtype=dtype)
    tgt = torch.tensor(target, device=pos.device, dtype=dtype)
    loss = ((fsi_torch.density_splat(pos, lo, hi, res) - tgt) ** 2).sum()
    loss.backward()
    st = fh.stats()
    assert st["replay_mismatches"] == 0 and st["capacity_bits"] == 0
    fh.close()
    gv = vel0.grad.cpu().numpy().astype(np.float64)
    gvisc, geps = float(vt.grad), float(et.grad)
    v64 = v0.astype(np.float64)
    base = (cfg, parts, v64, visc, eps, frames, target, lo, hi, res)
    names, adjs, jobs = [], [], []
    for k in range(ndirs):
        dv, dvisc, deps = unit_direction(np.random.default_rng(1000 + k), len(parts))
        names.append(f"direction {k}")
        adjs.append(float((gv * dv).sum() + gvisc * dvisc + geps * deps))
        jobs.append(base + (dv, dvisc, deps, 1e-4))
    for name, dvisc, deps, step in (("viscosity", 1.0, 0.0, 1e-4 * visc), ("eps", 0.0, 1.0, 1e-6 * eps)):
        names.append(name)
        adjs.append(gvisc * dvisc + geps * deps)
        jobs.append(base + (np.zeros_like(v64), dvisc, deps, step))
    errs = []
    for name, adj, fd in zip(names, adjs, _central_differences(jobs, workers, say)):
        rel = abs(adj - fd) / max(abs(adj), abs(fd))
        errs.append(rel)
        print(f"{name}: adjoint {adj:+.8e} fd {fd:+.8e} rel {rel:.2e}", flush=True)
    print(f"precision {precision} loss {loss.item():.6e} stats {st} max rel {max(errs):.2e}", flush=True)
    return errs


@pytest.mark.parametrize("n_iter", [2, 4])
def test_gradient_matches_fp64_central_differences(gpu, n_iter):
    cfg, parts, v0, frames = scene(n_iter=n_iter)
    errs = gradient_check(cfg, parts, v0, frames, (0.0, 0.0, 0.0), (1.2, 1.2, 1.2), RES, 6)
    assert max(errs) <= 1e-3, errs


@pytest.mark.parametrize("n_iter", [2, 4])
def test_fp64_trajectory_gradient_matches_central_differences(gpu, n_iter):
    """FP64 trajectory, states and loss: the adjoint and the differences see
    the same FP64 run (up to summation order), so they agree to the
    differences' own error."""
    cfg, parts, v0, frames = scene(n_iter=n_iter)
    errs = gradient_check(cfg, parts, v0, frames, (0.0, 0.0, 0.0), (1.2, 1.2, 1.2), RES, 6, precision=64)
    assert max(errs) <= 1e-6, errs


def _workers():
    import os

    return max(1, min(16, (os.cpu_count() or 2) - 1))


def _say_uncaptured(capsys):
    """A print that reaches the terminal while the test runs: the central
    differences of the 65,536-particle workload take minutes of CPU time, and
    a run that prints nothing for that long looks hung."""

    def say(msg):
        with capsys.disabled():
            print(msg, flush=True)

    return say


def test_differentiable_workload_65k_gradient_fp32_trajectory(gpu, capsys):
    """65,536 particles, 5 frames x 4 iterations x 2 substeps, 32^3 density
    loss, 16 seeded directions plus viscosity and eps.  The adjoint is exact
    for the FP32 trajectory; FP64 central differences see the FP64
    trajectory, and the FP32 rounding of the positions alone moves the FP64
    derivative by ~6e-4 per 1e-7 of noise per frame, so the bar here is 1e-2
    and the median 1e-3."""
    cfg, parts, v0, frames, lo, hi, res = c4_scene(frames=5)
    errs = gradient_check(cfg, parts, v0, frames, lo, hi, res, 16, workers=_workers(), say=_say_uncaptured(capsys))
    assert max(errs) <= 1e-2, errs
    assert float(np.median(errs)) <= 1e-3, errs


def test_differentiable_workload_65k_20_frames_fp64_trajectory(gpu, capsys):
    """The differentiable-PBF workload at full length: 65,536 particles, 20
    frames x 4 iterations x 2 substeps, 32^3 density loss; the FP64
    trajectory's adjoint against FP64 central differences (step 1e-4) along
    16 seeded unit directions plus viscosity and eps, each to 1e-3."""
    cfg, parts, v0, frames, lo, hi, res = c4_scene(frames=20)
    errs = gradient_check(cfg, parts, v0, frames, lo, hi, res, 16, precision=64, workers=_workers(),
                          say=_say_uncaptured(capsys))
    assert max(errs) <= 1e-3, errs


@pytest.mark.parametrize("precision", [32, 64])
def test_backward_is_bit_reproducible(gpu, precision):
    """The transposed rows fill in atomic order, which differs between runs;
    the gradient must not."""
    import fsi_torch

    cfg, parts, v0, frames = scene(dims=(12, 10, 8), frames=3, n_iter=3)
    dtype = torch.float64 if precision == 64 else torch.float32
    fh = fsi_torch.Frames(cfg, len(parts), frames, precision=precision)
    grads = []
    for _ in range(3):
        pos, vel, vel0, vt, et = run(fh, parts, v0, 0.01, 600.0, need_grad=True, dtype=dtype)
        (pos.square().sum() + vel.sin().sum()).backward()
        grads.append((vel0.grad.cpu().numpy().copy(), float(vt.grad), float(et.grad)))
    fh.close()
    for g in grads[1:]:
        assert np.array_equal(g[0], grads[0][0]) and g[1] == grads[0][1] and g[2] == grads[0][2]


def test_backward_needs_its_forward(gpu):
    import fsi_torch

    cfg, parts, v0, frames = scene(dims=(4, 4, 4), frames=1)
    fh = fsi_torch.Frames(cfg, len(parts), frames)
    pos, _, vel0, _, _ = run(fh, parts, v0, 0.01, 600.0, need_grad=True)
    run(fh, parts, v0, 0.01, 600.0)  # another forward on the same handle
    with pytest.raises(RuntimeError):
        pos.sum().backward()
    fh.close()
