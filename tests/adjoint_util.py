"""Helpers of the adjoint tests: the FP64 loss through the oracle, the same
quadratic B-spline density splat in numpy, central differences."""
import numpy as np


def bspline2_splat_np(pos, lo, hi, res):
    """numpy FP64 twin of fsi_torch.density_splat."""
    pos = np.asarray(pos, dtype=np.float64)
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    cell = (hi - lo) / res
    u = (pos - lo) / cell - 0.5
    i = np.floor(u + 0.5)
    f = u - i
    i = i.astype(np.int64)
    w = np.stack([0.5 * (0.5 - f) ** 2, 0.75 - f * f, 0.5 * (0.5 + f) ** 2], axis=-1)
    grid = np.zeros(res ** 3)
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                ix, iy, iz = i[:, 0] + dx - 1, i[:, 1] + dy - 1, i[:, 2] + dz - 1
                ww = w[:, 0, dx] * w[:, 1, dy] * w[:, 2, dz]
                ok = (ix >= 0) & (ix < res) & (iy >= 0) & (iy < res) & (iz >= 0) & (iz < res)
                np.add.at(grid, (iz[ok] * res + iy[ok]) * res + ix[ok], ww[ok])
    return grid.reshape(res, res, res)


def oracle_loss_f64(cfg, parts, v0, visc, eps, frames, target, lo, hi, res):
    """L = sum (splat(final positions) - target)^2 with the FP64 oracle."""
    import pyoracle

    sim = pyoracle.OracleSim(cfg, parts, precision="f64")
    sim.vel[:, 0:3] = np.asarray(v0, dtype=np.float64)
    sim.ubo.viscosity = float(visc)
    sim.ubo.eps = float(eps)
    sim.step(frames)
    d = bspline2_splat_np(sim.pos[:, 0:3], lo, hi, res)
    return float(((d - target) ** 2).sum())


def central_difference(cfg, parts, v0, visc, eps, frames, target, lo, hi, res, dv, dvisc, deps, step):
    args = (frames, target, lo, hi, res)
    lp = oracle_loss_f64(cfg, parts, v0 + step * dv, visc + step * dvisc, eps + step * deps, *args)
    lm = oracle_loss_f64(cfg, parts, v0 - step * dv, visc - step * dvisc, eps - step * deps, *args)
    return (lp - lm) / (2.0 * step)


def unit_direction(rng, n):
    """A seeded random direction in (v0 [n x 3], viscosity, eps), unit norm."""
    d = rng.normal(size=3 * n + 2)
    d /= np.linalg.norm(d)
    return d[:3 * n].reshape(n, 3), float(d[3 * n]), float(d[3 * n + 1])


def c4_scene(frames=20):
    """The differentiable-PBF workload: 65,536 particles (a 64x32x32 lattice
    block, spacing 0.05, seeded jitter) with initial velocities N(0, 0.1)
    (seed 1337), 4 iterations x 2 substeps per frame, dt = 1/120 s, viscosity
    0.01, eps 600; the loss grid is 32^3 over the 6.4^3 box.  The block falls
    freely in the middle of the box: it reaches no wall within 20 frames, so
    the loss is a smooth function of the inputs (a wall contact, with its
    velocity reflection, is a jump that central differences cannot follow)."""
    import fsi_pbf
    from conftest import dam_break_config

    box = (6.4, 6.4, 6.4)
    cfg = dam_break_config(n_iter=4, substeps=2, viscosity=0.01, box=box)
    parts = fsi_pbf.lattice(64, 32, 32, (1.6, 2.4, 2.4), 0.05, 0.001, 1337)
    v0 = np.random.default_rng(1337).normal(0.0, 0.1, size=(len(parts), 3)).astype(np.float32)
    parts[:, 8:11] = v0
    return cfg, parts, v0, frames, (0.0, 0.0, 0.0), box, 32
