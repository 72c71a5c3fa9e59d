"""Benchmark of the PBF substep on one MI355X.

Workload (the full-size plain configuration): 1,048,576 FP32 particles, a
128x64x128 lattice block (spacing 0.5h, h = 0.1, seeded jitter +-0.01h) in the
corner of a 12.8 x 6.4 x 12.8 box, with a static obstacle: an icosphere of
subdivision 5 (10,242 vertices, 20,480 triangles, radius 1.5, centred mid-box)
used as an SDF (cell 0.05, padding 0.025).  rho0 6378, eps 600, 6 solver
iterations, 2 substeps per frame, dt = 1/120 s, hash size 2^20.

A step is one frame (PBFluids::step: 2 substeps + vorticity).  The metric is
particle-iterations/s = N * substeps * iterations * frames / wall time, inputs
resident in HBM.  Per-pass kernel time comes from HIP events recorded on the
solver's stream around every launch during the timed region.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu-baseline]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "realtime-fsi-engine_amd")
sys.path.insert(0, PKG)

HBM_PEAK_GBS = 8000.0
N_SIDE = (128, 64, 128)
ITERS, SUBSTEPS, HASH = 6, 2, 1 << 20


def make_config(fsi):
    return fsi.FluidConfig(dt=1.0 / 120.0, gravity=(0.0, -9.81, 0.0), h=0.1, rho0=6378.0, solver_iterations=ITERS,
                           substep_iterations=SUBSTEPS, eps=600.0, bounds_min=(0.0, 0.0, 0.0),
                           bounds_max=(12.8, 6.4, 12.8), hash_size=HASH)


def make_scene(fsi):
    parts = fsi.lattice(*N_SIDE, (0.0, 0.0, 0.0), 0.05, 0.001, 1337)
    verts, faces = fsi.icosphere(5, (6.4, 3.2, 6.4), 1.5)
    sdf = fsi.sdf_build(verts, faces, 0.05)
    return parts, sdf


def algorithmic_bytes(n, sdf_bytes):
    """Compulsory HBM bytes per launch of each pass (FP32, each array touched
    once, the cell table of 8 bytes per bucket read once per gather pass)."""
    cells = 8 * HASH
    return {
        "predict": 52 * n,                     # R pos, vel; W vel, x*, key
        "sort": 20 * n,                        # counting-sort bound: R key, R+W (key, id)
        "tables": 8 * HASH + 4 * n + 8 * n,    # W table, R keys, W run ends
        "lambda": 20 * n + cells,              # R x; W lambda, rho
        "deltap": 28 * n + cells + sdf_bytes,  # R x, lambda; W x (apply fused)
        "integrate": 72 * n + cells + sdf_bytes,  # R pos, v*, x; W pos, vel, curl
        "vorticity": 48 * n + cells,           # R pos, vel, curl; W vel
    }


def cpu_baseline(fsi, cfg, parts, sdf, frames=2):
    """The CPU oracle (a C restatement of the reference shaders, OpenMP over
    particles) on the same workload, timed over `frames` frames."""
    threads = min(16, os.cpu_count() or 1)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    sim = pyoracle.OracleSim(cfg, parts, sdf=sdf, sdf_padding=0.025)
    t0 = time.perf_counter()
    sim.step(frames)
    wall = time.perf_counter() - t0
    n = len(parts)
    return {"value": n * SUBSTEPS * ITERS * frames / wall, "unit": "particle-iterations/s", "cores": threads,
            "kind": "port", "sample": f"{frames} frames of the same 1,048,576-particle workload ({wall:.1f} s)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the PBF substep of this build runs on a single MI355X "
              "(multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    import fsi_pbf as fsi

    cfg = make_config(fsi)
    parts, sdf = make_scene(fsi)
    n = len(parts)
    s = fsi.Solver(cfg)
    s.set_particles(parts)
    s.set_sdf(sdf, 0.025, False)
    s.step(args.warmup)
    s.sync()
    s.set_timing(True)
    s.timing_reset()
    t0 = time.perf_counter()
    s.step(args.steps)
    s.sync()
    wall = time.perf_counter() - t0
    tm = s.read_timing()
    s.set_timing(False)
    got = s.read_particles()
    assert np.all(np.isfinite(got[:, 0:3])), "non-finite positions"
    s.close()

    value = n * SUBSTEPS * ITERS * args.steps / wall
    sdf_bytes = sdf.grid.nbytes
    ab = algorithmic_bytes(n, sdf_bytes)
    passes = {}
    for name, ms in tm["ms"].items():
        cnt = tm["launches"][name]
        if cnt == 0:
            continue
        avg_ms = ms / cnt
        entry = {"launches": cnt, "avg_us": round(avg_ms * 1e3, 2), "share": round(ms / sum(tm["ms"].values()), 4)}
        if name in ab:
            entry["GBps"] = round(ab[name] / (avg_ms * 1e-3) / 1e9, 1)
        passes[name] = entry
    dom = max((p for p in passes if p in ("lambda", "deltap", "integrate", "vorticity")),
              key=lambda p: tm["ms"][p])
    achieved = passes[dom]["GBps"]
    roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "bytes_per_launch": ab[dom]}
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dom}_c2_1m.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            roofline["traffic"] = json.load(f).get("bytes_per_launch")

    out = {
        "metric": "PBF particle-iterations/s at 1M particles (plain/batched) + adjoint/forward ratio",
        "value": value, "unit": "particle-iterations/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded lattice + procedural icosphere SDF)",
        "config": {"workload": "plain PBF, 1,048,576 particles, icosphere SDF obstacle, I=6, S=2, H=2^20",
                   "particles": n, "iterations": ITERS, "substeps": SUBSTEPS, "hash_size": HASH},
        "roofline": roofline,
        "passes": passes,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(fsi, cfg, parts, sdf)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
