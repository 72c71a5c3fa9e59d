"""Benchmark of the PBF substep on one MI355X.

Workload (the full-size plain configuration): 1,048,576 FP32 particles, a
128x64x128 lattice block (spacing 0.5h, h = 0.1, seeded jitter +-0.01h) in the
corner of a 12.8 x 6.4 x 12.8 box, with a static obstacle: an icosphere of
subdivision 5 (10,242 vertices, 20,480 triangles, radius 1.5, centred mid-box)
used as an SDF (cell 0.05, padding 0.025).  rho0 6378, eps 600, 6 solver
iterations, 2 substeps per frame, dt = 1/120 s, hash size 2^20.

A step is one frame (PBFluids::step: 2 substeps + vorticity); the default
times the workload's 200 frames (the dam break gets cheaper per frame as the
block spreads, so a short window would not be representative).  The metric is
particle-iterations/s = N * substeps * iterations * frames / wall time, inputs
resident in HBM.  The dominant kernel's launch time comes from HIP events
recorded on the solver's stream around its launches during the timed region;
the per-pass table comes from a separate short run with events around every
pass.

Second leg (reported under "batched"): 256 independent dam-break scenes with
particle counts log-uniform in [4096, 131072] (seeded), each a lattice block
in the corner of its own box next to a static sphere obstacle of radius
U[0.3, 1.0] given as an SDF grid; 4 iterations, 2 substeps.  The multi-scene
handle (one launch per pass for all scenes, frames replayed from a HIP graph,
knobs -- gather form, radix digit width, workgroup size, tile shape -- picked
by its autotuner) is timed against a loop of one single-scene handle per
scene under the same knobs, each replaying its frames from a HIP graph,
twice: each handle on its own stream (frames of different scenes overlap on
the four hardware queues), and every handle on one stream (frames in
submission order, as on the reference's one GL queue).  Each form runs the
median of five windows of --batch-steps frames.

Third leg (reported under "adjoint"): the differentiable workload, 65,536
particles (a 64x32x32 lattice block falling freely in the middle of a 6.4^3
box, initial velocities N(0, 0.1) seeded), 20 frames of 4 iterations x 2 substeps,
viscosity 0.01, eps 600.  The forward (frames, per-frame checkpoints and each
substep's sorted permutation) and the backward (replay of every frame with
neighbour lists + reverse passes) of the
hand-written adjoint are timed on the same inputs; the ratio is backward /
forward.  The backward's input is the gradient of a 32^3 density-matching loss.

Fourth leg (reported under "fsi"): the FSI coupling workload, 262,144
particles (a 64^3 lattice block) falling onto a cantilever plate of 128 x 32
quads (4.0 x 1.0, 8,192 triangles, 4,257 vertices) whose SDF (cell 0.05) lives
on the device; every frame: contact forces from the last frame's candidates,
the host Laplacian deformation of the plate under them (first column
anchored), the SDF refresh, the fluid step and the contact filter.  It runs
twice: with the SDF refreshed on every 3rd deformation, the reference's
cadence, and (under "fsi.refit_every_frame") on every deformation.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--batch-steps K2]
                       [--fsi-steps K3] [--no-batched] [--no-adjoint] [--no-fsi]
                       [--no-cpu-baseline] [--gather G] [--order O] [--block B]
                       [--tile T] [--no-autotune]
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
VALU_SIMDS, VALU_CLOCK_HZ = 1024, 2.4e9  # MI355X: 256 CUs x 4 SIMDs, peak engine clock
N_SIDE = (128, 64, 128)
ITERS, SUBSTEPS, HASH = 6, 2, 1 << 20


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


BATCH_SCENES, BATCH_ITERS, BATCH_SUBSTEPS, BATCH_SEED = 256, 4, 2, 2024


def sphere_sdf(fsi, centre, r, cell=0.05, margin=4):
    """Signed distance |x - c| - r sampled at voxel centres origin + i*cell,
    covering the sphere plus `margin` cells (x fastest)."""
    ext = r + margin * cell
    n = int(np.ceil(2.0 * ext / cell)) + 1
    origin = np.asarray(centre, dtype=np.float64) - ext
    ax = [origin[a] + cell * np.arange(n) for a in range(3)]
    z, y, x = np.meshgrid(ax[2], ax[1], ax[0], indexing="ij")
    d = np.sqrt((x - centre[0]) ** 2 + (y - centre[1]) ** 2 + (z - centre[2]) ** 2) - r
    return fsi.SdfGrid(grid=d.astype(np.float32), origin=origin, cell=cell, dims=(n, n, n))


def make_batch_scenes(fsi):
    """256 seeded scenes: n ~ logU[4096, 131072], the first n particles of a
    near-cubic lattice block (spacing 0.05, jitter 0.001) in the corner of the
    box, a sphere of radius U[0.3, 1.0] beside it, hash size next_pow2(n)."""
    rng = np.random.default_rng(BATCH_SEED)
    scenes = []
    for s in range(BATCH_SCENES):
        n = int(round(np.exp(rng.uniform(np.log(4096.0), np.log(131072.0)))))
        r = float(rng.uniform(0.3, 1.0))
        side = int(np.ceil(np.cbrt(n)))
        ny = int(np.ceil(n / (side * side)))
        parts = fsi.lattice(side, ny, side, (0.0, 0.0, 0.0), 0.05, 0.001, 1337 + s)[:n]
        bx, by, bz = side * 0.05, ny * 0.05, side * 0.05
        box = (bx + 2.0 * r + 0.4, max(1.5 * by, 2.0 * r + 0.4), max(bz, 2.0 * r + 0.4))
        centre = (bx + r + 0.2, r + 0.1, 0.5 * box[2])
        H = 1 << int(np.ceil(np.log2(n)))
        scenes.append(fsi.Scene(particles=parts, hash_size=H, bounds_min=(0.0, 0.0, 0.0), bounds_max=box,
                                sdf=sphere_sdf(fsi, centre, r), sdf_padding=0.025))
    return scenes


def batch_config(fsi, box=(1.0, 1.0, 1.0), hash_size=1 << 17):
    return fsi.FluidConfig(dt=1.0 / 120.0, gravity=(0.0, -9.81, 0.0), h=0.1, rho0=6378.0,
                           solver_iterations=BATCH_ITERS, substep_iterations=BATCH_SUBSTEPS, eps=600.0,
                           bounds_min=(0.0, 0.0, 0.0), bounds_max=box, hash_size=hash_size)


def trial_summary(trials):
    """Fastest autotune trial per gather form (ms of its timed frames), and how many settings were timed."""
    best = {}
    for t in trials:
        g = str(t["gather"])
        if g not in best or t["ms"] < best[g]["ms"]:
            best[g] = {k: (round(v, 3) if k == "ms" else v) for k, v in t.items()}
    return {"settings_timed": len(trials), "fastest_by_gather_form": best}


def timed_runs(step, sync, steps, runs):
    """Wall times of `runs` consecutive windows of `steps` frames each."""
    walls = []
    for _ in range(runs):
        sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        sync()
        walls.append(time.perf_counter() - t0)
    return walls


def batched_leg(fsi, steps, warmup, runs=5, autotune=True, graphs=True, loop=True):
    """Multi-scene handle vs a loop of single-scene handles on the same scenes,
    every handle replaying its frames from a HIP graph under the knobs the
    multi-scene autotuner picked.  Each form runs warmup frames, then `runs`
    windows of `steps` frames; the median window is reported, and the final
    states are compared bit for bit."""
    scenes = make_batch_scenes(fsi)
    total = sum(len(sc.particles) for sc in scenes)
    work = total * BATCH_SUBSTEPS * BATCH_ITERS * steps
    counts = [len(x.particles) for x in scenes]
    offsets = np.cumsum([0] + counts)

    b = fsi.Batch(batch_config(fsi))
    b.set_scenes(scenes)
    tuned = None
    if autotune:
        b.autotune(1)
        tuned = trial_summary(b.autotune_trials())
    knobs = b.knobs()
    b.set_graph(graphs)
    b.step(warmup)
    walls_b = timed_runs(lambda: b.step(1), b.sync, steps, runs)
    got = b.read_particles()
    b.close()
    wall_b = float(np.median(walls_b))
    if not loop:
        return {"scenes": BATCH_SCENES, "particles": total, "iterations": BATCH_ITERS, "substeps": BATCH_SUBSTEPS,
                "steps": steps, "warmup": warmup, "runs": runs, "value": work / wall_b,
                "ms_per_step": wall_b / steps * 1e3, "graph": graphs, "loop": "not run"}

    solvers = []
    for sc in scenes:
        sv = fsi.Solver(batch_config(fsi, sc.bounds_max, sc.hash_size))
        sv.set_particles(sc.particles)
        sv.set_sdf(sc.sdf, sc.sdf_padding, False)
        for name, k in fsi.KNOBS.items():
            sv.set_tuning(k, knobs[name])
        sv.set_graph(graphs)
        solvers.append(sv)

    def loop_step():
        for sv in solvers:
            sv.step()

    def loop_sync():
        for sv in solvers:
            sv.sync()

    def same_as_batch():
        return all(np.array_equal(sv.read_particles(), got[o:o + c]) for sv, o, c in zip(solvers, offsets, counts))

    for _ in range(warmup):
        loop_step()
    walls_l = timed_runs(loop_step, loop_sync, steps, runs)
    same = same_as_batch()

    # The same loop with every handle on one stream: the scenes' frames run in
    # submission order, as the reference's dispatches on its one GL queue.
    shared = solvers[0].get_stream()
    for sv, sc in zip(solvers, scenes):
        sv.set_particles(sc.particles)
        sv.set_stream(shared)
    for _ in range(warmup):
        loop_step()
    walls_q = timed_runs(loop_step, solvers[0].sync, steps, runs)
    same_q = same_as_batch()
    for sv in solvers[1:]:
        sv.set_stream(0)
    for sv in solvers:
        sv.close()
    wall_b, wall_l, wall_q = (float(np.median(w)) for w in (walls_b, walls_l, walls_q))
    return {
        "scenes": BATCH_SCENES, "particles": total, "iterations": BATCH_ITERS, "substeps": BATCH_SUBSTEPS,
        "steps": steps, "warmup": warmup, "runs": runs, "of": "median window of `steps` frames over `runs` windows",
        "value": work / wall_b, "ms_per_step": wall_b / steps * 1e3,
        "ms_per_step_runs": [round(w / steps * 1e3, 3) for w in walls_b],
        "loop_value": work / wall_l, "loop_ms_per_step": wall_l / steps * 1e3,
        "loop_ms_per_step_runs": [round(w / steps * 1e3, 3) for w in walls_l],
        "speedup": wall_l / wall_b,
        "loop_one_stream_value": work / wall_q, "loop_one_stream_ms_per_step": wall_q / steps * 1e3,
        "speedup_vs_one_stream_loop": wall_q / wall_b,
        "autotuned" if autotune else "knobs": {"gather": knobs["gather"], "sort_digit_bits": knobs["sort_digit_bits"],
                                               "block": knobs["block"], "tile": knobs["tile"], "order": knobs["order"]},
        "autotune_trials": tuned,
        "graph": {"batch": graphs, "loop": graphs, "loop_one_stream": graphs},
        "identical_to_loop": bool(same and same_q),
    }


def adjoint_leg(fsi, reps=3):
    """Forward / backward wall time of the adjoint handle at 65,536 particles."""
    import ctypes as C

    import torch
    import fsi_torch

    box = (6.4, 6.4, 6.4)
    cfg = fsi.FluidConfig(dt=1.0 / 120.0, gravity=(0.0, -9.81, 0.0), h=0.1, rho0=6378.0, solver_iterations=4,
                          substep_iterations=2, eps=600.0, bounds_min=(0.0, 0.0, 0.0), bounds_max=box,
                          hash_size=1 << 17, viscosity=0.01)
    frames = 20
    parts = fsi.lattice(64, 32, 32, (1.6, 2.4, 2.4), 0.05, 0.001, 1337)
    n = len(parts)
    v0 = np.random.default_rng(1337).normal(0.0, 0.1, size=(n, 3)).astype(np.float32)
    dev = torch.device("cuda:0")
    pos0 = torch.zeros((n, 4), device=dev)
    pos0[:, 0:3] = torch.tensor(parts[:, 0:3], device=dev)
    vel0 = torch.zeros((n, 4), device=dev)
    vel0[:, 0:3] = torch.tensor(v0, device=dev)
    fh = fsi_torch.Frames(cfg, n, frames)
    lib, h = fh.lib, fh.h
    pos = torch.empty_like(pos0)
    vel = torch.empty_like(vel0)
    stream = C.c_void_p(torch.cuda.current_stream(dev).cuda_stream)
    ptr = lambda t: C.c_void_p(t.data_ptr())  # noqa: E731

    def forward():
        fsi._check(lib.fsi_adj_forward(h, ptr(pos0), ptr(vel0), 0.01, 600.0, ptr(pos), ptr(vel), stream))

    # the backward's input: d/dx of a 32^3 density loss against a perturbed run
    vt = vel0.clone()
    vt[:, 0:3] += torch.tensor(np.random.default_rng(99).normal(0.0, 0.05, size=(n, 3)).astype(np.float32),
                               device=dev)
    fsi._check(lib.fsi_adj_forward(h, ptr(pos0), ptr(vt), 0.01, 600.0, ptr(pos), ptr(vel), stream))
    target = fsi_torch.density_splat(pos[:, 0:3].clone(), (0.0, 0.0, 0.0), box, 32)
    forward()
    x = pos[:, 0:3].clone().requires_grad_(True)
    loss = ((fsi_torch.density_splat(x, (0.0, 0.0, 0.0), box, 32) - target) ** 2).sum()
    (gx,) = torch.autograd.grad(loss, x)
    gpos = torch.zeros_like(pos0)
    gpos[:, 0:3] = gx
    gp0 = torch.empty_like(pos0)
    gv0 = torch.empty_like(pos0)
    params = (C.c_double * 2)()

    def backward():
        fsi._check(lib.fsi_adj_backward(h, ptr(gpos), None, ptr(gp0), ptr(gv0), params, stream))

    backward()  # warm
    tf, tb = [], []
    for _ in range(reps):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        forward()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        backward()
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        tf.append(t1 - t0)
        tb.append(t2 - t1)
    st = fh.stats()
    # the same frames as an FP64 trajectory (states and adjoints in FP64)
    fh.set_precision(64)
    p64, v64 = pos0.double(), vel0.double()
    po64, vo64 = torch.empty_like(p64), torch.empty_like(v64)
    g64, gp64, gv64 = gpos.double(), torch.empty_like(p64), torch.empty_like(p64)
    t64f, t64b = [], []
    for _ in range(reps + 1):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        fsi._check(lib.fsi_adj_forward64(h, ptr(p64), ptr(v64), 0.01, 600.0, ptr(po64), ptr(vo64), stream))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        fsi._check(lib.fsi_adj_backward64(h, ptr(g64), None, ptr(gp64), ptr(gv64), params, stream))
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        t64f.append(t1 - t0)
        t64b.append(t2 - t1)
    fh.close()
    f_ms = float(np.median(tf)) * 1e3
    b_ms = float(np.median(tb)) * 1e3
    f64_ms = float(np.median(t64f[1:])) * 1e3
    b64_ms = float(np.median(t64b[1:])) * 1e3
    return {"particles": n, "frames": frames, "iterations": 4, "substeps": 2, "forward_ms": round(f_ms, 3),
            "backward_ms": round(b_ms, 3), "adjoint_forward_ratio": round(b_ms / f_ms, 3), "reps": reps,
            "loss": float(loss.item()), "grad_viscosity": params[0], "grad_eps": params[1],
            "replay_mismatches": st["replay_mismatches"], "max_neighbours": st["max_neighbours"],
            "fp64_trajectory": {"forward_ms": round(f64_ms, 3), "backward_ms": round(b64_ms, 3),
                                "backward_over_fp32_forward": round(b64_ms / f_ms, 3),
                                "backward_over_fp64_forward": round(b64_ms / f64_ms, 3)}}


def plate_mesh(nx=128, nz=32, size=(4.0, 1.0), origin=(1.2, 1.0, 2.7)):
    """The cantilever plate of the FSI workload: an nx x nz quad grid in the
    plane y = origin[1], two triangles per quad (normals +y)."""
    xs = np.linspace(0.0, size[0], nx + 1)
    zs = np.linspace(0.0, size[1], nz + 1)
    X, Z = np.meshgrid(xs, zs, indexing="xy")
    verts = np.stack([X.ravel() + origin[0], np.full(X.size, float(origin[1])), Z.ravel() + origin[2]], axis=1)
    a = (np.arange(nz)[:, None] * (nx + 1) + np.arange(nx)[None, :]).ravel()
    b, c, d = a + 1, a + nx + 1, a + nx + 2
    faces = np.stack([np.stack([a, d, b], 1), np.stack([a, c, d], 1)], 1).reshape(-1, 3).astype(np.int32)
    return verts, faces


def fsi_leg(fsi, frames, every=3):
    """The two-way coupling loop of src/main.cpp:1674-1707: contact forces on
    the device from the last frame's candidates, the host Laplacian
    deformation of the plate (deformStep, src/main.cpp:862-944: the plate's
    first column anchored, a cantilever), the SDF refresh of syncDeformedMesh
    (src/main.cpp:803-860: every 3rd deformation, a dirty refit on the device;
    every 100th refresh a full one), the fluid step and the contact filter.
    On frames whose deformation cannot refresh the SDF, the fluid step is
    enqueued before the host solve and runs beside it (same results)."""
    verts, faces = plate_mesh()
    cell, band, thresh, padding = 0.05, 0.12, 0.5, 0.025
    params = fsi.deform_params()  # stiffness 1, 0.02 per frame, 0.4 total, restore 0.05, thresholds 0.5 / 0.3
    full_every = 100  # g_sdfFullRebuildEvery; `every` is g_sdfUpdateEvery (3 in the reference)
    dirty_band = max(padding, 2.0 * cell) + every * float(params.max_displacement) + cell
    cfg = fsi.FluidConfig(dt=1.0 / 120.0, gravity=(0.0, -9.81, 0.0), h=0.1, rho0=6378.0, solver_iterations=6,
                          substep_iterations=2, eps=600.0, bounds_min=(0.0, 0.0, 0.0), bounds_max=(6.4, 6.4, 6.4),
                          hash_size=1 << 19)
    parts = fsi.lattice(64, 64, 64, (1.6, 2.0, 1.6), 0.05, 0.001, 1337)
    n = len(parts)
    anchors = np.nonzero(verts[:, 0] == verts[:, 0].min())[0]

    mesh = fsi.Mesh(verts, faces, cell)
    deform = fsi.Deform(verts, faces)
    for v in anchors:
        deform.add_anchor(int(v))
    s = fsi.Solver(cfg)
    s.set_particles(parts)
    s.set_sdf_from_mesh(mesh, padding=padding)
    t_contact = t_deform = t_refit = t_step = 0.0
    contacts = dirty = handles = solves = 0
    syncs = sdf_syncs = 0
    s.filter_contacts(band)
    s.sync()
    t0 = time.perf_counter()
    overlapped = 0
    for f in range(frames):
        a = time.perf_counter()
        forces, (nc, _) = s.contact_forces(mesh, band, True, thresh)
        b = time.perf_counter()
        # when this frame's deformation cannot refresh the SDF (syncDeformedMesh
        # refreshes on every `every`-th solve), the fluid step does not read
        # it: the step is enqueued first and the host solve runs beside it
        early = syncs % every != 0
        if early:
            s.step()
            s.filter_contacts(band)
            overlapped += 1
        nh, solved = deform.step(forces, cfg.dt, params)
        c = time.perf_counter()
        if solved:
            solves += 1
            if syncs % every == 0:
                cur = deform.vertices()
                if sdf_syncs % full_every == 0:
                    mesh.refit(cur)
                else:
                    nd, _, _ = mesh.refit_dirty(cur, dirty_band, 1e-4)
                    dirty += nd
                sdf_syncs += 1
                s.set_sdf_from_mesh(mesh, padding=padding)
            syncs += 1
        d = time.perf_counter()
        if not early:
            s.step()
            s.filter_contacts(band)
        s.sync()
        e = time.perf_counter()
        t_contact += b - a
        t_deform += c - b
        t_refit += d - c
        t_step += e - d
        contacts += nc
        handles += nh
    wall = time.perf_counter() - t0
    final = deform.vertices()
    info = deform.info()
    # a full refit (SDFBoundary::refitFromMesh) on the device vs the host build
    reps = 5
    a = time.perf_counter()
    for _ in range(reps):
        mesh.refit(final)
    full_ms = (time.perf_counter() - a) / reps * 1e3
    a = time.perf_counter()
    fsi.sdf_build(final, faces, cell)
    host_ms = (time.perf_counter() - a) * 1e3
    got = s.read_particles()
    s.close()
    deform.close()
    mesh.close()
    return {"particles": n, "frames": frames, "iterations": 6, "substeps": 2, "sdf_refresh_every": every,
            "plate": {"vertices": len(verts), "triangles": len(faces), "anchors": int(len(anchors)), "sdf_cell": cell,
                      "sdf_dims": list(mesh.dims)},
            "ms_per_frame": round(wall / frames * 1e3, 3),
            "value": n * 12 * frames / wall, "unit": "particle-iterations/s",
            "contact_ms_per_frame": round(t_contact / frames * 1e3, 3),
            "deform_ms_per_frame": round(t_deform / frames * 1e3, 3),
            "sdf_refit_ms_per_frame": round(t_refit / frames * 1e3, 3),
            "step_wait_ms_per_frame": round(t_step / frames * 1e3, 3),
            "mean_contacts": round(contacts / frames, 1), "mean_handles": round(handles / frames, 1),
            "deform_solves": solves, "factorizations": info["factorizations"], "sdf_refreshes": sdf_syncs,
            "frames_with_step_beside_deform": overlapped,
            "mean_dirty_voxels_per_refresh": round(dirty / max(1, sdf_syncs - 1), 1),
            "max_displacement_from_rest": round(float(np.abs(final - verts).max()), 4),
            "full_refit_ms": round(full_ms, 3), "host_sdf_build_ms": round(host_ms, 3),
            "host_sdf_build_threads": min(16, os.cpu_count() or 1),
            "finite": bool(np.all(np.isfinite(got[:, 0:3])))}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(fsi, cfg, parts, sdf, warmup=2, runs=5):
    """The CPU oracle (a C restatement of the reference shaders, OpenMP over
    particles; its -O3 x86-64-v3 timing build, same FP policy and bits) on the
    same workload: warm-up frames, then the median of one-frame runs, on every
    thread of the process's host share (OMP_NUM_THREADS where the host sets
    it, else the CPUs this process may run on)."""
    affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    threads = int(os.environ.get("OMP_NUM_THREADS") or affinity)
    os.environ["OMP_NUM_THREADS"] = str(threads)
    os.environ["FSI_ORACLE_LIB"] = os.path.join(ROOT, "oracle", "_build", "liboracle_fast.so")
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    sim = pyoracle.OracleSim(cfg, parts, sdf=sdf, sdf_padding=0.025)
    sim.step(warmup)
    times = []
    for _ in range(runs):
        t0 = time.perf_counter()
        sim.step(1)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    n = len(parts)
    return {"value": n * SUBSTEPS * ITERS / med, "unit": "particle-iterations/s", "cores": threads,
            "kind": "port",
            "sample": f"median of {runs} one-frame runs (frames {warmup + 1}-{warmup + runs}, after {warmup} "
                      f"warm-up frames) of the same 1,048,576-particle workload: {med:.2f} s per frame "
                      f"(min {min(times):.2f}, max {max(times):.2f})",
            "build": "gcc -O3 -march=x86-64-v3 -fopenmp -ffp-contract=off (oracle/_build/liboracle_fast.so)",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(), "process_cpus": affinity,
            "omp_threads": threads}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--batch-steps", type=int, default=8)
    ap.add_argument("--block", type=int, default=None, help="workgroup size of the gathers (128 or 256)")
    ap.add_argument("--tile", type=int, default=None, help="tile shape of the tiled gathers (0, 1, 2)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default knobs (plain and batched legs) instead of autotuning them")
    ap.add_argument("--no-batched", action="store_true")
    ap.add_argument("--no-batch-loop", action="store_true",
                    help="time only the multi-scene handle in the batched leg (no per-scene loops)")
    ap.add_argument("--no-graphs", action="store_true",
                    help="enqueue the batched leg's frames directly instead of replaying HIP graphs")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--fsi-steps", type=int, default=300)
    ap.add_argument("--no-fsi", action="store_true")
    ap.add_argument("--gather", type=int, default=None,
                    help="pin the gather form of the 1M solver (0, 1, 2; default: the library's, 1)")
    ap.add_argument("--order", type=int, default=None,
                    help="pin the processing order of the 1M solver (0 sorted slots, 1 Morton; default 1)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain device copy of --calibrate-mib MiB (3 launches), the reference point "
                         "that scales the FETCH_SIZE / WRITE_SIZE counters to bytes")
    ap.add_argument("--calibrate-mib", type=int, default=256)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the PBF substep of this build runs on a single MI355X "
              "(multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)

    if args.calibrate_copy:
        import torch

        n = args.calibrate_mib << 18  # floats
        a = torch.full((n,), 1.5, device="cuda:0", dtype=torch.float32)
        b = torch.empty_like(a)
        for _ in range(3):
            torch.mul(a, 1.0, out=b)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "copy", "bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 3}))
        return

    if not args.no_adjoint:
        import torch  # noqa: F401  (before the library: one HIP runtime in the process)
    import fsi_pbf as fsi

    cfg = make_config(fsi)
    parts, sdf = make_scene(fsi)
    n = len(parts)
    s = fsi.Solver(cfg)
    s.set_particles(parts)
    s.set_sdf(sdf, 0.025, False)
    fixed = {"gather": args.gather, "order": args.order, "block": args.block, "tile": args.tile}
    tuned = None
    if all(v is None for v in fixed.values()) and not args.no_autotune:
        s.autotune(2)  # times every knob setting from this state, which it restores
        tuned = trial_summary(s.autotune_trials())
    for name, v in fixed.items():
        if v is not None:
            s.set_tuning(fsi.KNOBS[name], v)
    knobs = s.knobs()
    knobs["autotuned"] = all(v is None for v in fixed.values()) and not args.no_autotune
    if tuned:
        knobs["autotune_trials"] = tuned
    s.step(args.warmup)
    s.sync()
    # timed region: events bracket the launches of the dominant pass (lambda)
    # only, so the wall clock stays that of an untimed run
    s.set_timing_passes(["lambda"])
    s.timing_reset()
    t0 = time.perf_counter()
    s.step(args.steps)
    s.sync()
    wall = time.perf_counter() - t0
    tm_dom = s.read_timing()
    # per-pass breakdown: a separate short run with events around every pass
    s.set_timing(True)
    s.timing_reset()
    s.step(min(args.steps, 5))
    s.sync()
    tm = s.read_timing()
    s.set_timing(False)
    got = s.read_particles()
    assert np.all(np.isfinite(got[:, 0:3])), "non-finite positions"
    # density-constraint error of the last lambda pass (CS_ComputeLambdas.glsl:
    # C = max(rho / rho0 - 1, -cohesion)), from the solver's own rho
    _, drho = s.read_solver()
    c_err = np.maximum(drho[:, 3].astype(np.float64) / float(cfg.rho0) - 1.0, -float(cfg.cohesion_strength))
    density_error = {"mean_abs_C": float(np.abs(c_err).mean()), "max_C": float(c_err.max()),
                     "of": f"the last lambda pass of frame {args.warmup + args.steps + min(args.steps, 5)}"}
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
    dom = "lambda"
    dom_avg_ms = tm_dom["ms"][dom] / max(1, tm_dom["launches"][dom])
    achieved = round(ab[dom] / (dom_avg_ms * 1e-3) / 1e9, 1)
    roofline = {"bound": "hbm", "kernel": "lambda pass (list builder + lambda walk, one launch pair per iteration)",
                "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                "bytes_per_launch": ab[dom], "avg_launch_us": round(dom_avg_ms * 1e3, 2),
                "launches": tm_dom["launches"][dom]}
    # HBM bytes per launch of the dominant pass, measured separately: TCC EA
    # read and write requests by size per dispatch, from rocprofv3 --pmc runs
    # of this workload (frames 1-3) with the default knobs (profiles/)
    tfile = os.path.join(ROOT, "profiles", "traffic_pbf_c2_1m.json")
    if os.path.exists(tfile):
        with open(tfile) as f:
            tp = json.load(f)["passes"].get(dom)
        if tp:
            roofline["traffic"] = round(tp["bytes_per_launch"])
            roofline["traffic_of"] = ("median bytes per dispatch of " + " + ".join(tp["kernels"]) +
                                      ", rocprofv3 --pmc TCC_EA0_RDREQ_{32B,64B,128B}_sum and TCC_EA0_WRREQ(_64B)_sum "
                                      "over frames 1-3 of this workload (profiles/traffic_pbf_c2_1m.md), "
                                      "not over the timed window")

    # issue rate of the dominant pass, which is VALU-bound: its VALU
    # instructions per launch (rocprofv3 SQ_INSTS_VALU of this workload,
    # profiles/counters_pbf_c2_1m.json) over the live launch time, at one wave
    # instruction per 4 cycles on each of the 1024 SIMDs at 2.4 GHz
    cfile = os.path.join(ROOT, "profiles", "counters_pbf_c2_1m.json")
    if os.path.exists(cfile):
        with open(cfile) as f:
            cp = json.load(f)["passes"].get(dom)
        if cp:
            issue_s = cp["valu_inst_per_launch"] * 4.0 / (VALU_SIMDS * VALU_CLOCK_HZ)
            roofline["valu_issue"] = {
                "frac": round(issue_s / (dom_avg_ms * 1e-3), 4),
                "valu_inst_per_launch": round(cp["valu_inst_per_launch"]),
                "simds": VALU_SIMDS, "clock_ghz": VALU_CLOCK_HZ / 1e9,
                "of": "VALU instructions per launch of " + " + ".join(cp["kernels"]) + " from rocprofv3 --pmc "
                      "SQ_INSTS_VALU over frames 1-3 of this workload (profiles/counters_pbf_c2_1m.md), "
                      "x 4 cycles / (1024 SIMDs x 2.4 GHz), over the live average launch time"}

    out = {
        "metric": "PBF particle-iterations/s at 1M particles (plain/batched) + adjoint/forward ratio",
        "value": value, "unit": "particle-iterations/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic (seeded lattice + procedural icosphere SDF)",
        "config": {"workload": "plain PBF, 1,048,576 particles, icosphere SDF obstacle, I=6, S=2, H=2^20",
                   "particles": n, "iterations": ITERS, "substeps": SUBSTEPS, "hash_size": HASH,
                   "knobs": knobs},
        "roofline": roofline,
        "density_error": density_error,
        "passes": passes,
        "passes_note": "separate run of min(steps, 5) frames with events around every pass",
    }
    progress(f"plain leg: {out['ms_per_step']:.3f} ms/frame")
    if not args.no_batched:
        out["batched"] = batched_leg(fsi, args.batch_steps, 1, autotune=not args.no_autotune,
                                     graphs=not args.no_graphs, loop=not args.no_batch_loop)
        progress(f"batched leg: {out['batched']['ms_per_step']:.3f} ms/frame, "
                 f"speed-up {out['batched'].get('speedup', float('nan')):.2f}")
    if not args.no_adjoint:
        out["adjoint"] = adjoint_leg(fsi)
        progress(f"adjoint leg: ratio {out['adjoint']['adjoint_forward_ratio']:.2f}")
    if not args.no_fsi:
        out["fsi"] = fsi_leg(fsi, args.fsi_steps, every=3)
        progress(f"fsi leg: {out['fsi']['ms_per_frame']:.3f} ms/frame")
        out["fsi"]["refit_every_frame"] = fsi_leg(fsi, args.fsi_steps, every=1)
        progress(f"fsi leg, refit every frame: {out['fsi']['refit_every_frame']['ms_per_frame']:.3f} ms/frame")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(fsi, cfg, parts, sdf)
        progress("cpu baseline done")
    print(json.dumps(out))


if __name__ == "__main__":
    main()
