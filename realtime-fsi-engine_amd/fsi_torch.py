"""torch binding of the differentiable frames (fsi_adj_* in include/fsi_pbf.h).

``Frames`` owns one adjoint handle for n particles and a fixed number of
frames; ``simulate`` runs them as a ``torch.autograd.Function`` whose backward
is the library's hand-written adjoint (checkpoint per frame, replay with
neighbour lists, reverse passes), so a loss on the final state can be
differentiated with respect to the initial positions and velocities, the XSPH
coefficient and the relaxation epsilon.

torch must be imported before the library is loaded: both bring a HIP runtime
with the same soname, and only the first one loaded is used.
"""
from __future__ import annotations

import ctypes as C

import torch

import fsi_pbf


def _ptr(t):
    return C.c_void_p(t.data_ptr()) if t is not None else None


def _stream(device):
    # torch's default stream is the null stream (handle 0): passed as NULL
    return C.c_void_p(torch.cuda.current_stream(device).cuda_stream)


class Frames:
    """`frames` frames of PBFluids::step for n particles under cfg (viscosity
    and eps are inputs of each run).  precision 32 differentiates the solver's
    FP32 trajectory, 64 an FP64 run of the same passes."""

    def __init__(self, cfg: fsi_pbf.FluidConfig, n: int, frames: int, precision: int = 32):
        self.lib = fsi_pbf.load()
        self.cfg = cfg
        self.n = int(n)
        self.frames = int(frames)
        self._c = cfg.to_c()
        h = C.c_void_p()
        fsi_pbf._check(self.lib.fsi_adj_create(C.byref(self._c), self.n, self.frames, C.byref(h)))
        self.h = h
        self.generation = 0  # bumped by every forward: a backward needs its own forward's checkpoints
        self.precision = 32
        if precision != 32:
            self.set_precision(precision)

    def set_precision(self, bits: int):
        """Trajectory precision, 32 or 64; drops the last forward's checkpoints."""
        fsi_pbf._check(self.lib.fsi_adj_set_precision(self.h, int(bits)))
        self.precision = int(bits)
        self.generation += 1

    def stats(self) -> dict:
        """Neighbour statistics and the replay check of the last backward."""
        out = (C.c_uint32 * 4)()
        fsi_pbf._check(self.lib.fsi_adj_stats(self.h, out))
        return {"capacity_bits": out[0], "replay_mismatches": out[1], "max_neighbours": out[2],
                "max_in_degree": out[3]}

    def frame_norms(self):
        """(frames, 2): largest |adjoint| of positions / velocities at each frame start, last backward."""
        import numpy as np

        out = np.zeros((self.frames, 2), dtype=np.float32)
        fsi_pbf._check(self.lib.fsi_adj_frame_norms(self.h, out.ctypes.data_as(C.POINTER(C.c_float))))
        return out

    def close(self):
        if self.h:
            self.lib.fsi_adj_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _Run(torch.autograd.Function):
    @staticmethod
    def forward(ctx, frames: Frames, pos0, vel0, viscosity, eps):
        pos0 = pos0.detach().contiguous()
        vel0 = vel0.detach().contiguous()
        if pos0.shape != (frames.n, 4) or vel0.shape != (frames.n, 4):
            raise ValueError("pos0 / vel0 must have shape (n, 4)")
        if pos0.dtype not in (torch.float32, torch.float64) or vel0.dtype != pos0.dtype:
            raise ValueError("pos0 / vel0 must both be float32, or both float64")
        if not pos0.is_cuda or not vel0.is_cuda:
            raise ValueError("pos0 / vel0 must be on the GPU")
        f64 = pos0.dtype == torch.float64
        pos = torch.empty_like(pos0)
        vel = torch.empty_like(vel0)
        fwd = frames.lib.fsi_adj_forward64 if f64 else frames.lib.fsi_adj_forward
        fsi_pbf._check(fwd(frames.h, _ptr(pos0), _ptr(vel0), float(viscosity), float(eps), _ptr(pos), _ptr(vel),
                           _stream(pos0.device)))
        frames.generation += 1
        ctx.frames = frames
        ctx.generation = frames.generation
        ctx.param_like = (viscosity, eps)
        ctx.f64 = f64
        return pos, vel

    @staticmethod
    def backward(ctx, gpos, gvel):
        frames = ctx.frames
        if frames.generation != ctx.generation:
            raise RuntimeError("the Frames handle ran another forward since this graph was built")
        dt = torch.float64 if ctx.f64 else torch.float32
        gpos = gpos.to(dt).contiguous() if gpos is not None else None
        gvel = gvel.to(dt).contiguous() if gvel is not None else None
        like = gpos if gpos is not None else gvel
        gp0 = torch.empty((frames.n, 4), dtype=dt, device=like.device)
        gv0 = torch.empty_like(gp0)
        params = (C.c_double * 2)()
        bwd = frames.lib.fsi_adj_backward64 if ctx.f64 else frames.lib.fsi_adj_backward
        fsi_pbf._check(bwd(frames.h, _ptr(gpos), _ptr(gvel), _ptr(gp0), _ptr(gv0), params, _stream(like.device)))
        out = []
        for val, ref in zip(params, ctx.param_like):
            out.append(torch.as_tensor(val, dtype=ref.dtype, device=ref.device) if torch.is_tensor(ref) else None)
        return None, gp0, gv0, out[0], out[1]


def simulate(frames: Frames, pos0, vel0, viscosity, eps):
    """Final (pos, vel), each (n, 3), after the frames from (pos0, vel0) (n, 3).

    viscosity and eps are 0-d tensors (differentiable) or floats.  float64
    states (FP64 trajectory only) keep the whole chain in FP64."""
    pad = torch.zeros((pos0.shape[0], 1), dtype=pos0.dtype, device=pos0.device)
    p4 = torch.cat([pos0, pad], dim=1)
    v4 = torch.cat([vel0, pad], dim=1)
    if not torch.is_tensor(viscosity):
        viscosity = torch.tensor(float(viscosity), dtype=torch.float64)
    if not torch.is_tensor(eps):
        eps = torch.tensor(float(eps), dtype=torch.float64)
    pos, vel = _Run.apply(frames, p4, v4, viscosity, eps)
    return pos[:, 0:3], vel[:, 0:3]


def bspline2_weights(u):
    """Quadratic B-spline weights of the three nodes around u (node units)."""
    i = torch.floor(u + 0.5)
    f = u - i
    w = torch.stack([0.5 * (0.5 - f) ** 2, 0.75 - f * f, 0.5 * (0.5 + f) ** 2], dim=-1)
    return i.long(), w


def density_splat(pos, lo, hi, res: int):
    """Particle count density on a res^3 grid over [lo, hi] (nodes at cell
    centres), quadratic B-spline weights (C1 in the positions), nodes outside
    the grid dropped.  Returns (res, res, res), z slowest."""
    lo = torch.as_tensor(lo, dtype=pos.dtype, device=pos.device)
    hi = torch.as_tensor(hi, dtype=pos.dtype, device=pos.device)
    cell = (hi - lo) / res
    u = (pos - lo) / cell - 0.5
    i, w = bspline2_weights(u)
    grid = torch.zeros(res ** 3, dtype=pos.dtype, device=pos.device)
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                ix, iy, iz = i[:, 0] + dx - 1, i[:, 1] + dy - 1, i[:, 2] + dz - 1
                ww = w[:, 0, dx] * w[:, 1, dy] * w[:, 2, dz]
                ok = (ix >= 0) & (ix < res) & (iy >= 0) & (iy < res) & (iz >= 0) & (iz < res)
                flat = (iz * res + iy) * res + ix
                grid = grid.index_add(0, flat[ok], ww[ok])
    return grid.view(res, res, res)
