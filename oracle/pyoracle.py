"""Python access to the CPU oracle (TEST INFRASTRUCTURE ONLY).

Loads oracle/_build/liboracle.so (oracle/Makefile) and drives the plain-C
restatement of the reference's PBF substep.  Only tests/, the smoke check of
__graft_entry__ and the CPU-baseline leg of bench.py use this module; the
product never does.

``OracleSim`` mirrors a PBFluids instance: the bounds/damping rule of
PBFluids::setBounds (src/fluid/PBFluids.cpp:299-306), the packed constants of
setParams (:39-109) and the dispatch order of step (:323-517).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import sys
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# FSI_ORACLE_LIB names another build of these same sources: the mutation test
# (tests/test_oracle_mutants.py) runs the pins against deliberately broken
# builds to show that each pin fails on the mistake it guards against.
LIB_PATH = os.environ.get("FSI_ORACLE_LIB") or os.path.join(HERE, "_build", "liboracle.so")
sys.path.insert(0, os.path.join(os.path.dirname(HERE), "realtime-fsi-engine_amd"))
from fsi_pbf import CFluidConfig, FluidConfig, SdfGrid  # noqa: E402  (data formats only)


class OracleUbo(C.Structure):
    _fields_ = [
        ("bounds_min", C.c_double * 3), ("bounds_max", C.c_double * 3), ("gravity", C.c_double * 3),
        ("sub_dt", C.c_double), ("h", C.c_double), ("rho0", C.c_double), ("eps", C.c_double), ("wq", C.c_double),
        ("k_corr", C.c_double), ("n_corr", C.c_double), ("viscosity", C.c_double), ("bound_damping", C.c_double),
        ("hash_size", C.c_uint32), ("particle_count", C.c_uint32), ("enable_scorr", C.c_uint32),
        ("enable_viscosity", C.c_uint32), ("cohesion_strength", C.c_double), ("interaction_radius", C.c_double),
        ("interaction_strength", C.c_double), ("w0_self", C.c_double), ("hash_mask", C.c_uint32),
        ("min_lod", C.c_uint32), ("max_lod", C.c_uint32), ("enable_apbf", C.c_uint32),
        ("poly6_coeff", C.c_double), ("spiky_coeff", C.c_double), ("inv_rho0", C.c_double),
        ("lod_max_dist", C.c_double),
    ]


class OracleUniforms(C.Structure):
    _fields_ = [
        ("substeps", C.c_uint32), ("iterations", C.c_uint32), ("enable_vorticity", C.c_uint32),
        ("is_interacting", C.c_uint32), ("vorticity_epsilon", C.c_double), ("interaction_pos", C.c_double * 3),
        ("camera_pos", C.c_double * 3), ("enable_sdf", C.c_uint32), ("sdf_mode", C.c_uint32),
        ("sdf_dims", C.c_int32 * 3), ("sdf_origin", C.c_double * 3), ("sdf_cell", C.c_double),
        ("sdf_padding", C.c_double), ("sdf_grid", C.POINTER(C.c_float)),
    ]


class OracleState(C.Structure):
    _fields_ = [
        ("n", C.c_uint32), ("hash_size", C.c_uint32), ("pos", C.c_void_p), ("vel", C.c_void_p),
        ("solver", C.c_void_p), ("keys", C.POINTER(C.c_uint32)), ("ids", C.POINTER(C.c_uint32)),
        ("cells", C.POINTER(C.c_int32)), ("lod", C.POINTER(C.c_uint32)), ("vsnap", C.c_void_p),
        ("cand_count", C.POINTER(C.c_uint32)), ("nbr_count", C.POINTER(C.c_uint32)),
    ]


class CContact(C.Structure):
    _fields_ = [("pos", C.c_float * 4), ("vel", C.c_float * 4)]


_lib: Optional[C.CDLL] = None


def build():
    subprocess.run(["make", "-C", HERE, "-s"], check=True)


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    lib = C.CDLL(LIB_PATH)
    U, S, K = C.POINTER(OracleUbo), C.POINTER(OracleState), C.POINTER(OracleUniforms)
    f32p = C.POINTER(C.c_float)
    d3 = C.POINTER(C.c_double)
    lib.oracle_cell_hash.argtypes = [C.c_int32, C.c_int32, C.c_int32, C.c_uint32]
    lib.oracle_cell_hash.restype = C.c_uint32
    for p in ("f32", "f64"):
        getattr(lib, f"oracle_pack_ubo_{p}").argtypes = [C.POINTER(CFluidConfig), f32p, f32p, C.c_float, C.c_uint32, U]
        getattr(lib, f"oracle_predict_hash_{p}").argtypes = [U, K, S]
        getattr(lib, f"oracle_sort_{p}").argtypes = [S]
        getattr(lib, f"oracle_build_cells_{p}").argtypes = [S]
        getattr(lib, f"oracle_lod_{p}").argtypes = [U, K, S]
        getattr(lib, f"oracle_lambda_{p}").argtypes = [U, S, C.c_uint32]
        getattr(lib, f"oracle_deltap_{p}").argtypes = [U, S, C.c_uint32]
        getattr(lib, f"oracle_apply_{p}").argtypes = [U, K, S, C.c_uint32]
        getattr(lib, f"oracle_integrate_{p}").argtypes = [U, K, S]
        getattr(lib, f"oracle_vorticity_{p}").argtypes = [U, K, S]
        getattr(lib, f"oracle_substep_{p}").argtypes = [U, K, S]
        getattr(lib, f"oracle_step_{p}").argtypes = [U, K, S]
        fn = getattr(lib, f"oracle_contact_filter_{p}")
        fn.argtypes = [U, K, S, C.c_double, C.POINTER(CContact), C.c_uint32]
        fn.restype = C.c_uint32
        fn = getattr(lib, f"oracle_poly6_{p}")
        fn.argtypes = [U, d3]
        fn.restype = C.c_double
        getattr(lib, f"oracle_grad_spiky_{p}").argtypes = [U, d3, d3]
        fn = getattr(lib, f"oracle_sdf_query_{p}")
        fn.argtypes = [K, d3]
        fn.restype = C.c_double
        getattr(lib, f"oracle_sdf_gradient_{p}").argtypes = [K, d3, d3]
    lib.oracle_density_error_f64.argtypes = [f32p, C.c_uint32, U, d3, d3, d3]
    i32p, u32p, u8p = C.POINTER(C.c_int32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint8)
    lib.oracle_mesh_sdf.argtypes = [d3, d3, C.c_uint32, i32p, C.c_uint32, d3, C.c_double, i32p, f32p]
    lib.oracle_dirty_mask.argtypes = [d3, d3, C.c_uint32, C.c_double, C.c_double, d3, C.c_double, i32p, u8p, i32p]
    lib.oracle_dirty_mask.restype = C.c_uint32
    lib.oracle_contact_forces.argtypes = [d3, C.c_uint32, i32p, C.c_uint32, f32p, C.c_uint32, C.c_float, C.c_int,
                                          C.c_float, f32p, u32p]
    lib.oracle_closest_triangles.argtypes = [d3, i32p, C.c_uint32, f32p, C.c_uint32, i32p, f32p,
                                             C.POINTER(C.c_double)]
    _lib = lib
    return lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def mesh_sdf(verts, faces, origin, cell, dims, verts_build=None) -> np.ndarray:
    """Brute-force signed distance grid (dims[2], dims[1], dims[0]; x fastest
    in memory) of a mesh; verts_build: the vertices its edge / vertex
    pseudo-normals were built from (default: verts)."""
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1, 3)
    vb = v if verts_build is None else np.ascontiguousarray(verts_build, dtype=np.float64).reshape(-1, 3)
    f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
    o = np.ascontiguousarray(origin, dtype=np.float64)
    d = np.ascontiguousarray(dims, dtype=np.int32)
    grid = np.zeros((int(d[2]), int(d[1]), int(d[0])), dtype=np.float32)
    load().oracle_mesh_sdf(_p(v, C.c_double), _p(vb, C.c_double), len(v), _p(f, C.c_int32), len(f),
                           _p(o, C.c_double), float(cell), _p(d, C.c_int32), _p(grid, C.c_float))
    return grid


def dirty_mask(old_verts, new_verts, band, move_eps, origin, cell, dims):
    """(count, mask (z, y, x), box[6] = min xyz, max xyz) of a dirty refit."""
    ov = np.ascontiguousarray(old_verts, dtype=np.float64).reshape(-1, 3)
    nv = np.ascontiguousarray(new_verts, dtype=np.float64).reshape(-1, 3)
    o = np.ascontiguousarray(origin, dtype=np.float64)
    d = np.ascontiguousarray(dims, dtype=np.int32)
    mask = np.zeros((int(d[2]), int(d[1]), int(d[0])), dtype=np.uint8)
    box = np.zeros(6, dtype=np.int32)
    n = load().oracle_dirty_mask(_p(ov, C.c_double), _p(nv, C.c_double), len(ov), float(band), float(move_eps),
                                 _p(o, C.c_double), float(cell), _p(d, C.c_int32), _p(mask, C.c_uint8),
                                 _p(box, C.c_int32))
    return int(n), mask, box


def closest_triangles(verts, faces, points):
    """(face (n,) int32, closest point (n, 3) float32, squared distance (n,)
    float64) of every float point, by brute force over the triangles."""
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1, 3)
    f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
    p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
    face = np.zeros(len(p), dtype=np.int32)
    out = np.zeros((len(p), 3), dtype=np.float32)
    d2 = np.zeros(len(p), dtype=np.float64)
    load().oracle_closest_triangles(_p(v, C.c_double), _p(f, C.c_int32), len(f), _p(p, C.c_float), len(p),
                                    _p(face, C.c_int32), _p(out, C.c_float), _p(d2, C.c_double))
    return face, out, d2


def contact_forces(verts, faces, cand, band, use_vel=True, threshold=0.0):
    """Per-vertex forces (nv, 3) float32 and (contacts, control points) of the
    candidate records cand (ncand, 8) float32 = pos xyzw, vel xyzw."""
    v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1, 3)
    f = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
    c = np.ascontiguousarray(cand, dtype=np.float32).reshape(-1, 8)
    out = np.zeros((len(v), 3), dtype=np.float32)
    counts = np.zeros(2, dtype=np.uint32)
    load().oracle_contact_forces(_p(v, C.c_double), len(v), _p(f, C.c_int32), len(f), _p(c, C.c_float), len(c),
                                 float(band), 1 if use_vel else 0, float(threshold), _p(out, C.c_float),
                                 _p(counts, C.c_uint32))
    return out, (int(counts[0]), int(counts[1]))


def f32(x) -> float:
    return float(np.float32(x))


def padded_bounds(bmin, bmax, damping, padding=0.0):
    """PBFluids::setBounds (src/fluid/PBFluids.cpp:299-306) in FP32."""
    pad = np.float32(padding)
    mn = (np.asarray(bmin, dtype=np.float32) + pad).astype(np.float32)
    mx = (np.asarray(bmax, dtype=np.float32) - pad).astype(np.float32)
    d = np.float32(min(max(np.float32(damping), np.float32(0.0)), np.float32(1.0)))
    return mn, mx, d


def pack_ubo(cfg: FluidConfig, n: int, precision: str = "f32", bounds=None, damping=None,
             collision_padding: float = 0.0) -> OracleUbo:
    lib = load()
    bmin, bmax = (cfg.bounds_min, cfg.bounds_max) if bounds is None else bounds
    mn, mx, d = padded_bounds(bmin, bmax, cfg.bound_damping if damping is None else damping, collision_padding)
    u = OracleUbo()
    c = cfg.to_c()
    getattr(lib, f"oracle_pack_ubo_{precision}")(C.byref(c), mn.ctypes.data_as(C.POINTER(C.c_float)),
                                                   mx.ctypes.data_as(C.POINTER(C.c_float)), float(d), n, C.byref(u))
    return u


class OracleSim:
    """The reference's PBF substep on the CPU, state in particle-id order."""

    def __init__(self, cfg: FluidConfig, particles: np.ndarray, precision: str = "f32", bounds=None,
                 damping=None, collision_padding: float = 0.0, sdf: Optional[SdfGrid] = None,
                 sdf_padding: float = 0.0, containment: bool = False, interaction=None, camera=(0.0, 0.0, 0.0),
                 taps: bool = False):
        self.lib = load()
        self.cfg = cfg
        self.p = precision
        dt = np.float32 if precision == "f32" else np.float64
        parts = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1, 12)
        n = parts.shape[0]
        self.n = n
        self.H = cfg.hash_size
        self.initial = parts.copy()
        self.pos = np.ascontiguousarray(parts[:, 0:4].astype(dt))
        self.vel = np.ascontiguousarray(parts[:, 8:12].astype(dt))
        self.solver = np.zeros((n, 8), dtype=dt)
        self.keys = np.zeros(n, dtype=np.uint32)
        self.ids = np.zeros(n, dtype=np.uint32)
        self.cells = np.full((self.H, 2), -1, dtype=np.int32)
        self.lod = np.full(n, cfg.max_lod if cfg.enable_apbf else cfg.solver_iterations, dtype=np.uint32)
        self.vsnap = np.zeros((n, 4), dtype=dt)
        self.cand = np.zeros(n, dtype=np.uint32) if taps else None
        self.nbr = np.zeros(n, dtype=np.uint32) if taps else None
        self.ubo = pack_ubo(cfg, n, precision, bounds, damping, collision_padding)
        u = OracleUniforms()
        u.substeps = cfg.substep_iterations
        u.iterations = cfg.solver_iterations
        u.enable_vorticity = 1 if cfg.enable_vorticity else 0
        u.vorticity_epsilon = f32(cfg.vorticity_epsilon)
        if interaction is not None:
            u.is_interacting = 1
            for i in range(3):
                u.interaction_pos[i] = f32(interaction[i])
        for i in range(3):
            u.camera_pos[i] = f32(camera[i])
        self.uni = u
        self._grid = None
        if sdf is not None:
            self.set_sdf(sdf, sdf_padding, containment)
        st = OracleState()
        st.n = n
        st.hash_size = self.H
        st.pos = self.pos.ctypes.data
        st.vel = self.vel.ctypes.data
        st.solver = self.solver.ctypes.data
        st.keys = self.keys.ctypes.data_as(C.POINTER(C.c_uint32))
        st.ids = self.ids.ctypes.data_as(C.POINTER(C.c_uint32))
        st.cells = self.cells.ctypes.data_as(C.POINTER(C.c_int32))
        st.lod = self.lod.ctypes.data_as(C.POINTER(C.c_uint32))
        st.vsnap = self.vsnap.ctypes.data
        if taps:
            st.cand_count = self.cand.ctypes.data_as(C.POINTER(C.c_uint32))
            st.nbr_count = self.nbr.ctypes.data_as(C.POINTER(C.c_uint32))
        self.st = st

    def set_sdf(self, sdf: Optional[SdfGrid], padding: float = 0.0, containment: bool = False):
        """PBFluids::setSDFBoundary / setSDFPadding / setSDFContainment; None disables."""
        u = self.uni
        if sdf is None:
            self._grid = None
            u.enable_sdf = 0
            u.sdf_grid = None
            return
        self._grid = np.ascontiguousarray(sdf.grid, dtype=np.float32)
        u.enable_sdf = 1
        u.sdf_mode = 1 if containment else 0
        for i in range(3):
            u.sdf_dims[i] = sdf.dims[i]
            u.sdf_origin[i] = f32(sdf.origin[i])
        u.sdf_cell = f32(sdf.cell)
        u.sdf_padding = f32(padding)
        u.sdf_grid = self._grid.ctypes.data_as(C.POINTER(C.c_float))

    def _call(self, name, *args):
        return getattr(self.lib, f"oracle_{name}_{self.p}")(*args)

    def step(self, frames: int = 1):
        for _ in range(frames):
            self._call("step", C.byref(self.ubo), C.byref(self.uni), C.byref(self.st))

    def substep(self):
        self._call("substep", C.byref(self.ubo), C.byref(self.uni), C.byref(self.st))

    def predict(self):
        self._call("predict_hash", C.byref(self.ubo), C.byref(self.uni), C.byref(self.st))

    def sort(self):
        self._call("sort", C.byref(self.st))

    def build_cells(self):
        self._call("build_cells", C.byref(self.st))

    def lambda_pass(self, it: int = 1):
        self._call("lambda", C.byref(self.ubo), C.byref(self.st), it)

    def deltap_pass(self, it: int = 1):
        self._call("deltap", C.byref(self.ubo), C.byref(self.st), it)

    def apply_pass(self, it: int = 1):
        self._call("apply", C.byref(self.ubo), C.byref(self.uni), C.byref(self.st), it)

    def lod_pass(self):
        self._call("lod", C.byref(self.ubo), C.byref(self.uni), C.byref(self.st))

    def integrate(self):
        self._call("integrate", C.byref(self.ubo), C.byref(self.uni), C.byref(self.st))

    def vorticity(self):
        self._call("vorticity", C.byref(self.ubo), C.byref(self.uni), C.byref(self.st))

    def particles(self) -> np.ndarray:
        out = self.initial.copy()
        out[:, 0:4] = self.pos.astype(np.float32)
        out[:, 8:12] = self.vel.astype(np.float32)
        return out

    def contacts(self, band: float) -> np.ndarray:
        buf = (CContact * max(1, self.n))()
        cnt = self._call("contact_filter", C.byref(self.ubo), C.byref(self.uni), C.byref(self.st), f32(band),
                         buf, self.n)
        arr = np.frombuffer(buf, dtype=np.float32).reshape(-1, 8)
        return arr[:cnt].copy()


def density_error_f64(x: np.ndarray, cfg: FluidConfig) -> dict:
    """Density-constraint error of positions x (n,>=3) evaluated in FP64."""
    lib = load()
    x4 = np.zeros((len(x), 4), dtype=np.float32)
    x4[:, :3] = x[:, :3]
    u = pack_ubo(cfg, len(x), "f64")
    a, b, c = C.c_double(), C.c_double(), C.c_double()
    lib.oracle_density_error_f64(x4.ctypes.data_as(C.POINTER(C.c_float)), len(x), C.byref(u), C.byref(a),
                                 C.byref(b), C.byref(c))
    return {"mean_abs_C": a.value, "max_C": b.value, "mean_rho": c.value}
