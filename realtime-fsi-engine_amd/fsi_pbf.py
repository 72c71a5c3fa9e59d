"""ctypes binding of the gfx950 PBF substep library (include/fsi_pbf.h).

The binding mirrors the reference's ``PBFluids`` solver interface
(include/fluid/PBFluids.h:88-268): ``Solver`` has ``set_params``,
``set_particles``, ``set_bounds``, ``set_sdf``, ``step``, ``read_particles``,
``filter_contacts`` / ``read_contacts`` and the inspection taps.  Particle
arrays are ``(n, 12)`` float32 numpy arrays in the reference's 48-byte record
layout (include/fluid/Particle.h:8-12): pos.xyz, pad0, predPos.xyz, pad1,
vel.xyz, pad2.

The library must be built (``make -C realtime-fsi-engine_amd``); loading fails
loudly otherwise.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field, fields
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libfsi_pbf.so")

FSI_OK = 0
TUNE_GATHER, TUNE_SORT_DIGIT_BITS, TUNE_ORDER, TUNE_BLOCK, TUNE_TILE = 0, 1, 2, 3, 4
KNOBS = {"gather": TUNE_GATHER, "sort_digit_bits": TUNE_SORT_DIGIT_BITS, "order": TUNE_ORDER,
         "block": TUNE_BLOCK, "tile": TUNE_TILE}


class CTuneTrial(C.Structure):
    _fields_ = [("gather", C.c_int32), ("sort_digit_bits", C.c_int32), ("order", C.c_int32), ("block", C.c_int32),
                ("tile", C.c_int32), ("ms", C.c_float)]


def _read_trials(fn, h):
    n = C.c_uint32(0)
    _check(fn(h, None, 0, C.byref(n)))
    arr = (CTuneTrial * max(1, n.value))()
    _check(fn(h, C.cast(arr, C.c_void_p), n.value, C.byref(n)))
    return [{f: getattr(arr[i], f) for f, _ in CTuneTrial._fields_} for i in range(n.value)]
PASS_NAMES = ["predict", "sort", "tables", "lambda", "deltap", "integrate", "vorticity", "lod"]


class FsiError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"fsi error {code}: {msg}")
        self.code = code


class CFluidConfig(C.Structure):
    _fields_ = [
        ("dt", C.c_float), ("gravity", C.c_float * 3), ("h", C.c_float), ("rho0", C.c_float),
        ("solver_iterations", C.c_uint32), ("substep_iterations", C.c_uint32), ("eps", C.c_float),
        ("particle_count", C.c_uint32), ("spawn_min", C.c_float * 3), ("spawn_max", C.c_float * 3),
        ("spawn_random", C.c_uint32), ("spacing", C.c_float), ("initial_velocity", C.c_float * 3),
        ("bounds_min", C.c_float * 3), ("bounds_max", C.c_float * 3), ("bound_damping", C.c_float),
        ("hash_size", C.c_uint32), ("enable_scorr", C.c_uint32), ("k_corr", C.c_float), ("n_corr", C.c_float),
        ("delta_q", C.c_float), ("cohesion_strength", C.c_float), ("enable_viscosity", C.c_uint32),
        ("viscosity", C.c_float), ("interaction_radius", C.c_float), ("interaction_strength", C.c_float),
        ("enable_vorticity", C.c_uint32), ("vorticity_epsilon", C.c_float), ("enable_apbf", C.c_uint32),
        ("min_lod", C.c_uint32), ("max_lod", C.c_uint32), ("lod_max_dist", C.c_float),
    ]


class CConstants(C.Structure):
    _fields_ = [
        ("bounds_min", C.c_float * 4), ("bounds_max", C.c_float * 4), ("gravity_dt", C.c_float * 4),
        ("h", C.c_float), ("rho0", C.c_float), ("eps", C.c_float), ("wq", C.c_float),
        ("k_corr", C.c_float), ("n_corr", C.c_float), ("viscosity", C.c_float), ("bound_damping", C.c_float),
        ("hash_size", C.c_uint32), ("particle_count", C.c_uint32), ("enable_scorr", C.c_uint32),
        ("enable_viscosity", C.c_uint32), ("cohesion_strength", C.c_float), ("interaction_radius", C.c_float),
        ("interaction_strength", C.c_float), ("w0_self", C.c_float), ("hash_mask", C.c_uint32),
        ("poly6_coeff", C.c_float), ("spiky_coeff", C.c_float), ("inv_rho0", C.c_float),
        ("min_lod", C.c_uint32), ("max_lod", C.c_uint32), ("lod_max_dist", C.c_float),
        ("enable_apbf", C.c_uint32),
    ]


class CSceneDesc(C.Structure):
    _fields_ = [
        ("particle_offset", C.c_uint32), ("particle_count", C.c_uint32), ("hash_size", C.c_uint32),
        ("bounds_min", C.c_float * 3), ("bounds_max", C.c_float * 3), ("sdf_grid", C.POINTER(C.c_float)),
        ("sdf_dims", C.c_int32 * 3), ("sdf_origin", C.c_float * 3), ("sdf_cell", C.c_float),
        ("sdf_padding", C.c_float), ("sdf_containment", C.c_uint32),
    ]


class CDeformParams(C.Structure):
    _fields_ = [("stiffness", C.c_float), ("max_displacement", C.c_float), ("max_total_disp", C.c_float),
                ("restore_enabled", C.c_uint32), ("restore_strength", C.c_float), ("control_threshold", C.c_float),
                ("control_threshold_low", C.c_float)]


class CTiming(C.Structure):
    _fields_ = [("ms", C.c_double * 8), ("launches", C.c_uint64 * 8), ("substeps", C.c_uint64),
                ("frames", C.c_uint64)]


@dataclass
class FluidConfig:
    """The reference's FluidConfig (include/core/config.h:9-63), same order and defaults."""
    dt: float = 1.0 / 60.0
    gravity: tuple = (0.0, -3.81, 0.0)
    h: float = 0.15
    rho0: float = 1200.0
    solver_iterations: int = 3
    substep_iterations: int = 2
    eps: float = 0.0001
    particle_count: int = 1 << 12
    spawn_min: tuple = (2.0, 0.2, 2.0)
    spawn_max: tuple = (5.0, 5.0, 5.0)
    spawn_random: bool = False
    spacing: float = 0.1
    initial_velocity: tuple = (0.0, 0.0, 0.0)
    bounds_min: tuple = (0.0, 0.0, 0.0)
    bounds_max: tuple = (6.0, 6.0, 6.0)
    bound_damping: float = 0.5
    hash_size: int = 1 << 17
    enable_scorr: bool = True
    k_corr: float = 0.001
    n_corr: float = 4.0
    delta_q: float = 0.3
    cohesion_strength: float = 0.001
    enable_viscosity: bool = True
    viscosity: float = 0.1
    interaction_radius: float = 1.5
    interaction_strength: float = 25.0
    enable_vorticity: bool = True
    vorticity_epsilon: float = 0.05
    enable_apbf: bool = False
    min_lod: int = 2
    max_lod: int = 7
    lod_max_dist: float = 15.0

    def to_c(self) -> CFluidConfig:
        c = CFluidConfig()
        for f in fields(self):
            v = getattr(self, f.name)
            if isinstance(v, (tuple, list)):
                arr = getattr(c, f.name)
                for i, x in enumerate(v):
                    arr[i] = float(np.float32(x))
            elif isinstance(v, bool):
                setattr(c, f.name, 1 if v else 0)
            else:
                setattr(c, f.name, v)
        return c

    def f32(self, name: str) -> float:
        """The FP32 value the library sees for a float field."""
        return float(np.float32(getattr(self, name)))


_lib: Optional[C.CDLL] = None


def load() -> C.CDLL:
    """Load lib/libfsi_pbf.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(
            f"{LIB_PATH} is missing: build it with `make -C {HERE}` (or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    P = C.c_void_p
    u32p = C.POINTER(C.c_uint32)
    i32p = C.POINTER(C.c_int32)
    f32p = C.POINTER(C.c_float)
    f64p = C.POINTER(C.c_double)
    sig = {
        "fsi_last_error": ([], C.c_char_p),
        "fsi_abi_version": ([], C.c_int),
        "fsi_fluid_config_default": ([C.POINTER(CFluidConfig)], C.c_int),
        "fsi_device_count": ([C.POINTER(C.c_int)], C.c_int),
        "fsi_pbf_create": ([C.POINTER(CFluidConfig), C.POINTER(P)], C.c_int),
        "fsi_pbf_destroy": ([P], C.c_int),
        "fsi_pbf_set_params": ([P, C.POINTER(CFluidConfig)], C.c_int),
        "fsi_pbf_get_params": ([P, C.POINTER(CFluidConfig)], C.c_int),
        "fsi_pbf_get_constants": ([P, C.POINTER(CConstants)], C.c_int),
        "fsi_pbf_set_particles": ([P, P, C.c_uint32], C.c_int),
        "fsi_pbf_particle_count": ([P, u32p], C.c_int),
        "fsi_pbf_read_particles": ([P, P, C.c_uint32, u32p], C.c_int),
        "fsi_pbf_write_render_buffer": ([P, P, P], C.c_int),
        "fsi_spawn_particles": ([C.POINTER(CFluidConfig), P, C.c_uint32, u32p], C.c_int),
        "fsi_load_off": ([C.c_char_p, f64p, C.c_uint32, i32p, C.c_uint32, u32p, u32p], C.c_int),
        "fsi_mesh_fit_to_domain": ([f64p, C.c_uint32, f32p, C.c_double], C.c_int),
        "fsi_pbf_set_collision_padding": ([P, C.c_float], C.c_int),
        "fsi_pbf_set_bounds": ([P, f32p, f32p, C.c_float], C.c_int),
        "fsi_pbf_set_sdf": ([P, f32p, i32p, f32p, C.c_float], C.c_int),
        "fsi_pbf_set_sdf_padding": ([P, C.c_float], C.c_int),
        "fsi_pbf_set_sdf_containment": ([P, C.c_int], C.c_int),
        "fsi_pbf_set_interaction": ([P, C.c_int, f32p], C.c_int),
        "fsi_pbf_set_camera": ([P, f32p], C.c_int),
        "fsi_pbf_step": ([P], C.c_int),
        "fsi_pbf_sync": ([P], C.c_int),
        "fsi_pbf_filter_contacts": ([P, C.c_float], C.c_int),
        "fsi_pbf_read_contacts": ([P, P, C.c_uint32, u32p], C.c_int),
        "fsi_pbf_read_grid": ([P, u32p, u32p, i32p], C.c_int),
        "fsi_pbf_read_solver": ([P, f32p, f32p], C.c_int),
        "fsi_pbf_set_taps": ([P, C.c_int], C.c_int),
        "fsi_pbf_read_counts": ([P, u32p, u32p], C.c_int),
        "fsi_pbf_set_timing": ([P, C.c_int], C.c_int),
        "fsi_pbf_set_timing_passes": ([P, C.c_uint32], C.c_int),
        "fsi_batch_set_timing_passes": ([P, C.c_uint32], C.c_int),
        "fsi_pbf_timing_reset": ([P], C.c_int),
        "fsi_pbf_read_timing": ([P, C.POINTER(CTiming)], C.c_int),
        "fsi_pbf_get_stream": ([P, C.POINTER(P)], C.c_int),
        "fsi_pbf_set_stream": ([P, P], C.c_int),
        "fsi_pbf_set_tuning": ([P, C.c_int, C.c_int], C.c_int),
        "fsi_pbf_get_tuning": ([P, C.c_int, C.POINTER(C.c_int)], C.c_int),
        "fsi_pbf_autotune": ([P, C.c_uint32, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
        "fsi_pbf_read_autotune": ([P, P, C.c_uint32, u32p], C.c_int),
        "fsi_pbf_set_graph": ([P, C.c_int], C.c_int),
        "fsi_batch_create": ([C.POINTER(CFluidConfig), C.POINTER(P)], C.c_int),
        "fsi_batch_destroy": ([P], C.c_int),
        "fsi_batch_set_scenes": ([P, P, C.c_uint32, C.POINTER(CSceneDesc), C.c_uint32], C.c_int),
        "fsi_batch_step": ([P], C.c_int),
        "fsi_batch_sync": ([P], C.c_int),
        "fsi_batch_read_particles": ([P, P, C.c_uint32, u32p], C.c_int),
        "fsi_batch_read_grid": ([P, C.c_uint32, u32p, u32p, i32p], C.c_int),
        "fsi_batch_set_graph": ([P, C.c_int], C.c_int),
        "fsi_batch_set_tuning": ([P, C.c_int, C.c_int], C.c_int),
        "fsi_batch_get_tuning": ([P, C.c_int, C.POINTER(C.c_int)], C.c_int),
        "fsi_batch_autotune": ([P, C.c_uint32, C.POINTER(C.c_int), C.POINTER(C.c_int)], C.c_int),
        "fsi_batch_read_autotune": ([P, P, C.c_uint32, u32p], C.c_int),
        "fsi_batch_set_timing": ([P, C.c_int], C.c_int),
        "fsi_batch_timing_reset": ([P], C.c_int),
        "fsi_batch_read_timing": ([P, C.POINTER(CTiming)], C.c_int),
        "fsi_adj_create": ([C.POINTER(CFluidConfig), C.c_uint32, C.c_uint32, C.POINTER(P)], C.c_int),
        "fsi_adj_destroy": ([P], C.c_int),
        "fsi_adj_set_precision": ([P, C.c_uint32], C.c_int),
        "fsi_adj_forward": ([P, P, P, C.c_float, C.c_float, P, P, P], C.c_int),
        "fsi_adj_backward": ([P, P, P, P, P, f64p, P], C.c_int),
        "fsi_adj_forward64": ([P, P, P, C.c_double, C.c_double, P, P, P], C.c_int),
        "fsi_adj_backward64": ([P, P, P, P, P, f64p, P], C.c_int),
        "fsi_adj_stats": ([P, u32p], C.c_int),
        "fsi_adj_frame_norms": ([P, f32p], C.c_int),
        "fsi_scene_lattice": ([C.c_uint32, C.c_uint32, C.c_uint32, f32p, C.c_float, C.c_float, C.c_uint32, f32p, P],
                              C.c_int),
        "fsi_mesh_icosphere": ([C.c_uint32, f64p, C.c_double, f64p, i32p], C.c_int),
        "fsi_sdf_build": ([f64p, C.c_uint32, i32p, C.c_uint32, C.c_double, f64p, i32p, f32p], C.c_int),
        "fsi_mesh_create": ([f64p, C.c_uint32, i32p, C.c_uint32, C.c_double, C.POINTER(P)], C.c_int),
        "fsi_mesh_destroy": ([P], C.c_int),
        "fsi_mesh_grid_info": ([P, f64p, f64p, i32p], C.c_int),
        "fsi_mesh_read_sdf": ([P, f32p], C.c_int),
        "fsi_mesh_refit": ([P, f64p], C.c_int),
        "fsi_mesh_refit_dirty": ([P, f64p, C.c_double, C.c_double, i32p, i32p, u32p], C.c_int),
        "fsi_mesh_read_forces": ([P, f32p], C.c_int),
        "fsi_mesh_closest": ([P, f32p, C.c_uint32, i32p, f32p], C.c_int),
        "fsi_pbf_set_sdf_from_mesh": ([P, P], C.c_int),
        "fsi_pbf_contact_forces": ([P, P, C.c_float, C.c_int, C.c_float, f32p, u32p], C.c_int),
        "fsi_mesh_contact_forces_noalloc": ([P, P, C.c_uint32, C.c_float, C.c_int, C.c_float, P, P, P], C.c_int),
        "fsi_deform_params_default": ([C.POINTER(CDeformParams)], C.c_int),
        "fsi_deform_create": ([f64p, C.c_uint32, i32p, C.c_uint32, C.POINTER(P)], C.c_int),
        "fsi_deform_destroy": ([P], C.c_int),
        "fsi_deform_set_mode": ([P, C.c_int, C.c_int, C.c_int], C.c_int),
        "fsi_deform_add_anchor": ([P, C.c_int32], C.c_int),
        "fsi_deform_set_control_points": ([P, i32p, f64p, C.c_uint32], C.c_int),
        "fsi_deform_precompute": ([P], C.c_int),
        "fsi_deform_solve": ([P], C.c_int),
        "fsi_deform_read_vertices": ([P, f64p], C.c_int),
        "fsi_deform_step": ([P, f32p, C.c_float, C.POINTER(CDeformParams), u32p, C.POINTER(C.c_int)], C.c_int),
        "fsi_deform_info": ([P, u32p], C.c_int),
        "fsi_deform_verify_incremental": ([P, f64p], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    _lib = lib
    return lib


def _check(rc: int):
    if rc != FSI_OK:
        raise FsiError(rc, load().fsi_last_error().decode())


def _ptr(a: np.ndarray, ctype):
    return a.ctypes.data_as(C.POINTER(ctype))


def _f3(v) -> C.Array:
    return (C.c_float * 3)(*[float(np.float32(x)) for x in v])


def device_count() -> int:
    n = C.c_int(0)
    _check(load().fsi_device_count(C.byref(n)))
    return n.value


# ---------------------------------------------------------------- scenes

def lattice(nx: int, ny: int, nz: int, origin, spacing: float, jitter: float, seed: int = 1337,
            v0=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Lattice block of nx*ny*nz particles (ids x-fastest) with seeded jitter."""
    out = np.zeros((nx * ny * nz, 12), dtype=np.float32)
    _check(load().fsi_scene_lattice(nx, ny, nz, _f3(origin), float(np.float32(spacing)), float(np.float32(jitter)),
                                    seed, _f3(v0), out.ctypes.data_as(C.c_void_p)))
    return out


def spawn_particles(cfg: "FluidConfig") -> np.ndarray:
    """spawnParticles (src/main.cpp:160-188) of cfg: (n, 12) float32 records."""
    lib = load()
    cc = cfg.to_c()
    n = C.c_uint32(0)
    _check(lib.fsi_spawn_particles(C.byref(cc), None, 0, C.byref(n)))
    out = np.zeros((n.value, 12), dtype=np.float32)
    _check(lib.fsi_spawn_particles(C.byref(cc), out.ctypes.data_as(C.c_void_p), n.value, C.byref(n)))
    return out


def load_off(path: str):
    """loadOFF (src/main.cpp:364-415): (verts (V, 3) float64, faces (F, 3) int32)."""
    lib = load()
    nv, nf = C.c_uint32(0), C.c_uint32(0)
    p = os.fsencode(path)
    _check(lib.fsi_load_off(p, None, 0, None, 0, C.byref(nv), C.byref(nf)))
    verts = np.zeros((nv.value, 3), dtype=np.float64)
    faces = np.zeros((nf.value, 3), dtype=np.int32)
    _check(lib.fsi_load_off(p, _ptr(verts, C.c_double), nv.value, _ptr(faces, C.c_int32), nf.value,
                            C.byref(nv), C.byref(nf)))
    return verts, faces


def fit_mesh_to_domain(verts: np.ndarray, center, target_size: float) -> np.ndarray:
    """fitMeshToDomain (src/main.cpp:419-432) on a copy of verts."""
    v = np.array(verts, dtype=np.float64, order="C", copy=True)
    _check(load().fsi_mesh_fit_to_domain(_ptr(v, C.c_double), len(v), _f3(center), float(target_size)))
    return v


def icosphere(subdivisions: int, centre, radius: float):
    V = 10 * 4 ** subdivisions + 2
    F = 20 * 4 ** subdivisions
    verts = np.zeros((V, 3), dtype=np.float64)
    faces = np.zeros((F, 3), dtype=np.int32)
    c = (C.c_double * 3)(*centre)
    _check(load().fsi_mesh_icosphere(subdivisions, c, float(radius), _ptr(verts, C.c_double), _ptr(faces, C.c_int32)))
    return verts, faces


@dataclass
class SdfGrid:
    grid: np.ndarray      # (nz, ny, nx) float32, x fastest in memory
    origin: np.ndarray    # (3,) float64
    cell: float
    dims: tuple           # (nx, ny, nz)


def sdf_build(verts: np.ndarray, faces: np.ndarray, cell: float) -> SdfGrid:
    verts = np.ascontiguousarray(verts, dtype=np.float64)
    faces = np.ascontiguousarray(faces, dtype=np.int32)
    lib = load()
    origin = np.zeros(3, dtype=np.float64)
    dims = np.zeros(3, dtype=np.int32)
    args = (_ptr(verts, C.c_double), len(verts), _ptr(faces, C.c_int32), len(faces), float(cell),
            _ptr(origin, C.c_double), _ptr(dims, C.c_int32))
    _check(lib.fsi_sdf_build(*args, None))
    grid = np.zeros((int(dims[2]), int(dims[1]), int(dims[0])), dtype=np.float32)
    _check(lib.fsi_sdf_build(*args, _ptr(grid, C.c_float)))
    return SdfGrid(grid=grid, origin=origin, cell=float(cell), dims=(int(dims[0]), int(dims[1]), int(dims[2])))


def deform_params(**kw) -> CDeformParams:
    """deformStep's tuning (src/main.cpp:251, 300-304, 343) with overrides."""
    p = CDeformParams()
    _check(load().fsi_deform_params_default(C.byref(p)))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class Deform:
    """LaplacianDeformation on the host (fsi_deform_*): cotangent Laplacian,
    local patch / global soft / hard-constraint (incremental) solves, and the
    per-frame deformStep driven by contact forces."""

    def __init__(self, verts: np.ndarray, faces: np.ndarray, use_local: bool = True, kring: int = 3,
                 incremental: bool = False):
        self.lib = load()
        v = np.ascontiguousarray(verts, dtype=np.float64)
        f = np.ascontiguousarray(faces, dtype=np.int32)
        self.nv = len(v)
        self.h = C.c_void_p()
        _check(self.lib.fsi_deform_create(_ptr(v, C.c_double), len(v), _ptr(f, C.c_int32), len(f), C.byref(self.h)))
        _check(self.lib.fsi_deform_set_mode(self.h, int(use_local), int(kring), int(incremental)))

    def close(self):
        if self.h:
            self.lib.fsi_deform_destroy(self.h)
            self.h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_mode(self, use_local: bool, kring: int = 3, incremental: bool = False):
        _check(self.lib.fsi_deform_set_mode(self.h, int(use_local), int(kring), int(incremental)))

    def add_anchor(self, v: int):
        _check(self.lib.fsi_deform_add_anchor(self.h, int(v)))

    def set_control_points(self, idx, targets):
        i = np.ascontiguousarray(idx, dtype=np.int32).reshape(-1)
        t = np.ascontiguousarray(targets, dtype=np.float64).reshape(-1, 3)
        _check(self.lib.fsi_deform_set_control_points(self.h, _ptr(i, C.c_int32), _ptr(t, C.c_double), len(i)))

    def precompute(self):
        _check(self.lib.fsi_deform_precompute(self.h))

    def solve(self):
        _check(self.lib.fsi_deform_solve(self.h))

    def vertices(self) -> np.ndarray:
        out = np.zeros((self.nv, 3), dtype=np.float64)
        _check(self.lib.fsi_deform_read_vertices(self.h, _ptr(out, C.c_double)))
        return out

    def step(self, forces: np.ndarray, dt: float, params: Optional[CDeformParams] = None):
        """deformStep: returns (handle count, solved)."""
        f = np.ascontiguousarray(forces, dtype=np.float32).reshape(self.nv, 3)
        p = params if params is not None else deform_params()
        nh, solved = C.c_uint32(0), C.c_int(0)
        _check(self.lib.fsi_deform_step(self.h, _ptr(f, C.c_float), float(np.float32(dt)), C.byref(p), C.byref(nh),
                                        C.byref(solved)))
        return nh.value, bool(solved.value)

    def info(self) -> dict:
        out = np.zeros(4, dtype=np.uint32)
        _check(self.lib.fsi_deform_info(self.h, _ptr(out, C.c_uint32)))
        return {"patch": int(out[0]), "factor_steps": int(out[1]), "handles": int(out[2]), "factorizations": int(out[3])}

    def verify_incremental(self) -> float:
        d = C.c_double(0.0)
        _check(self.lib.fsi_deform_verify_incremental(self.h, C.byref(d)))
        return d.value


class Mesh:
    """The obstacle mesh on the device (SDFBoundary with its grid in HBM):
    signed-distance grid, refit after a deformation, contact forces."""

    def __init__(self, verts: np.ndarray, faces: np.ndarray, cell: float):
        self.lib = load()
        self.verts = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1, 3)
        self.faces = np.ascontiguousarray(faces, dtype=np.int32).reshape(-1, 3)
        h = C.c_void_p()
        _check(self.lib.fsi_mesh_create(_ptr(self.verts, C.c_double), len(self.verts), _ptr(self.faces, C.c_int32),
                                        len(self.faces), float(cell), C.byref(h)))
        self.h = h
        origin = np.zeros(3, dtype=np.float64)
        c = np.zeros(1, dtype=np.float64)
        dims = np.zeros(3, dtype=np.int32)
        _check(self.lib.fsi_mesh_grid_info(h, _ptr(origin, C.c_double), _ptr(c, C.c_double), _ptr(dims, C.c_int32)))
        self.origin = origin
        self.cell = float(c[0])
        self.dims = (int(dims[0]), int(dims[1]), int(dims[2]))

    def read_sdf(self) -> SdfGrid:
        grid = np.zeros((self.dims[2], self.dims[1], self.dims[0]), dtype=np.float32)
        _check(self.lib.fsi_mesh_read_sdf(self.h, _ptr(grid, C.c_float)))
        return SdfGrid(grid=grid, origin=self.origin.copy(), cell=self.cell, dims=self.dims)

    def refit(self, verts: np.ndarray):
        v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1, 3)
        if v.shape != self.verts.shape:
            raise ValueError("refit keeps the vertex count")
        _check(self.lib.fsi_mesh_refit(self.h, _ptr(v, C.c_double)))
        self.verts = v.copy()

    def refit_dirty(self, verts: np.ndarray, band_radius: float, move_eps: float):
        """Returns (recomputed voxels, box min xyz, box max xyz)."""
        v = np.ascontiguousarray(verts, dtype=np.float64).reshape(-1, 3)
        if v.shape != self.verts.shape:
            raise ValueError("refit keeps the vertex count")
        lo = np.zeros(3, dtype=np.int32)
        hi = np.zeros(3, dtype=np.int32)
        n = C.c_uint32(0)
        _check(self.lib.fsi_mesh_refit_dirty(self.h, _ptr(v, C.c_double), float(band_radius), float(move_eps),
                                             _ptr(lo, C.c_int32), _ptr(hi, C.c_int32), C.byref(n)))
        self.verts = v.copy()
        return n.value, lo, hi

    def read_forces(self) -> np.ndarray:
        out = np.zeros((len(self.verts), 3), dtype=np.float32)
        _check(self.lib.fsi_mesh_read_forces(self.h, _ptr(out, C.c_float)))
        return out

    def closest(self, points: np.ndarray):
        """SDFBoundary::closestTriangle of every point: (face index per point, closest points (n, 3) float32)."""
        p = np.ascontiguousarray(points, dtype=np.float32).reshape(-1, 3)
        faces = np.zeros(len(p), dtype=np.int32)
        out = np.zeros((len(p), 3), dtype=np.float32)
        _check(self.lib.fsi_mesh_closest(self.h, _ptr(p, C.c_float), len(p), _ptr(faces, C.c_int32),
                                         _ptr(out, C.c_float)))
        return faces, out

    def close(self):
        if self.h:
            self.lib.fsi_mesh_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- solver

class Solver:
    """One solver handle (the reference's PBFluids)."""

    def __init__(self, cfg: FluidConfig):
        self.lib = load()
        self.cfg = cfg
        self._c = cfg.to_c()
        h = C.c_void_p()
        _check(self.lib.fsi_pbf_create(C.byref(self._c), C.byref(h)))
        self.h = h
        self.n = 0

    def close(self):
        if self.h:
            self.lib.fsi_pbf_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_params(self, cfg: FluidConfig):
        self.cfg = cfg
        self._c = cfg.to_c()
        _check(self.lib.fsi_pbf_set_params(self.h, C.byref(self._c)))

    def constants(self) -> CConstants:
        k = CConstants()
        _check(self.lib.fsi_pbf_get_constants(self.h, C.byref(k)))
        return k

    def set_particles(self, particles: np.ndarray):
        p = np.ascontiguousarray(particles, dtype=np.float32).reshape(-1, 12)
        self.n = p.shape[0]
        _check(self.lib.fsi_pbf_set_particles(self.h, p.ctypes.data_as(C.c_void_p), self.n))

    def read_particles(self) -> np.ndarray:
        out = np.zeros((self.n, 12), dtype=np.float32)
        n = C.c_uint32(0)
        _check(self.lib.fsi_pbf_read_particles(self.h, out.ctypes.data_as(C.c_void_p), self.n, C.byref(n)))
        return out

    def write_render_buffer(self, device_ptr: int, stream: int = 0):
        """bindParticlesForRendering: the 48-byte records into device memory at
        device_ptr (n * 48 bytes); `stream` (a hipStream_t as int) waits for it."""
        _check(self.lib.fsi_pbf_write_render_buffer(self.h, C.c_void_p(device_ptr), C.c_void_p(stream or None)))

    def set_collision_padding(self, p: float):
        _check(self.lib.fsi_pbf_set_collision_padding(self.h, float(p)))

    def set_bounds(self, mn, mx, damping: float = 0.0):
        _check(self.lib.fsi_pbf_set_bounds(self.h, _f3(mn), _f3(mx), float(np.float32(damping))))

    def set_sdf(self, sdf: Optional[SdfGrid], padding: float = 0.0, containment: bool = False):
        if sdf is None:
            _check(self.lib.fsi_pbf_set_sdf(self.h, None, None, None, 0.0))
            return
        g = np.ascontiguousarray(sdf.grid, dtype=np.float32)
        dims = (C.c_int32 * 3)(*sdf.dims)
        _check(self.lib.fsi_pbf_set_sdf(self.h, _ptr(g, C.c_float), dims, _f3(sdf.origin),
                                        float(np.float32(sdf.cell))))
        _check(self.lib.fsi_pbf_set_sdf_padding(self.h, float(np.float32(padding))))
        _check(self.lib.fsi_pbf_set_sdf_containment(self.h, 1 if containment else 0))

    def set_interaction(self, active: bool, pos=(0.0, 0.0, 0.0)):
        _check(self.lib.fsi_pbf_set_interaction(self.h, 1 if active else 0, _f3(pos)))

    def set_camera(self, pos):
        _check(self.lib.fsi_pbf_set_camera(self.h, _f3(pos)))

    def step(self, frames: int = 1):
        for _ in range(frames):
            _check(self.lib.fsi_pbf_step(self.h))

    def sync(self):
        _check(self.lib.fsi_pbf_sync(self.h))

    def filter_contacts(self, band: float):
        _check(self.lib.fsi_pbf_filter_contacts(self.h, float(np.float32(band))))

    def set_sdf_from_mesh(self, mesh: "Mesh", padding: float = 0.0, containment: bool = False):
        _check(self.lib.fsi_pbf_set_sdf_from_mesh(self.h, mesh.h))
        _check(self.lib.fsi_pbf_set_sdf_padding(self.h, float(np.float32(padding))))
        _check(self.lib.fsi_pbf_set_sdf_containment(self.h, 1 if containment else 0))

    def contact_forces(self, mesh: "Mesh", band: float, use_vel_magnitude: bool = True,
                       control_threshold: float = 0.0):
        """computeContactForces on the device: ((nv, 3) forces, (contacts, control points))."""
        out = np.zeros((len(mesh.verts), 3), dtype=np.float32)
        counts = np.zeros(2, dtype=np.uint32)
        _check(self.lib.fsi_pbf_contact_forces(self.h, mesh.h, float(np.float32(band)), 1 if use_vel_magnitude else 0,
                                               float(np.float32(control_threshold)), _ptr(out, C.c_float),
                                               _ptr(counts, C.c_uint32)))
        return out, (int(counts[0]), int(counts[1]))

    def read_contacts(self) -> np.ndarray:
        count = C.c_uint32(0)
        _check(self.lib.fsi_pbf_read_contacts(self.h, None, 0, C.byref(count)))
        out = np.zeros((count.value, 8), dtype=np.float32)
        if count.value:
            _check(self.lib.fsi_pbf_read_contacts(self.h, out.ctypes.data_as(C.c_void_p), count.value,
                                                  C.byref(count)))
        return out

    def read_grid(self):
        keys = np.zeros(self.n, dtype=np.uint32)
        ids = np.zeros(self.n, dtype=np.uint32)
        cells = np.zeros((self.cfg.hash_size, 2), dtype=np.int32)
        _check(self.lib.fsi_pbf_read_grid(self.h, _ptr(keys, C.c_uint32), _ptr(ids, C.c_uint32),
                                          _ptr(cells, C.c_int32)))
        return keys, ids, cells

    def read_solver(self):
        a = np.zeros((self.n, 4), dtype=np.float32)
        b = np.zeros((self.n, 4), dtype=np.float32)
        _check(self.lib.fsi_pbf_read_solver(self.h, _ptr(a, C.c_float), _ptr(b, C.c_float)))
        return a, b

    def set_taps(self, on: bool):
        _check(self.lib.fsi_pbf_set_taps(self.h, 1 if on else 0))

    def read_counts(self):
        c = np.zeros(self.n, dtype=np.uint32)
        nb = np.zeros(self.n, dtype=np.uint32)
        _check(self.lib.fsi_pbf_read_counts(self.h, _ptr(c, C.c_uint32), _ptr(nb, C.c_uint32)))
        return c, nb

    def set_tuning(self, knob: int, value: int):
        _check(self.lib.fsi_pbf_set_tuning(self.h, knob, value))

    def get_tuning(self, knob: int) -> int:
        v = C.c_int(0)
        _check(self.lib.fsi_pbf_get_tuning(self.h, knob, C.byref(v)))
        return v.value

    def set_graph(self, on: bool):
        _check(self.lib.fsi_pbf_set_graph(self.h, 1 if on else 0))

    def get_stream(self) -> int:
        """hipStream_t the handle enqueues on, as an integer."""
        s = C.c_void_p()
        _check(self.lib.fsi_pbf_get_stream(self.h, C.byref(s)))
        return s.value or 0

    def set_stream(self, stream: int = 0):
        """Enqueue on the given hipStream_t of this device (0: the handle's own)."""
        _check(self.lib.fsi_pbf_set_stream(self.h, C.c_void_p(stream or None)))

    def autotune(self, frames: int = 3):
        """Time every (gather form, sort digit bits, workgroup size / tile shape) setting; keeps the fastest and
        returns (gather form, digit bits); knobs() has the rest. Particle state is restored."""
        g, d = C.c_int(0), C.c_int(0)
        _check(self.lib.fsi_pbf_autotune(self.h, frames, C.byref(g), C.byref(d)))
        return g.value, d.value

    def knobs(self) -> dict:
        """Every launch-shape knob's current value."""
        return {name: self.get_tuning(k) for name, k in KNOBS.items()}

    def autotune_trials(self) -> list:
        """The settings the last autotune timed, with their milliseconds."""
        return _read_trials(self.lib.fsi_pbf_read_autotune if isinstance(self, Solver)
                            else self.lib.fsi_batch_read_autotune, self.h)

    def set_timing(self, on: bool):
        _check(self.lib.fsi_pbf_set_timing(self.h, 1 if on else 0))

    def set_timing_passes(self, names):
        """Time only the named passes (see PASS_NAMES)."""
        mask = 0
        for nm in names:
            mask |= 1 << PASS_NAMES.index(nm)
        _check(self.lib.fsi_pbf_set_timing_passes(self.h, mask))

    def timing_reset(self):
        _check(self.lib.fsi_pbf_timing_reset(self.h))

    def read_timing(self) -> dict:
        t = CTiming()
        _check(self.lib.fsi_pbf_read_timing(self.h, C.byref(t)))
        return _timing_dict(t)


def _timing_dict(t: CTiming) -> dict:
    return {
        "ms": {PASS_NAMES[i]: t.ms[i] for i in range(8)},
        "launches": {PASS_NAMES[i]: int(t.launches[i]) for i in range(8)},
        "substeps": int(t.substeps),
        "frames": int(t.frames),
    }


@dataclass
class Scene:
    """One scene of a Batch: its particles, cell-table size, box and optional SDF."""
    particles: np.ndarray
    hash_size: int
    bounds_min: tuple
    bounds_max: tuple
    sdf: Optional[SdfGrid] = None
    sdf_padding: float = 0.0
    sdf_containment: bool = False


class Batch:
    """Many independent scenes advanced together, one launch per pass.

    Every scene follows the same frame as a ``Solver`` (PBFluids::step,
    src/fluid/PBFluids.cpp:323-517) with the shared ``FluidConfig`` (its
    particle_count, hash_size and bounds are taken per scene; bounds are the
    box after padding).
    """

    def __init__(self, cfg: FluidConfig):
        self.lib = load()
        self.cfg = cfg
        self._c = cfg.to_c()
        h = C.c_void_p()
        _check(self.lib.fsi_batch_create(C.byref(self._c), C.byref(h)))
        self.h = h
        self.counts: list = []
        self.offsets: list = []
        self.hash_sizes: list = []
        self.n = 0

    def close(self):
        if self.h:
            self.lib.fsi_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_scenes(self, scenes: list):
        parts = [np.ascontiguousarray(s.particles, dtype=np.float32).reshape(-1, 12) for s in scenes]
        allp = np.ascontiguousarray(np.concatenate(parts, axis=0)) if parts else np.zeros((0, 12), np.float32)
        descs = (CSceneDesc * len(scenes))()
        keep = []
        off = 0
        self.counts, self.offsets, self.hash_sizes = [], [], []
        for i, (s, p) in enumerate(zip(scenes, parts)):
            d = descs[i]
            d.particle_offset = off
            d.particle_count = p.shape[0]
            d.hash_size = int(s.hash_size)
            for a in range(3):
                d.bounds_min[a] = float(np.float32(s.bounds_min[a]))
                d.bounds_max[a] = float(np.float32(s.bounds_max[a]))
            if s.sdf is not None:
                g = np.ascontiguousarray(s.sdf.grid, dtype=np.float32)
                keep.append(g)
                d.sdf_grid = _ptr(g, C.c_float)
                for a in range(3):
                    d.sdf_dims[a] = s.sdf.dims[a]
                    d.sdf_origin[a] = float(np.float32(s.sdf.origin[a]))
                d.sdf_cell = float(np.float32(s.sdf.cell))
                d.sdf_padding = float(np.float32(s.sdf_padding))
                d.sdf_containment = 1 if s.sdf_containment else 0
            self.counts.append(p.shape[0])
            self.offsets.append(off)
            self.hash_sizes.append(int(s.hash_size))
            off += p.shape[0]
        self.n = off
        _check(self.lib.fsi_batch_set_scenes(self.h, allp.ctypes.data_as(C.c_void_p), off, descs, len(scenes)))

    def step(self, frames: int = 1):
        for _ in range(frames):
            _check(self.lib.fsi_batch_step(self.h))

    def sync(self):
        _check(self.lib.fsi_batch_sync(self.h))

    def read_particles(self) -> np.ndarray:
        out = np.zeros((self.n, 12), dtype=np.float32)
        n = C.c_uint32(0)
        _check(self.lib.fsi_batch_read_particles(self.h, out.ctypes.data_as(C.c_void_p), self.n, C.byref(n)))
        return out

    def scene_particles(self, out: np.ndarray, s: int) -> np.ndarray:
        return out[self.offsets[s]:self.offsets[s] + self.counts[s]]

    def read_grid(self, s: int):
        keys = np.zeros(self.counts[s], dtype=np.uint32)
        ids = np.zeros(self.counts[s], dtype=np.uint32)
        cells = np.zeros((self.hash_sizes[s], 2), dtype=np.int32)
        _check(self.lib.fsi_batch_read_grid(self.h, s, _ptr(keys, C.c_uint32), _ptr(ids, C.c_uint32),
                                            _ptr(cells, C.c_int32)))
        return keys, ids, cells

    def set_graph(self, on: bool):
        _check(self.lib.fsi_batch_set_graph(self.h, 1 if on else 0))

    def set_tuning(self, knob: int, value: int):
        _check(self.lib.fsi_batch_set_tuning(self.h, knob, value))

    def get_tuning(self, knob: int) -> int:
        v = C.c_int(0)
        _check(self.lib.fsi_batch_get_tuning(self.h, knob, C.byref(v)))
        return v.value

    def autotune(self, frames: int = 3):
        """As Solver.autotune, over every scene at once."""
        g, d = C.c_int(0), C.c_int(0)
        _check(self.lib.fsi_batch_autotune(self.h, frames, C.byref(g), C.byref(d)))
        return g.value, d.value

    def knobs(self) -> dict:
        """Every launch-shape knob's current value."""
        return {name: self.get_tuning(k) for name, k in KNOBS.items()}

    def autotune_trials(self) -> list:
        """The settings the last autotune timed, with their milliseconds."""
        return _read_trials(self.lib.fsi_pbf_read_autotune if isinstance(self, Solver)
                            else self.lib.fsi_batch_read_autotune, self.h)

    def set_timing(self, on: bool):
        _check(self.lib.fsi_batch_set_timing(self.h, 1 if on else 0))

    def timing_reset(self):
        _check(self.lib.fsi_batch_timing_reset(self.h))

    def read_timing(self) -> dict:
        t = CTiming()
        _check(self.lib.fsi_batch_read_timing(self.h, C.byref(t)))
        return _timing_dict(t)
