"""Pins of the CPU oracle against facts that do not come from its own code:
known-answer hashes, closed forms of the SPH kernels and of the coefficient
block, a library stable sort, O(N^2) brute-force neighbour visits (with the
reference's aliased-bucket duplicates), isolated-particle / particle-pair /
ballistic / wall closed forms, the conservation of the position corrections,
and exactness of trilinear sampling on linear fields.  The reference ships no
tests or golden vectors for this path, so these are the pins."""
import ctypes as C
import math

import numpy as np
import pytest

from conftest import dam_break_config


def O():
    import pyoracle

    return pyoracle


def fsi():
    import fsi_pbf

    return fsi_pbf


def call3(fn, u, r):
    out = (C.c_double * 3)()
    fn(C.byref(u), (C.c_double * 3)(*r), out)
    return np.array(out[:])


# ---------------------------------------------------------------- hash


def test_hash_known_answers():
    h = O().load().oracle_cell_hash
    m = 0xFFFFFFFF
    assert h(1, 0, 0, m) == 73856093
    assert h(0, 1, 0, m) == 19349663
    assert h(0, 0, 1, m) == 83492791
    assert h(1, 1, 1, m) == 73856093 ^ 19349663 ^ 83492791
    assert h(-1, 0, 0, m) == 2 ** 32 - 73856093
    # 30 * 73856093 overflows int32; the wrapped bits are kept
    assert h(30, 0, 0, m) == (30 * 73856093) % 2 ** 32
    assert h(-7, 5, -3, m) == ((-7 * 73856093) % 2 ** 32) ^ ((5 * 19349663) % 2 ** 32) ^ ((-3 * 83492791) % 2 ** 32)
    assert h(1, 1, 1, (1 << 17) - 1) == (73856093 ^ 19349663 ^ 83492791) & ((1 << 17) - 1)
    assert h(0, 0, 0, m) == 0


# ---------------------------------------------------------------- constants and kernels


@pytest.mark.parametrize("h,rho0,dq", [(0.1, 6378.0, 0.3), (0.15, 1200.0, 0.3), (0.137, 999.0, 0.2)])
def test_constants_closed_forms(h, rho0, dq):
    cfg = fsi().FluidConfig(h=h, rho0=rho0, delta_q=dq, dt=1 / 120, substep_iterations=2)
    u = O().pack_ubo(cfg, 10, "f32")
    hf = float(np.float32(h))
    assert u.w0_self == pytest.approx(315 / (64 * math.pi * hf ** 3), rel=3e-7)
    assert u.poly6_coeff == pytest.approx(315 / (64 * math.pi * hf ** 9), rel=6e-7)
    assert u.spiky_coeff == pytest.approx(45 / (math.pi * hf ** 6), rel=5e-7)
    assert u.inv_rho0 == pytest.approx(1 / float(np.float32(rho0)), rel=1e-7)
    r = float(np.float32(np.float32(dq) * np.float32(h)))
    assert u.wq == pytest.approx(315 / (64 * math.pi * hf ** 9) * (hf * hf - r * r) ** 3, rel=1e-6)
    assert u.sub_dt == pytest.approx(1 / 240, rel=1e-7)
    assert u.hash_mask == (1 << 17) - 1
    # the FP32 coefficients are FP32 values
    for v in (u.w0_self, u.poly6_coeff, u.spiky_coeff, u.wq, u.sub_dt):
        assert float(np.float32(v)) == v


def test_kernel_closed_forms_and_normalisation():
    lib = O().load()
    cfg = dam_break_config()
    u = O().pack_ubo(cfg, 1, "f32")
    h = float(np.float32(0.1))
    poly6 = lambda r: lib.oracle_poly6_f32(C.byref(u), (C.c_double * 3)(*r))
    assert poly6((0, 0, 0)) == pytest.approx(315 / (64 * math.pi * h ** 3), rel=1e-6)
    assert poly6((h * 1.0001, 0, 0)) == 0.0
    assert poly6((0.05, 0.0, 0.0)) == pytest.approx(315 / (64 * math.pi * h ** 9) * (h * h - 0.0025) ** 3, rel=1e-5)
    for r in ([0.03, -0.04, 0.0], [-0.01, 0.02, 0.06]):
        g = call3(lib.oracle_grad_spiky_f32, u, r)
        d = np.linalg.norm(r)
        mag = 45 / (math.pi * h ** 6) * (h - d) ** 2
        np.testing.assert_allclose(g, -mag * np.array(r) / d, rtol=2e-5)
    assert np.all(call3(lib.oracle_grad_spiky_f32, u, [1e-6, 0, 0]) == 0.0)  # r^2 < 1e-10
    assert np.all(call3(lib.oracle_grad_spiky_f32, u, [0.2, 0, 0]) == 0.0)
    # integral of poly6 over its support is 1 (midpoint rule, step h/40)
    s = h / 40
    ax = (np.arange(-40, 40) + 0.5) * s
    X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
    r2 = X ** 2 + Y ** 2 + Z ** 2
    W = np.where(r2 <= h * h, u.poly6_coeff * (h * h - r2) ** 3, 0.0)
    assert W.sum() * s ** 3 == pytest.approx(1.0, abs=2e-3)
    # and the oracle's own poly6 agrees on a sample of those points
    rng = np.random.default_rng(0)
    idx = rng.integers(0, X.size, 200)
    for i in idx:
        p = (X.flat[i], Y.flat[i], Z.flat[i])
        assert poly6(p) == pytest.approx(W.flat[i], rel=2e-5, abs=1e-3)


# ---------------------------------------------------------------- sort, cells, neighbours


def random_scene(n, box, seed):
    rng = np.random.default_rng(seed)
    p = np.zeros((n, 12), dtype=np.float32)
    p[:, 0:3] = rng.uniform(0.05, 1.0, (n, 3)) * np.asarray(box, dtype=np.float32)
    p[:, 4:7] = p[:, 0:3]
    p[:, 8:11] = rng.normal(0, 0.3, (n, 3))
    return p


@pytest.mark.parametrize("H", [1 << 4, 1 << 10, 1 << 17])
def test_sort_and_cells_match_library_stable_sort(H):
    cfg = dam_break_config(hash_size=H, box=(1.0, 1.0, 1.0))
    parts = random_scene(3000, (1.0, 1.0, 1.0), H)
    sim = O().OracleSim(cfg, parts)
    sim.predict()
    raw = sim.keys.copy()
    assert np.array_equal(sim.ids, np.arange(3000))
    sim.sort()
    order = np.argsort(raw, kind="stable")
    assert np.array_equal(sim.ids, order.astype(np.uint32))
    assert np.array_equal(sim.keys, raw[order])
    sim.build_cells()
    occ = sim.cells[:, 0] >= 0
    assert np.array_equal(occ, sim.cells[:, 1] >= 0)
    assert np.array_equal(np.nonzero(occ)[0], np.unique(raw))
    for k in np.unique(raw)[:50]:
        idx = np.nonzero(sim.keys == k)[0]
        assert sim.cells[k, 0] == idx[0] and sim.cells[k, 1] == idx[-1] + 1


@pytest.mark.parametrize("H", [1 << 3, 1 << 6, 1 << 16])
def test_neighbour_visits_brute_force(H):
    """O(N^2): particle i visits j once per bucket among its 27 (duplicates
    when buckets alias) that holds j's key."""
    cfg = dam_break_config(hash_size=H, box=(1.0, 1.0, 1.0))
    parts = random_scene(400, (0.4, 0.4, 0.4), 5)
    sim = O().OracleSim(cfg, parts, taps=True)
    sim.predict()
    sim.sort()
    sim.build_cells()
    sim.lambda_pass(1)
    x = sim.solver[:, 0:3].astype(np.float32)
    h = np.float32(0.1)
    cell = np.floor(x / h).astype(np.int64)
    mask = H - 1
    prim = np.array([73856093, 19349663, 83492791], dtype=np.int64)

    def key(c):
        v = [(int(c[a]) * int(prim[a])) % 2 ** 32 for a in range(3)]
        return (v[0] ^ v[1] ^ v[2]) & mask

    pkeys = np.array([key(c) for c in cell])
    offs = [(dx, dy, dz) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]
    for i in range(len(x)):
        visit = [key(cell[i] + np.array(o)) for o in offs]
        mult = np.array([visit.count(k) for k in pkeys])
        mult[i] = 0
        d = x[i] - x
        r2 = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]) + d[:, 2] * d[:, 2]
        assert sim.cand[i] == mult.sum()
        assert sim.nbr[i] == mult[r2 <= h * h].sum()


def test_lattice_interior_has_32_neighbours():
    """Unjittered lattice at spacing 0.049 < h/2: lattice vectors n with
    |n|^2 <= 4 lie within h, |n|^2 = 5 does not; 33 points minus self."""
    cfg = dam_break_config(hash_size=1 << 20, box=(2.0, 2.0, 2.0), gravity=(0.0, 0.0, 0.0))
    parts = fsi().lattice(12, 12, 12, (0.3, 0.3, 0.3), 0.049, 0.0)
    sim = O().OracleSim(cfg, parts, taps=True)
    sim.predict()
    sim.sort()
    sim.build_cells()
    sim.lambda_pass(1)
    ijk = np.stack(np.meshgrid(np.arange(12), np.arange(12), np.arange(12), indexing="ij"), -1)
    ijk = ijk.transpose(2, 1, 0, 3).reshape(-1, 3)  # ids x-fastest
    interior = np.all((ijk >= 2) & (ijk <= 9), axis=1)
    assert np.all(sim.nbr[interior] == 32)


# ---------------------------------------------------------------- closed forms of the passes


def test_isolated_particle():
    cfg = dam_break_config(substeps=1, n_iter=3, enable_vorticity=False)
    parts = fsi().lattice(1, 1, 1, (1.0, 1.0, 1.0), 0.1, 0.0)
    parts[0, 8:11] = [0.3, -0.2, 0.1]
    sim = O().OracleSim(cfg, parts, taps=True)
    sim.predict()
    sim.sort()
    sim.build_cells()
    sim.lambda_pass(1)
    u = sim.ubo
    rho = sim.solver[0, 7]
    assert rho == np.float32(u.w0_self)
    C_ = max(u.w0_self / u.rho0 - 1.0, -u.cohesion_strength)
    assert sim.solver[0, 3] == pytest.approx(-C_ / u.eps, rel=1e-6)
    sim.deltap_pass(1)
    assert np.all(sim.solver[0, 4:7] == 0.0)
    # a whole step is ballistic: v1 = v0 + g dt, x1 = x0 + v1 dt
    sim2 = O().OracleSim(cfg, parts)
    sim2.step(1)
    dt = 1 / 120
    v1 = np.array([0.3, -0.2 - 9.81 * dt, 0.1])
    x1 = np.array([1.05, 1.05, 1.05]) + v1 * dt
    np.testing.assert_allclose(sim2.pos[0, 0:3], x1, rtol=0, atol=2e-6)
    np.testing.assert_allclose(sim2.vel[0, 0:3], v1, rtol=0, atol=2e-4)


def test_particle_pair_symmetry():
    cfg = dam_break_config(substeps=1, enable_scorr=True, hash_size=1 << 20)
    parts = np.zeros((2, 12), dtype=np.float32)
    parts[0, 0:3] = [1.0, 1.0, 1.0]
    parts[1, 0:3] = [1.03, 1.02, 0.99]
    parts[:, 4:7] = parts[:, 0:3]
    cfg.gravity = (0.0, 0.0, 0.0)
    sim = O().OracleSim(cfg, parts)
    sim.predict()
    sim.sort()
    sim.build_cells()
    sim.lambda_pass(1)
    sim.deltap_pass(1)
    u = sim.ubo
    r = np.linalg.norm(parts[0, 0:3].astype(np.float64) - parts[1, 0:3])
    h = 0.1
    W = u.poly6_coeff * (h * h - r * r) ** 3
    assert sim.solver[0, 7] == pytest.approx(u.w0_self + W, rel=2e-5)
    assert sim.solver[0, 7] == sim.solver[1, 7]
    assert sim.solver[0, 3] == sim.solver[1, 3]
    np.testing.assert_array_equal(sim.solver[0, 4:7], -sim.solver[1, 4:7])
    # dp_0 is parallel to x0 - x1 (the spiky gradient direction)
    d = (parts[0, 0:3] - parts[1, 0:3]).astype(np.float64)
    c = np.cross(sim.solver[0, 4:7].astype(np.float64), d)
    assert np.abs(c).max() <= 1e-5 * np.linalg.norm(sim.solver[0, 4:7]) * np.linalg.norm(d)


def test_position_corrections_sum_to_zero():
    """Sum_i dp_i = 0: each pair adds g_ij (l_i + l_j + s_ij) and its mirror
    g_ji = -g_ij (no aliasing at H = 2^22, no clamping before apply)."""
    cfg = dam_break_config(hash_size=1 << 22, box=(1.0, 1.0, 1.0))
    parts = fsi().lattice(10, 10, 10, (0.2, 0.2, 0.2), 0.05, 0.002, 3)
    sim = O().OracleSim(cfg, parts, precision="f64")
    sim.predict()
    sim.sort()
    sim.build_cells()
    sim.lambda_pass(1)
    sim.deltap_pass(1)
    dp = sim.solver[:, 4:7]
    assert np.abs(dp.sum(0)).max() <= 1e-12 * np.abs(dp).sum()


def test_wall_clamp_and_damping():
    cfg = dam_break_config(substeps=1, n_iter=1, box=(1.0, 1.0, 1.0), bound_damping=0.25, gravity=(0.0, 0.0, 0.0),
                           enable_viscosity=False, enable_vorticity=False)
    parts = np.zeros((1, 12), dtype=np.float32)
    parts[0, 0:3] = [0.99, 0.5, 0.5]
    parts[0, 8:11] = [6.0, 0.0, 0.0]  # crosses x = 1 within one substep
    sim = O().OracleSim(cfg, parts)
    sim.step(1)
    assert sim.pos[0, 0] == np.float32(1.0)
    v_pos = (np.float32(1.0) - np.float32(0.99)) * np.float32(120.0)
    assert sim.vel[0, 0] == pytest.approx(-0.25 * v_pos, rel=1e-5)
    assert sim.vel[0, 1] == 0.0 and sim.vel[0, 2] == 0.0


# ---------------------------------------------------------------- SDF sampling


def linear_sdf(dims, origin, cell, a, b):
    import fsi_pbf

    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    p = np.stack([x, y, z], -1) * cell + origin
    g = (p @ np.asarray(a) + b).astype(np.float32)
    return fsi_pbf.SdfGrid(grid=g, origin=np.asarray(origin, dtype=np.float64), cell=cell, dims=dims)


def test_trilinear_is_exact_on_linear_fields_and_clamps():
    pyo = O()
    lib = pyo.load()
    sdf = linear_sdf((20, 16, 12), (0.25, 0.5, 0.75), 0.0625, [0.5, -0.25, 0.125], 0.1)
    cfg = dam_break_config()
    sim = pyo.OracleSim(cfg, np.zeros((1, 12), np.float32), precision="f64", sdf=sdf)
    rng = np.random.default_rng(1)
    lo = np.array([0.25, 0.5, 0.75])
    hi = lo + (np.array([20, 16, 12]) - 1) * 0.0625
    for _ in range(200):
        p = rng.uniform(lo, hi)
        v = lib.oracle_sdf_query_f64(C.byref(sim.uni), (C.c_double * 3)(*p))
        assert v == pytest.approx(p @ [0.5, -0.25, 0.125] + 0.1, abs=2e-6)
    # clamp to edge: beyond the last voxel centre the value stays at the border
    far = hi + [1.0, 0.0, 0.0]
    edge = hi.copy()
    v_far = lib.oracle_sdf_query_f64(C.byref(sim.uni), (C.c_double * 3)(*far))
    v_edge = lib.oracle_sdf_query_f64(C.byref(sim.uni), (C.c_double * 3)(*edge))
    assert v_far == pytest.approx(v_edge, abs=1e-6)
    # the central-difference gradient of a linear field is its slope
    p = (lo + hi) / 2
    g = (C.c_double * 3)()
    lib.oracle_sdf_gradient_f64(C.byref(sim.uni), (C.c_double * 3)(*p), g)
    np.testing.assert_allclose(g[:], [0.5, -0.25, 0.125], atol=1e-5)


def test_sdf_projection_pushes_out_of_sphere():
    """A particle inside the obstacle ends at distance >= padding (within
    the trilinear error) after apply."""
    f = fsi()
    v, fc = f.icosphere(4, (1.0, 1.0, 1.0), 0.5)
    sdf = f.sdf_build(v, fc, 0.05)
    cfg = dam_break_config(substeps=1, n_iter=1, gravity=(0.0, 0.0, 0.0), box=(2.0, 2.0, 2.0))
    parts = np.zeros((3, 12), np.float32)
    parts[:, 0:3] = [[1.0, 1.45, 1.0], [1.3, 1.0, 1.3], [0.2, 0.2, 0.2]]
    sim = O().OracleSim(cfg, parts, sdf=sdf, sdf_padding=0.025)
    sim.step(1)
    d = np.linalg.norm(sim.pos[:, 0:3] - 1.0, axis=1) - 0.5
    assert np.all(d[:2] >= 0.025 - 3e-3)
    np.testing.assert_allclose(sim.pos[2, 0:3], [0.2, 0.2, 0.2], atol=1e-7)


# ---------------------------------------------------------------- FP64 twin and density error


def test_fp64_twin_tracks_fp32_over_one_frame():
    cfg = dam_break_config()
    parts = fsi().lattice(16, 16, 16, (0.0, 0.0, 0.0), 0.05, 0.001, 1337)
    a = O().OracleSim(cfg, parts, "f32")
    b = O().OracleSim(cfg, parts, "f64")
    a.step(1)
    b.step(1)
    assert np.abs(a.pos[:, 0:3] - b.pos[:, 0:3]).max() < 1e-4 * cfg.h
    ea = O().density_error_f64(a.pos, cfg)
    eb = O().density_error_f64(b.pos, cfg)
    assert ea["mean_abs_C"] == pytest.approx(eb["mean_abs_C"], rel=1e-3)


def test_density_error_of_isolated_particles_is_the_cohesion_floor():
    cfg = dam_break_config()
    x = np.array([[1.0, 1.0, 1.0], [2.0, 2.0, 2.0]], dtype=np.float32)
    e = O().density_error_f64(x, cfg)
    # rho = W(0) << rho0, so C is clamped to -cohesion
    assert e["mean_abs_C"] == pytest.approx(0.001, rel=1e-6)
    assert e["mean_rho"] == pytest.approx(315 / (64 * math.pi * 0.1 ** 3), rel=1e-6)


def test_contact_filter_band_and_order():
    sdf = linear_sdf((16, 16, 16), (0.0, 0.0, 0.0), 0.0625, [1.0, 0.0, 0.0], -0.5)  # d = x - 0.5
    cfg = dam_break_config()
    parts = random_scene(500, (0.9, 0.9, 0.9), 9)
    sim = O().OracleSim(cfg, parts, sdf=sdf)
    got = sim.contacts(0.1)
    sel = np.abs(parts[:, 0] - 0.5) < 0.1 - 1e-5
    assert len(got) >= sel.sum() and len(got) <= (np.abs(parts[:, 0] - 0.5) < 0.1 + 1e-5).sum()
    assert np.all(np.abs(got[:, 0] - 0.5) < 0.1 + 1e-5)
    assert np.all(got[:, 3] == 0.0)
    # id order: the x positions appear in the order of the input
    ids = [int(np.nonzero(np.all(parts[:, 0:3] == g[0:3], axis=1))[0][0]) for g in got]
    assert ids == sorted(ids)


# ---------------------------------------------------------------- XSPH, curl, vorticity, LOD, interaction


def _pair(cfg, xi, xj, vi, vj):
    parts = np.zeros((2, 12), dtype=np.float32)
    parts[0, 0:3], parts[1, 0:3] = xi, xj
    parts[0, 8:11], parts[1, 8:11] = vi, vj
    parts[:, 4:7] = parts[:, 0:3]
    return parts


def test_xsph_pair_closed_form_and_momentum():
    """Two particles, no gravity, lambda ~ 0 (eps = 1e30), no s_corr: the
    positions are x* = x + v dt and the velocity is v + c/rho0 (v_j - v_i) W(r)
    (CS_Integrate.glsl:135-280), so the pair's momentum is conserved."""
    c_visc = 0.2
    cfg = dam_break_config(substeps=1, n_iter=1, gravity=(0.0, 0.0, 0.0), eps=1e30, enable_scorr=False,
                           enable_vorticity=False, viscosity=c_visc, box=(4.0, 4.0, 4.0), hash_size=1 << 20)
    vi, vj = np.array([0.3, -0.1, 0.2]), np.array([-0.2, 0.4, 0.1])
    parts = _pair(cfg, (1.0, 1.0, 1.0), (1.04, 1.01, 0.98), vi, vj)
    sim = O().OracleSim(cfg, parts, precision="f64")
    sim.step(1)
    u = sim.ubo
    x0 = parts[:, 0:3].astype(np.float64)
    v0 = parts[:, 8:11].astype(np.float64)
    xs = x0 + v0 * u.sub_dt
    r2 = float(((xs[0] - xs[1]) ** 2).sum())
    W = u.poly6_coeff * (u.h * u.h - r2) ** 3
    assert W > 0
    expect_i = v0[0] + c_visc * u.inv_rho0 * (v0[1] - v0[0]) * W
    np.testing.assert_allclose(sim.vel[0, 0:3], expect_i, rtol=0, atol=1e-9)
    np.testing.assert_allclose(sim.vel[0, 0:3] + sim.vel[1, 0:3], v0[0] + v0[1], rtol=0, atol=1e-9)
    np.testing.assert_allclose(sim.pos[:, 0:3], xs, rtol=0, atol=1e-12)


def _still_lattice(cfg, side, origin, vel_fn):
    parts = fsi().lattice(side, side, side, origin, 0.05, 0.0)
    x = parts[:, 0:3].astype(np.float64)
    parts[:, 8:11] = vel_fn(x)
    parts[:, 4:7] = parts[:, 0:3]
    return parts


def test_curl_of_rigid_rotation_is_minus_two_omega():
    """For a rigid rotation v = W x (x - c), (v_j - v_i) = -W x r and, with
    V_j = 1/rho0, the SPH identity sum_j V_j r (x) grad W ~ -(rho/rho0) I
    gives curl = -2 W rho/rho0 for an interior particle (the gradient is
    taken with r = x_i - x_j, CS_Integrate.glsl:171-193).  The lattice sums
    of the spiky gradient and of poly6 differ from their integrals by a few
    per cent (5.3% at spacing h/2); a flipped sign or cross order would give
    +2 W, a dropped 1/rho0 a factor of rho0."""
    cfg = dam_break_config(substeps=1, n_iter=1, gravity=(0.0, 0.0, 0.0), eps=1e30, enable_scorr=False,
                           box=(4.0, 4.0, 4.0), hash_size=1 << 20)
    omega = np.array([0.0, 0.0, 2.0])
    centre = np.array([1.4, 1.4, 1.4])
    parts = _still_lattice(cfg, 16, (1.0, 1.0, 1.0), lambda x: np.cross(omega, x - centre))
    sim = O().OracleSim(cfg, parts, precision="f64")
    sim.step(1)
    x = parts[:, 0:3].astype(np.float64)
    interior = (np.abs(x - centre) < 0.4 - 0.2 + 1e-9).all(1)
    assert interior.sum() > 100
    curl = sim.solver[interior, 4:7]
    rho = sim.solver[interior, 7]
    expect = -2.0 * omega[None, :] * (rho / sim.ubo.rho0)[:, None]
    np.testing.assert_allclose(curl[:, 0:2], 0.0, atol=1e-3 * 2.0)
    rel = np.abs(curl[:, 2] / expect[:, 2] - 1.0)
    assert rel.max() < 0.08, rel.max()


def test_vorticity_confinement_points_up_the_vorticity_gradient():
    """|w| = a x on a still lattice: eta = inv sum |w_j| grad W ~ (rho/rho0)
    grad|w| points along +x, so v gains eps_v dt (x^ x w) = -eps_v dt a x y^
    (CS_VorticityConfinement.glsl:133-167), up to the direction of eta (the
    FP32 lattice coordinates are symmetric to ~1e-8)."""
    cfg = dam_break_config(substeps=1, n_iter=1, gravity=(0.0, 0.0, 0.0), box=(4.0, 4.0, 4.0), hash_size=1 << 20)
    a = 3.0
    parts = _still_lattice(cfg, 16, (1.0, 1.0, 1.0), lambda x: np.zeros_like(x))
    sim = O().OracleSim(cfg, parts, precision="f64")
    sim.predict()
    sim.sort()
    sim.build_cells()
    x = parts[:, 0:3].astype(np.float64)
    sim.solver[:, 4:7] = 0.0
    sim.solver[:, 6] = a * x[:, 0]
    sim.vorticity()
    interior = (np.abs(x - 1.4) < 0.2 + 1e-9).all(1)
    dv = sim.vel[interior, 0:3]
    expect_y = -cfg.f32("vorticity_epsilon") * sim.ubo.sub_dt * a * x[interior, 0]
    np.testing.assert_allclose(dv[:, 1], expect_y, rtol=1e-6)
    assert np.abs(dv[:, [0, 2]]).max() <= 1e-3 * np.abs(expect_y).max()


def test_lod_closed_form():
    """CS_ComputeLOD.glsl:66-79: round(mix(max, min, clamp(d / maxDist, 0, 1)))."""
    cfg = dam_break_config(enable_apbf=True, min_lod=1, max_lod=7, lod_max_dist=4.0, gravity=(0.0, 0.0, 0.0))
    cam = np.array([0.5, 0.5, 0.5])
    d = np.array([0.0, 1.0, 2.0, 2.25, 3.99, 10.0])
    parts = np.zeros((len(d), 12), dtype=np.float32)
    parts[:, 0] = cam[0] + d
    parts[:, 1:3] = cam[1:3]
    parts[:, 4:7] = parts[:, 0:3]
    sim = O().OracleSim(cfg, parts, precision="f64", camera=tuple(cam))
    sim.predict()
    sim.lod_pass()
    t = np.clip(d / 4.0, 0.0, 1.0)
    m = 7.0 * (1.0 - t) + 1.0 * t
    expect = np.clip(np.floor(m + 0.5), 1, 7)  # half away from zero (m > 0)
    np.testing.assert_array_equal(sim.lod, expect.astype(np.uint32))
    assert list(sim.lod) == [7, 6, 4, 4, 1, 1]


def test_interaction_force_closed_form():
    """CS_PredictAndHash.glsl:105-112: v += normalize(p_i - x) * strength * (1 - d / r) * dt."""
    cfg = dam_break_config(substeps=1, gravity=(0.0, -9.81, 0.0), interaction_radius=1.5, interaction_strength=25.0)
    parts = np.zeros((3, 12), dtype=np.float32)
    parts[:, 0:3] = [[1.0, 1.0, 1.0], [2.0, 1.0, 1.0], [3.0, 1.0, 1.0]]
    parts[:, 4:7] = parts[:, 0:3]
    ip = np.array([1.5, 1.3, 1.0])
    sim = O().OracleSim(cfg, parts, precision="f64", interaction=tuple(ip))
    sim.predict()
    dt = sim.ubo.sub_dt
    g = np.array(cfg.gravity, dtype=np.float32).astype(np.float64)
    for i in range(3):
        p = parts[i, 0:3].astype(np.float64)
        to = ip.astype(np.float32).astype(np.float64) - p
        dist = np.linalg.norm(to)
        v = g * dt
        if 1e-5 < dist < 1.5:
            v = v + to / dist * 25.0 * (1.0 - dist / 1.5) * dt
        np.testing.assert_allclose(sim.vel[i, 0:3], v, rtol=0, atol=1e-12)
    assert np.linalg.norm(parts[2, 0:3] - ip) > 1.5  # the third is outside the radius: gravity only


# ---------------------------------------------------------------- value pins of lambda, delta-p, the two clamps
#
# Closed forms from the published method (Mueller et al. 2003 for the kernels,
# Macklin & Mueller 2013 for the constraint), evaluated here in FP64 from the
# FP32 inputs: W(r) = 315/(64 pi h^9) (h^2 - r^2)^3, grad W(d) = -45/(pi h^6)
# (h - r)^2 d/r, lambda_i = -C_i / (sum_k |grad_k C_i|^2 + eps) with
# grad_i C_i = (1/rho0) sum_j grad W_ij and grad_j C_i = -(1/rho0) grad W_ij,
# s_corr = -k (W(r)/W(dq h))^n and
# dp_i = (1/rho0) sum_j (lambda_i + lambda_j + s_corr) grad W_ij.
# Each case is chosen so that a plausible slip (a dropped 1/rho0, a missing
# |grad_i C_i|^2 term, a flipped s_corr sign, exponent n - 1, a missing clamp)
# moves the pinned value by far more than the tolerance;
# tests/test_oracle_mutants.py plants those slips in the oracle and checks
# that these tests then fail.

PRECISIONS = [("f64", 1e-12), ("f32", 3e-5)]


def _f(x):
    return float(np.float32(x))


def _w_poly6(r, h):
    return 315.0 / (64.0 * math.pi * h ** 9) * (h * h - r * r) ** 3 if r <= h else 0.0


def _grad_spiky(d, h):
    r = float(np.linalg.norm(d))
    return -45.0 / (math.pi * h ** 6) * (h - r) ** 2 * d / r


def _pair_case(prec, d, **kw):
    """Two particles at rest, no gravity, x1 = x0 - d; returns the oracle after
    predict, sort and cell tables (x* = x) and the FP64 inputs."""
    kw.setdefault("gravity", (0.0, 0.0, 0.0))
    cfg = dam_break_config(substeps=1, n_iter=1, box=(4.0, 4.0, 4.0), hash_size=1 << 20, **kw)
    parts = np.zeros((2, 12), dtype=np.float32)
    parts[0, 0:3] = [1.5, 1.5, 1.5]
    parts[1, 0:3] = parts[0, 0:3] - np.asarray(d, dtype=np.float32)
    parts[:, 4:7] = parts[:, 0:3]
    sim = O().OracleSim(cfg, parts, precision=prec)
    sim.predict()
    sim.sort()
    sim.build_cells()
    x = parts[:, 0:3].astype(np.float64)
    return cfg, sim, x


def _pair_lambda(cfg, x):
    h, rho0, eps = _f(cfg.h), _f(cfg.rho0), _f(cfg.eps)
    d = x[0] - x[1]
    r = float(np.linalg.norm(d))
    rho = _w_poly6(0.0, h) + _w_poly6(r, h)
    g = _grad_spiky(d, h)
    C_ = max(rho / rho0 - 1.0, -_f(cfg.cohesion_strength))
    sum_j = float(g @ g) / rho0 ** 2          # sum_j |grad_j C_i|^2 (one neighbour)
    grad_i = g / rho0                         # grad_i C_i
    lam = -C_ / (sum_j + float(grad_i @ grad_i) + eps)
    return rho, lam, g, d, r


def _rel(got, exp):
    got, exp = np.asarray(got, dtype=np.float64), np.asarray(exp, dtype=np.float64)
    return float(np.linalg.norm(got - exp) / np.linalg.norm(exp))


@pytest.mark.parametrize("prec,tol", PRECISIONS)
@pytest.mark.parametrize("rho0,eps,coh", [(1000.0, 100.0, 0.001), (6378.0, 1.0, 0.001), (6378.0, 600.0, 0.2)])
def test_lambda_of_a_pair_closed_form(prec, tol, rho0, eps, coh):
    """CS_ComputeLambdas.glsl:136-158: rho = W(0) + W(r) and
    lambda = -max(rho/rho0 - 1, -cohesion) / (2 |grad W|^2 / rho0^2 + eps).
    The first case is compressed (C > 0), the others sit on the cohesion floor."""
    cfg, sim, x = _pair_case(prec, (0.031, -0.017, 0.009), rho0=rho0, eps=eps, cohesion_strength=coh)
    sim.lambda_pass(1)
    rho, lam, g, _, _ = _pair_lambda(cfg, x)
    # the case exercises both denominator terms against eps
    assert float(g @ g) / rho0 ** 2 > 0.5 * eps or rho0 == 6378.0 and eps == 600.0
    for i in (0, 1):
        assert float(sim.solver[i, 7]) == pytest.approx(rho, rel=tol)
        assert float(sim.solver[i, 3]) == pytest.approx(lam, rel=tol)


@pytest.mark.parametrize("prec,tol", PRECISIONS)
@pytest.mark.parametrize("k,n,dq,scorr", [(1e-4, 4.0, 0.3, True), (2e-4, 2.5, 0.5, True), (1e-4, 4.0, 0.3, False)])
def test_position_correction_of_a_pair_closed_form(prec, tol, k, n, dq, scorr):
    """CS_ComputeDeltaP.glsl:94-104,150-165: dp_0 = (1/rho0) (2 lambda + s) grad W
    with s = -k (W(r)/W(dq h))^n (integer and fractional n), and dp_1 = -dp_0."""
    cfg, sim, x = _pair_case(prec, (0.012, -0.009, 0.011), rho0=1000.0, eps=100.0, k_corr=k, n_corr=n,
                             delta_q=dq, enable_scorr=scorr)
    sim.lambda_pass(1)
    sim.deltap_pass(1)
    h, rho0 = _f(cfg.h), _f(cfg.rho0)
    _, lam, g, _, r = _pair_lambda(cfg, x)
    dqh = _f(dq) * h  # FP64 product of the FP32 inputs; its FP32 rounding is below the f32 tolerance
    s = -_f(k) * (_w_poly6(r, h) / _w_poly6(dqh, h)) ** _f(n) if scorr else 0.0
    if scorr:
        assert abs(s) > 0.5 * abs(2 * lam) and _w_poly6(r, h) > 1.1 * _w_poly6(dqh, h)
    dp0 = (2.0 * lam + s) * g / rho0
    assert _rel(sim.solver[0, 4:7], dp0) <= tol
    assert _rel(sim.solver[1, 4:7], -dp0) <= tol


@pytest.mark.parametrize("prec,tol", PRECISIONS)
@pytest.mark.parametrize("eps", [1e-3, 1e5])
def test_position_correction_is_clamped_to_half_h(prec, tol, eps):
    """CS_ApplyDeltaP.glsl:92-104: |dp| > h/2 is scaled back to h/2 along dp;
    below it x + dp is applied as is.  eps = 1e-3 gives |dp| ~ 0.4 = 8 h/2,
    eps = 1e5 gives |dp| ~ 0.01 < h/2."""
    cfg, sim, x = _pair_case(prec, (0.06, 0.05, 0.03), rho0=100.0, eps=eps, enable_scorr=False)
    sim.lambda_pass(1)
    sim.deltap_pass(1)
    h, rho0 = _f(cfg.h), _f(cfg.rho0)
    _, lam, g, _, _ = _pair_lambda(cfg, x)
    dp0 = 2.0 * lam * g / rho0
    assert _rel(sim.solver[0, 4:7], dp0) <= tol
    big = np.linalg.norm(dp0) > 0.5 * h
    assert big == (eps < 1.0)
    step = dp0 * (0.5 * h / np.linalg.norm(dp0)) if big else dp0
    sim.apply_pass(1)
    for i, sgn in ((0, 1.0), (1, -1.0)):
        moved = sim.solver[i, 0:3].astype(np.float64) - x[i]
        assert _rel(moved, sgn * step) <= max(tol, 2e-5 if prec == "f32" else tol)


@pytest.mark.parametrize("prec,tol", [("f64", 1e-9), ("f32", 1e-4)])
@pytest.mark.parametrize("v0", [(60.0, 30.0, -20.0), (40.0, 20.0, -10.0), (0.0, -51.0, 0.0)])
def test_speed_is_clamped_to_50(prec, tol, v0):
    """CS_Integrate.glsl:244-256: v = (x* - x)/dt, and a speed above 50 is
    scaled back to 50 along v (72.8 and 51 m/s); 45.8 m/s passes unchanged."""
    cfg = dam_break_config(substeps=1, n_iter=1, box=(100.0, 100.0, 100.0), gravity=(0.0, 0.0, 0.0),
                           hash_size=1 << 20)
    parts = np.zeros((1, 12), dtype=np.float32)
    parts[0, 0:3] = [1.0, 2.0, 1.0]
    parts[0, 4:7] = parts[0, 0:3]
    parts[0, 8:11] = v0
    sim = O().OracleSim(cfg, parts, precision=prec)
    sim.step(1)
    v = np.asarray(v0, dtype=np.float64)
    speed = float(np.linalg.norm(v))
    expect = v * (50.0 / speed) if speed > 50.0 else v
    assert _rel(sim.vel[0, 0:3], expect) <= tol
    assert _rel(sim.pos[0, 0:3], parts[0, 0:3] + v * _f(cfg.dt)) <= tol  # x* = x + v dt, dt the FP32 1/120
