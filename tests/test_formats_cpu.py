"""The reference's host formats on the CPU: spawnParticles (src/main.cpp:160-188),
loadOFF (src/main.cpp:364-415) and fitMeshToDomain (src/main.cpp:419-432).

The expected values are restated here independently: the grid spawn as the
same FP32 loop in numpy, the random spawn from numpy's legacy MT19937 (the
engine std::mt19937 specifies, pinned by the standard's 10000th-output value)
mapped to floats as libstdc++'s generate_canonical<float, 24> does, the OFF
files by hand."""
import numpy as np
import pytest

import fsi_pbf


def test_numpy_legacy_mt19937_is_the_standard_engine():
    # [rand.predef]: the 10000th invocation of a default-constructed
    # std::mt19937 (seed 5489) produces 4123659995
    draws = np.random.RandomState(5489).randint(0, 2 ** 32, size=10000, dtype=np.uint64)
    assert int(draws[-1]) == 4123659995


def grid_spawn(cfg):
    """The grid branch of spawnParticles, FP32 accumulation as the C++ loop."""
    f32 = np.float32
    dx = max(f32(1e-6), f32(cfg.spacing))
    lo = [f32(v) for v in cfg.spawn_min]
    hi = [f32(v) for v in cfg.spawn_max]
    out = []
    z = lo[2]
    while z <= hi[2] and len(out) < cfg.particle_count:
        y = lo[1]
        while y <= hi[1] and len(out) < cfg.particle_count:
            x = lo[0]
            while x <= hi[0] and len(out) < cfg.particle_count:
                out.append((x, y, z))
                x = f32(x + dx)
            y = f32(y + dx)
        z = f32(z + dx)
    return np.array(out, dtype=np.float32)


@pytest.mark.parametrize("count", [4096, 20, 1 << 20])
def test_grid_spawn(count):
    cfg = fsi_pbf.FluidConfig(particle_count=count, spacing=0.1, spawn_min=(2.0, 0.2, 2.0), spawn_max=(5.0, 5.0, 5.0),
                              initial_velocity=(0.5, -1.0, 0.25))
    got = fsi_pbf.spawn_particles(cfg)
    want = grid_spawn(cfg)
    assert got.shape == (len(want), 12)
    assert len(want) == min(count, 31 * 49 * 31)  # 2..5 step 0.1 in FP32 (31), 0.2..5 (49)
    assert np.array_equal(got[:, 0:3], want) and np.array_equal(got[:, 4:7], want)
    assert np.array_equal(got[:, 8:11], np.tile(np.float32([0.5, -1.0, 0.25]), (len(want), 1)))
    assert not got[:, [3, 7, 11]].any()


def test_random_spawn():
    f32 = np.float32
    cfg = fsi_pbf.FluidConfig(particle_count=1000, spacing=0.1, spawn_random=True, spawn_min=(2.0, 0.2, 2.0),
                              spawn_max=(5.0, 5.0, 5.0))
    got = fsi_pbf.spawn_particles(cfg)
    draws = np.random.RandomState(1337).randint(0, 2 ** 32, size=3 * 1000, dtype=np.uint64)
    u = draws.astype(np.float32) / f32(2.0 ** 32)  # generate_canonical<float, 24>: one draw, FP32 quotient
    u = np.where(u >= f32(1.0), np.nextafter(f32(1.0), f32(0.0)), u).reshape(1000, 3)
    lo = np.float32([2.0, 0.2, 2.0]) + f32(0.1)
    hi = np.float32([5.0, 5.0, 5.0]) - f32(0.1)
    want = (lo + (hi - lo) * u).astype(np.float32)
    assert np.array_equal(got[:, 0:3], want)
    assert np.array_equal(got[:, 4:7], want)
    assert (got[:, 0:3] >= lo).all() and (got[:, 0:3] <= hi).all()


OFF_TEXT = """# exported by a tool
# second comment

OFF  \r
# counts follow
5 3 0
0 0 0  0 0 1  255 0 0
1 0 0  0 0 1
# a comment between vertices
1 1 0
0 1 0
0.5 0.5 1.25
4 0 1 2 3
3 0 1 4
5 0 1 2 3 4
"""


def test_off_loader_fan_triangulates_and_skips_comments(tmp_path):
    path = tmp_path / "mesh.off"
    path.write_text(OFF_TEXT)
    verts, faces = fsi_pbf.load_off(str(path))
    assert np.array_equal(verts, np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0.5, 0.5, 1.25]], dtype=float))
    assert faces.tolist() == [[0, 1, 2], [0, 2, 3], [0, 1, 4], [0, 1, 2], [0, 2, 3], [0, 3, 4]]


def test_off_loader_short_file_and_errors(tmp_path):
    short = tmp_path / "short.off"
    short.write_text("OFF\n3 2 0\n1 2 3\n4 5 6\n")  # one vertex and both faces missing
    verts, faces = fsi_pbf.load_off(str(short))
    assert np.array_equal(verts, np.array([[1, 2, 3], [4, 5, 6], [0, 0, 0]], dtype=float))
    assert faces.shape == (0, 3)
    bad = tmp_path / "bad.off"
    bad.write_text("PLY\n3 1 0\n")
    with pytest.raises(fsi_pbf.FsiError, match="not an OFF file"):
        fsi_pbf.load_off(str(bad))
    with pytest.raises(fsi_pbf.FsiError, match="cannot open"):
        fsi_pbf.load_off(str(tmp_path / "missing.off"))


def test_fit_mesh_to_domain():
    rng = np.random.default_rng(3)
    verts = rng.normal(size=(200, 3)) * [3.0, 1.0, 0.5] + [10.0, -4.0, 2.0]
    out = fsi_pbf.fit_mesh_to_domain(verts, (1.0, 2.0, 3.0), 1.5)
    lo, hi = out.min(axis=0), out.max(axis=0)
    assert np.allclose(0.5 * (lo + hi), [1.0, 2.0, 3.0], atol=1e-12)
    assert np.isclose((hi - lo).max(), 1.5, rtol=1e-12)
    # a uniform scale: ratios of extents are kept
    assert np.allclose((hi - lo) / (hi - lo).max(), np.ptp(verts, axis=0) / np.ptp(verts, axis=0).max())
    # a single point has extent 1e-9 and lands on the centre
    one = fsi_pbf.fit_mesh_to_domain(np.array([[5.0, 5.0, 5.0]]), (1.0, 2.0, 3.0), 2.0)
    assert np.array_equal(one, np.array([[1.0, 2.0, 3.0]], dtype=float))
