"""The oracle's two builds (oracle/Makefile: liboracle.so for the tests,
liboracle_fast.so, -O3 for AVX2, for the CPU baseline of bench.py) compute
the same bits: same FP policy (no contraction, no fast-math)."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import sys, numpy as np
sys.path[:0] = [{tests!r}, {oracle!r}]
from conftest import dam_break_config
import fsi_pbf, pyoracle
cfg = dam_break_config(n_iter=3, substeps=2)
parts = fsi_pbf.lattice(16, 16, 16, (0.0, 0.0, 0.0), 0.05, 0.001, 5)
verts, faces = fsi_pbf.icosphere(2, (1.2, 0.4, 0.4), 0.3)
sdf = fsi_pbf.sdf_build(verts, faces, 0.05)
sim = pyoracle.OracleSim(cfg, parts, sdf=sdf, sdf_padding=0.025)
sim.step(3)
np.save(sys.argv[1], sim.particles())
"""


def _run(lib, out):
    env = dict(os.environ, FSI_ORACLE_LIB=lib, OMP_NUM_THREADS="4")
    code = SCRIPT.format(tests=os.path.join(ROOT, "tests"), oracle=os.path.join(ROOT, "oracle"))
    subprocess.run([sys.executable, "-c", code, out], check=True, env=env, timeout=300)
    return np.load(out)


def test_timing_build_gives_the_same_bits(tmp_path):
    build = os.path.join(ROOT, "oracle", "_build")
    a = _run(os.path.join(build, "liboracle.so"), str(tmp_path / "a.npy"))
    b = _run(os.path.join(build, "liboracle_fast.so"), str(tmp_path / "b.npy"))
    assert np.isfinite(a).all() and np.abs(a[:, 0:3] - a[0, 0:3]).max() > 0
    assert np.array_equal(a, b)
