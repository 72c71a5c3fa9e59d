"""The value pins of tests/test_oracle_pins.py bite: each case below plants one
plausible slip in a copy of the oracle's sources (a dropped 1/rho0, a missing
constraint-gradient term, a flipped s_corr sign, exponent n - 1, a missing
clamp, the cohesion floor taken as a minimum), builds that copy with the
oracle's own flags, and runs the pins against it through FSI_ORACLE_LIB.  The
pins must fail on every mutant and pass on an unmutated copy built the same
way.  CPU only (gcc)."""
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
CFLAGS = ["-O2", "-std=c11", "-fPIC", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-march=x86-64-v2",
          "-w"]

LAMBDA = "test_lambda_of_a_pair_closed_form"
DELTAP = "test_position_correction_of_a_pair_closed_form"
CLAMP = "test_position_correction_is_clamped_to_half_h"
SPEED = "test_speed_is_clamped_to_50"

# (name, text in pbf_oracle_body.h, replacement, pins that must fail)
MUTANTS = [
    ("lambda_drops_inv_rho0_squared", "sum_g2 = sum_g2 + gj2 * (inv * inv);", "sum_g2 = sum_g2 + gj2;", LAMBDA),
    ("lambda_drops_grad_i_term", "sum_g2 = sum_g2 + FN(dot)(gi, gi);", "sum_g2 = sum_g2;", LAMBDA),
    ("lambda_cohesion_floor_as_min", "C = C > floorc ? C : floorc;", "C = C < floorc ? C : floorc;", LAMBDA),
    ("deltap_drops_inv_rho0", "dp = FN(scl)(dp, inv);", "dp = dp;", DELTAP),
    ("scorr_sign_flipped", "return -(REAL)u->k_corr * FN(powr)(ratio, (REAL)u->n_corr);",
     "return (REAL)u->k_corr * FN(powr)(ratio, (REAL)u->n_corr);", DELTAP),
    ("scorr_exponent_n_minus_1", "return -(REAL)u->k_corr * FN(powr)(ratio, (REAL)u->n_corr);",
     "return -(REAL)u->k_corr * FN(powr)(ratio, (REAL)u->n_corr - (REAL)1);", DELTAP),
    ("deltap_clamp_removed", "if (l2 > max2) {", "if (0) {", CLAMP),
    ("speed_clamp_removed", "if (sp2 > RC(2500.0))", "if (0)", SPEED),
]


def _build(tmp, mutation):
    src = os.path.join(tmp, "oracle")
    inc = os.path.join(tmp, "include")
    os.makedirs(src)
    os.makedirs(inc)
    for f in ("pbf_oracle.c", "pbf_oracle.h", "pbf_oracle_body.h", "mesh_oracle.c", "mesh_oracle.h"):
        shutil.copy(os.path.join(ORACLE, f), src)
    shutil.copy(os.path.join(ROOT, "include", "fsi_pbf.h"), inc)
    if mutation is not None:
        old, new = mutation
        body = os.path.join(src, "pbf_oracle_body.h")
        text = open(body).read()
        assert text.count(old) == 1, f"mutation site not found once: {old!r}"
        open(body, "w").write(text.replace(old, new))
    lib = os.path.join(src, "liboracle.so")
    subprocess.run(["gcc", *CFLAGS, "-shared", "-o", lib, os.path.join(src, "pbf_oracle.c"),
                    os.path.join(src, "mesh_oracle.c"), "-lm"], check=True)
    return lib


def _run_pins(lib, selector):
    env = dict(os.environ, FSI_ORACLE_LIB=lib)
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle_pins.py"), "-k", selector],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    return r.returncode, r.stdout + r.stderr


def test_unmutated_copy_passes_the_pins(tmp_path):
    lib = _build(str(tmp_path), None)
    rc, out = _run_pins(lib, " or ".join((LAMBDA, DELTAP, CLAMP, SPEED)))
    assert rc == 0, out[-2000:]


@pytest.mark.parametrize("name,old,new,pins", MUTANTS, ids=[m[0] for m in MUTANTS])
def test_pins_fail_on_mutant(tmp_path, name, old, new, pins):
    lib = _build(str(tmp_path), (old, new))
    rc, out = _run_pins(lib, pins)
    # rc 1: tests ran and some failed (not a collection or usage error)
    assert rc == 1, f"{name}: pins did not fail (rc {rc})\n{out[-2000:]}"
    assert " failed" in out and "AssertionError" in out, out[-2000:]
