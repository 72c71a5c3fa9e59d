"""CPU-side checks of the product library: it loads, exports every entry point
declared in include/fsi_pbf.h, and its host-only helpers (defaults, lattice
spawn, icosphere, FP64 SDF build, config validation) behave."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from conftest import PKG, ROOT


def declared_functions():
    text = open(os.path.join(ROOT, "include", "fsi_pbf.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(fsi_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    import fsi_pbf

    lib = fsi_pbf.load()
    names = declared_functions()
    assert len(names) >= 30
    out = subprocess.run(["nm", "-D", "--defined-only", fsi_pbf.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (fsi_\w+)", out))
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    for n in names:
        assert getattr(lib, n) is not None
    assert lib.fsi_abi_version() == 1


def test_struct_sizes_match_reference_layouts():
    import fsi_pbf

    # FluidConfigUBO is 144 bytes (include/fluid/PBFluids.h:21-61)
    assert C.sizeof(fsi_pbf.CConstants) == 144
    cfg_fields = [f[0] for f in fsi_pbf.CFluidConfig._fields_]
    # FluidConfig field order (include/core/config.h:9-63)
    assert cfg_fields[:8] == ["dt", "gravity", "h", "rho0", "solver_iterations", "substep_iterations", "eps",
                              "particle_count"]
    assert cfg_fields[-4:] == ["enable_apbf", "min_lod", "max_lod", "lod_max_dist"]


def test_defaults_are_the_reference_fluid_config():
    import fsi_pbf

    c = fsi_pbf.CFluidConfig()
    assert fsi_pbf.load().fsi_fluid_config_default(C.byref(c)) == 0
    d = fsi_pbf.FluidConfig().to_c()
    for name, _ in fsi_pbf.CFluidConfig._fields_:
        a, b = getattr(c, name), getattr(d, name)
        if hasattr(a, "__len__"):
            assert list(a) == list(b), name
        else:
            assert a == b, name
    # spot values of include/core/config.h
    assert c.hash_size == 1 << 17 and c.solver_iterations == 3 and c.substep_iterations == 2
    assert np.float32(c.gravity[1]) == np.float32(-3.81) and np.float32(c.h) == np.float32(0.15)


def test_create_rejects_bad_configs_before_touching_a_device():
    import fsi_pbf

    lib = fsi_pbf.load()
    h = C.c_void_p()
    for bad in (dict(hash_size=1000), dict(hash_size=0), dict(substep_iterations=0), dict(h=0.0), dict(rho0=0.0)):
        cfg = fsi_pbf.FluidConfig(**bad).to_c()
        rc = lib.fsi_pbf_create(C.byref(cfg), C.byref(h))
        assert rc == -1, (bad, rc)
        assert lib.fsi_last_error().decode()


def test_lattice_spawn_layout_and_determinism():
    import fsi_pbf

    a = fsi_pbf.lattice(4, 3, 2, (1.0, 2.0, 3.0), 0.05, 0.001, 1337)
    b = fsi_pbf.lattice(4, 3, 2, (1.0, 2.0, 3.0), 0.05, 0.001, 1337)
    c = fsi_pbf.lattice(4, 3, 2, (1.0, 2.0, 3.0), 0.05, 0.001, 1338)
    assert a.shape == (24, 12) and np.array_equal(a, b) and not np.array_equal(a, c)
    base = fsi_pbf.lattice(4, 3, 2, (1.0, 2.0, 3.0), 0.05, 0.0, 1337)
    # ids x-fastest, then y, then z, at origin + (i + 0.5) * spacing
    i, j, k = 1, 2, 1
    idx = i + 4 * (j + 3 * k)
    np.testing.assert_allclose(base[idx, 0:3], [1.0 + 1.5 * 0.05, 2.0 + 2.5 * 0.05, 3.0 + 1.5 * 0.05], rtol=1e-6)
    off = a[:, 0:3] - base[:, 0:3]
    assert np.abs(off).max() <= 0.001 + 1e-6 and np.abs(off).max() > 0.0005
    assert np.array_equal(a[:, 4:7], a[:, 0:3])  # predPos = pos
    assert np.all(a[:, 8:12] == 0.0)


def test_mt19937_known_answer():
    """The lattice jitter comes from MT19937; its 10,000th output for the
    default seed 5489 is 4123659995 (the standard's conformance value).  With
    jitter = 1 and spacing = 0 the x offset of particle 3333 is (2u - 1) with
    u = (draw_10000 >> 8) * 2^-24 (three draws per particle, x first)."""
    import fsi_pbf

    a = fsi_pbf.lattice(3334, 1, 1, (0.0, 0.0, 0.0), 0.0, 1.0, 5489)
    u = np.float32((4123659995 >> 8) * (1.0 / 16777216.0))
    assert a[3333, 0] == np.float32(2.0 * u - 1.0)


def test_icosphere_counts_radius_orientation():
    import fsi_pbf

    for s in (0, 1, 3, 5):
        v, f = fsi_pbf.icosphere(s, (6.4, 3.2, 6.4), 1.5)
        assert v.shape == (10 * 4 ** s + 2, 3) and f.shape == (20 * 4 ** s, 3)
        np.testing.assert_allclose(np.linalg.norm(v - [6.4, 3.2, 6.4], axis=1), 1.5, rtol=1e-12)
        a, b, c = v[f[:, 0]], v[f[:, 1]], v[f[:, 2]]
        n = np.cross(b - a, c - a)
        cen = (a + b + c) / 3 - [6.4, 3.2, 6.4]
        assert np.all((n * cen).sum(1) > 0), "faces wind outward"
        # closed 2-manifold: every edge shared by exactly two faces
        e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), axis=1)
        _, cnt = np.unique(e, axis=0, return_counts=True)
        assert np.all(cnt == 2)


def test_sdf_build_matches_analytic_sphere():
    """SDF of the subdivision-5 icosphere (r=1.5) against |x - c| - r: the
    inscribed polyhedron deviates from the sphere by at most r(1 - cos a) with
    a the largest edge's half-angle (~2.5e-4 here); the sign is negative inside."""
    import fsi_pbf

    c = np.array([6.4, 3.2, 6.4])
    v, f = fsi_pbf.icosphere(5, c, 1.5)
    sdf = fsi_pbf.sdf_build(v, f, 0.05)
    nx, ny, nz = sdf.dims
    assert (nx, ny, nz) == (70, 70, 70)
    np.testing.assert_allclose(sdf.origin, c - 1.5 - 0.25, atol=1e-12)
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    p = np.stack([x, y, z], -1) * 0.05 + sdf.origin
    exact = np.linalg.norm(p - c, axis=-1) - 1.5
    err = np.abs(sdf.grid - exact)
    assert err.max() < 1.5e-3, err.max()
    assert sdf.grid[nz // 2, ny // 2, nx // 2] < -1.4
    assert np.all(np.sign(sdf.grid[np.abs(exact) > 0.01]) == np.sign(exact[np.abs(exact) > 0.01]))


def test_cpp_mirror_class_compiles():
    """The C++ PBFluids / SDFBoundary / LaplacianDeformation mirrors build against the header."""
    for name in ("PBFluids.cpp", "SDFBoundary.cpp", "LaplacianDeformation.cpp"):
        src = os.path.join(PKG, "host", name)
        subprocess.run(["g++", "-std=c++17", "-fsyntax-only", "-I", os.path.join(ROOT, "include"), src], check=True)


def test_deform_params_layout_matches_header(tmp_path):
    """The ctypes mirror of fsi_deform_params has the C compiler's offsets and size."""
    import fsi_pbf

    names = [f[0] for f in fsi_pbf.CDeformParams._fields_]
    body = "\n".join(f'    printf("%zu\\n", offsetof(fsi_deform_params, {n}));' for n in names)
    src = tmp_path / "layout.c"
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include \"fsi_pbf.h\"\nint main(void) {\n" + body +
                   '\n    printf("%zu\\n", sizeof(fsi_deform_params));\n    return 0;\n}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [getattr(fsi_pbf.CDeformParams, n).offset for n in names] + [C.sizeof(fsi_pbf.CDeformParams)]
    assert got == want


def test_scene_desc_layout_matches_header(tmp_path):
    """The ctypes mirror of fsi_scene_desc has the C compiler's offsets and size."""
    import fsi_pbf

    names = [f[0] for f in fsi_pbf.CSceneDesc._fields_]
    body = "\n".join(f'    printf("%zu\\n", offsetof(fsi_scene_desc, {n}));' for n in names)
    src = tmp_path / "layout.c"
    src.write_text("#include <stddef.h>\n#include <stdio.h>\n#include \"fsi_pbf.h\"\nint main(void) {\n" + body +
                   '\n    printf("%zu\\n", sizeof(fsi_scene_desc));\n    return 0;\n}\n')
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    got = [int(x) for x in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split()]
    want = [getattr(fsi_pbf.CSceneDesc, n).offset for n in names] + [C.sizeof(fsi_pbf.CSceneDesc)]
    assert got == want
