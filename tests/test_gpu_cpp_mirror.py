"""The C++ PBFluids / SDFBoundary mirror (realtime-fsi-engine_amd/host/), driven
the way the reference's src/main.cpp drives its PBFluids and SDFBoundary
(tests/cpp/pbfluids_driver.cpp, built with the package), gives the same
particles and contact candidates, bit for bit, as the same scene run through
the ctypes binding of the C-ABI: first on the static obstacle, then over
coupling frames that dent it (refitDirty / uploadDirtyRegion, refitFromMesh /
uploadToGPU), with closestTriangle of every final candidate checked against
the device query and the brute-force oracle."""
import os
import subprocess

import numpy as np
import pytest

from conftest import dam_break_config

pytestmark = pytest.mark.gpu

FRAMES = 6
COUPLE = 6


def test_cpp_mirror_matches_the_c_abi(gpu, tmp_path):
    import fsi_pbf as fsi

    exe = os.path.join(os.path.dirname(fsi.LIB_PATH), "pbfluids_driver")
    assert os.path.exists(exe), "build the package first (make -C realtime-fsi-engine_amd)"
    print(f"running {exe} for {FRAMES} + {COUPLE} frames", flush=True)
    r = subprocess.run([exe, str(tmp_path), str(FRAMES), str(COUPLE)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(tmp_path / "particles.bin", dtype=np.float32).reshape(-1, 12)
    got_c = np.fromfile(tmp_path / "contacts.bin", dtype=np.float32).reshape(-1, 8)

    # the same scene: a 16 x 12 x 12 block at 0.025 + 0.05 i (FP32, as the
    # driver computes it), an icosphere obstacle, 3 iterations x 2 substeps
    cfg = dam_break_config(n_iter=3, substeps=2, box=(2.4, 1.6, 1.6), hash_size=1 << 14)
    f32 = np.float32
    parts = np.zeros((16 * 12 * 12, 12), dtype=np.float32)
    i = 0
    for z in range(12):
        for y in range(12):
            for x in range(16):
                p = [f32(0.025) + f32(0.05) * f32(x), f32(0.025) + f32(0.05) * f32(y), f32(0.025) + f32(0.05) * f32(z)]
                parts[i, 0:3] = p
                parts[i, 4:7] = p
                i += 1
    centre = (0.9, 0.2, 0.3)
    verts, faces = fsi.icosphere(2, centre, 0.25)
    sdf = fsi.sdf_build(verts, faces, 0.05)  # the host build: the mirror's device build must equal it
    mesh = fsi.Mesh(verts, faces, 0.05)
    with fsi.Solver(cfg) as s:
        s.set_collision_padding(0.0)
        s.set_bounds((0.0, 0.0, 0.0), (2.4, 1.6, 1.6), 0.5)
        s.set_sdf(sdf, 0.025, False)
        s.set_particles(parts)
        s.step(FRAMES)
        s.filter_contacts(0.1)
        exp = s.read_particles()
        exp_c = s.read_contacts()
        assert got.shape == exp.shape and np.array_equal(got, exp)
        assert len(exp_c) > 0, "the scene should touch the obstacle"
        assert got_c.shape == exp_c.shape and np.array_equal(got_c, exp_c)

        # the coupling frames through the C-ABI
        dirty_rows = []
        cur = verts.copy()
        up = verts[:, 1] > centre[1]
        for c in range(COUPLE):
            cur[up, 1] = verts[up, 1] - 0.01 * (c + 1) * (verts[up, 1] - centre[1])
            if c % 3 == 2:
                mesh.refit(cur)
                dirty_rows.append((0, 0, 0, 0, -1, -1, -1))
            else:
                n, lo, hi = mesh.refit_dirty(cur, 0.2, 1e-4)
                dirty_rows.append((1, *lo, *hi) if n > 0 else (0, 0, 0, 0, -1, -1, -1))
            s.set_sdf_from_mesh(mesh, 0.025, False)
            s.step(1)
            s.filter_contacts(0.1)
        exp2 = s.read_particles()
        exp2_c = s.read_contacts()
    grid = mesh.read_sdf()
    got2 = np.fromfile(tmp_path / "coupled_particles.bin", dtype=np.float32).reshape(-1, 12)
    got2_c = np.fromfile(tmp_path / "coupled_contacts.bin", dtype=np.float32).reshape(-1, 8)
    assert np.array_equal(got2, exp2)
    assert len(exp2_c) > 0 and np.array_equal(got2_c, exp2_c)
    rows = [tuple(int(t) for t in line.split()) for line in (tmp_path / "dirty.txt").read_text().splitlines()]
    assert rows == dirty_rows
    assert sum(r[0] for r in rows) > 0, "the dent should mark voxels"
    # distances(): x * Ny * Nz + y * Nz + z, the device grid transposed
    dims = grid.dims
    host = np.fromfile(tmp_path / "grid.bin", dtype=np.float32).reshape(dims[0], dims[1], dims[2])
    assert np.array_equal(host, np.asarray(grid.grid).reshape(dims[2], dims[1], dims[0]).transpose(2, 1, 0))

    # closestTriangle of every final candidate: the device query, and the
    # oracle's brute force over the dented triangles
    rec = np.fromfile(tmp_path / "closest.bin", dtype=np.dtype([("face", "<i4"), ("p", "<f4", 3)]))
    assert len(rec) == len(exp2_c)
    dev_face, dev_p = mesh.closest(exp2_c[:, 0:3])
    assert np.array_equal(rec["face"], dev_face) and np.array_equal(rec["p"], dev_p)
    import pyoracle

    of, op, od2 = pyoracle.closest_triangles(cur, faces, exp2_c[:, 0:3])
    q = exp2_c[:, 0:3].astype(np.float64)
    assert np.abs(rec["p"] - op).max() <= 1e-6
    d2 = np.sum((q - rec["p"].astype(np.float64)) ** 2, axis=1)
    assert np.abs(d2 - od2).max() <= 1e-6
    # a different face only where it is as close as the oracle's (a shared edge or vertex)
    for i in np.nonzero(rec["face"] != of)[0]:
        _, _, d2f = pyoracle.closest_triangles(cur, faces[rec["face"][i]][None], exp2_c[i, 0:3][None])
        assert abs(np.sqrt(d2f[0]) - np.sqrt(od2[i])) <= 1e-12, (i, rec["face"][i], of[i])
    mesh.close()
