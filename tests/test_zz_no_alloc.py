"""fsi_mesh_contact_forces_noalloc, the contact forces without allocation,
through the C-ABI: on the same candidates it returns the bits of
fsi_pbf_contact_forces, on two calls in a row (the second with fewer
candidates, so the shared scratch holds the first call's contents), and it
rejects more candidates than its limit before touching the device."""
import ctypes as C

import numpy as np
import pytest

from conftest import dam_break_config

torch = pytest.importorskip("torch")  # before the library is loaded: one HIP runtime

pytestmark = pytest.mark.gpu

NOALLOC_MAX_CONTACTS = 65536
ERR_INVALID_ARG = -1  # FSI_ERR_INVALID_ARG (include/fsi_pbf.h)


def _scene():
    import fsi_pbf

    centre, radius = (2.0, 2.0, 2.0), 0.6
    verts, faces = fsi_pbf.icosphere(3, centre, radius)
    cfg = dam_break_config(n_iter=2, substeps=1, box=(4.0, 4.0, 4.0), hash_size=1 << 14)
    n_side = 24
    spacing = 2.2 * radius / n_side
    lo = np.asarray(centre) - 1.1 * radius
    parts = fsi_pbf.lattice(n_side, n_side, n_side, lo, spacing, 0.1 * spacing, 7)
    rng = np.random.default_rng(7)
    parts[:, 8:11] = rng.normal(0.0, 0.5, size=(len(parts), 3)).astype(np.float32)
    return verts, faces, cfg, parts


def test_noalloc_contact_forces_equal_the_allocating_entry_point(gpu):
    import fsi_pbf

    verts, faces, cfg, parts = _scene()
    lib = fsi_pbf.load()
    dev = torch.device("cuda:0")
    # torch's current stream (its handle 0, the null stream, reads as NULL:
    # the mesh's own stream); the device synchronisations below order both
    stream = torch.cuda.current_stream(dev)
    counts_seen = []
    with fsi_pbf.Mesh(verts, faces, 0.05) as m, fsi_pbf.Solver(cfg) as s:
        s.set_particles(parts)
        s.set_sdf_from_mesh(m)
        for band in (0.08, 0.03):  # the second call has fewer candidates
            s.filter_contacts(band)
            cand = s.read_contacts()
            ref_forces, ref_counts = s.contact_forces(m, band, True, 0.05)
            cd = torch.tensor(cand, device=dev)
            out = torch.full((len(verts), 4), float("nan"), device=dev)  # every entry must be written
            cnt = torch.full((2,), -1, dtype=torch.int32, device=dev)
            torch.cuda.synchronize()
            print(f"band {band}: {len(cand)} candidates", flush=True)
            rc = lib.fsi_mesh_contact_forces_noalloc(m.h, C.c_void_p(cd.data_ptr()), len(cand), float(band), 1, 0.05,
                                                     C.c_void_p(out.data_ptr()), C.c_void_p(cnt.data_ptr()),
                                                     C.c_void_p(stream.cuda_stream))
            assert rc == 0, lib.fsi_last_error().decode()
            torch.cuda.synchronize()
            got = out.cpu().numpy()
            got_counts = tuple(int(v) for v in cnt.cpu().numpy().view(np.uint32))
            assert got_counts == ref_counts
            assert np.array_equal(got[:, 0:3], ref_forces)
            assert np.all(got[:, 3] == 0.0)
            counts_seen.append((len(cand), got_counts))
        # more candidates than the limit: rejected on the host, nothing enqueued
        big = NOALLOC_MAX_CONTACTS + 1
        rc = lib.fsi_mesh_contact_forces_noalloc(m.h, C.c_void_p(cd.data_ptr()), big, 0.08, 1, 0.05,
                                                 C.c_void_p(out.data_ptr()), C.c_void_p(cnt.data_ptr()),
                                                 C.c_void_p(stream.cuda_stream))
        assert rc == ERR_INVALID_ARG and "at most 65536" in lib.fsi_last_error().decode()
    print("calls:", counts_seen)
    assert counts_seen[0][0] > counts_seen[1][0] > 0
    assert counts_seen[0][1][0] > 0 and counts_seen[1][1][0] > 0
