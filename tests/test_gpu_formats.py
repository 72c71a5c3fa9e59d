"""The render buffer of bindParticlesForRendering (fsi_pbf_write_render_buffer):
the 48-byte std430 Particle records the particle vertex shader reads, written
on the device, equal to what read_particles copies back."""
import numpy as np
import pytest

from conftest import dam_break_config

torch = pytest.importorskip("torch")  # before the library is loaded: one HIP runtime

pytestmark = pytest.mark.gpu


def test_render_buffer_is_the_particle_records(gpu):
    import fsi_pbf

    cfg = dam_break_config(n_iter=2, substeps=1)
    parts = fsi_pbf.lattice(16, 8, 8, (0.2, 0.2, 0.2), 0.05, 0.001, 7)
    parts[:, 4:7] += 0.125  # predPos differs from pos: it must come through as set
    parts[:, 3], parts[:, 7], parts[:, 11] = 1.5, 2.5, 3.5  # pads survive
    dev = torch.device("cuda:0")
    with fsi_pbf.Solver(cfg) as s:
        s.set_particles(parts)
        s.step(3)
        buf = torch.full((len(parts), 12), float("nan"), device=dev, dtype=torch.float32)
        stream = torch.cuda.current_stream(dev)
        s.write_render_buffer(buf.data_ptr(), stream.cuda_stream)
        stream.synchronize()
        got = buf.cpu().numpy()
        want = s.read_particles()
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    assert np.array_equal(got[:, 4:8], parts[:, 4:8])
    assert not np.array_equal(got[:, 0:3], parts[:, 0:3])  # the state moved
