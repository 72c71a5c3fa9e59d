"""CPU checks of the adjoint's test helpers: the torch splat and its numpy FP64
twin agree, and the FP64 oracle loss is smooth enough for central differences
(a quadratic fit through five steps)."""
import numpy as np
import pytest

from adjoint_util import bspline2_splat_np, oracle_loss_f64
from conftest import dam_break_config


def test_torch_and_numpy_splats_agree():
    torch = pytest.importorskip("torch")
    import fsi_torch

    rng = np.random.default_rng(0)
    pos = rng.uniform(-0.1, 1.3, size=(500, 3))
    a = fsi_torch.density_splat(torch.tensor(pos), (0, 0, 0), (1.2, 1.2, 1.2), 8).numpy()
    b = bspline2_splat_np(pos, (0, 0, 0), (1.2, 1.2, 1.2), 8)
    np.testing.assert_allclose(a, b, rtol=0, atol=1e-12)
    inside = (pos > 0.15 + 1e-9).all(1) & (pos < 1.05 - 1e-9).all(1)
    # weights of a particle whose nodes are all inside sum to one
    c = bspline2_splat_np(pos[inside], (0, 0, 0), (1.2, 1.2, 1.2), 8)
    assert abs(c.sum() - inside.sum()) < 1e-9


def test_splat_gradient_is_the_analytic_derivative():
    torch = pytest.importorskip("torch")
    import fsi_torch

    rng = np.random.default_rng(1)
    pos = torch.tensor(rng.uniform(0.3, 0.9, size=(40, 3)), requires_grad=True)
    w = torch.tensor(rng.normal(size=(8, 8, 8)))
    (fsi_torch.density_splat(pos, (0, 0, 0), (1.2, 1.2, 1.2), 8) * w).sum().backward()
    p = pos.detach().numpy()
    e = np.zeros_like(p)
    e[3, 1] = 1.0
    h = 1e-6
    fd = ((bspline2_splat_np(p + h * e, (0, 0, 0), (1.2,) * 3, 8) * w.numpy()).sum() -
          (bspline2_splat_np(p - h * e, (0, 0, 0), (1.2,) * 3, 8) * w.numpy()).sum()) / (2 * h)
    assert abs(pos.grad[3, 1].item() - fd) < 1e-6 * max(1.0, abs(fd))


def test_fp64_loss_is_smooth_in_eps():
    import fsi_pbf

    cfg = dam_break_config(n_iter=2, substeps=2, box=(1.2, 1.2, 1.2), hash_size=1 << 12, viscosity=0.01)
    parts = fsi_pbf.lattice(6, 5, 5, (0.0, 0.0, 0.0), 0.05, 0.001, 3)
    v0 = np.random.default_rng(3).normal(0, 0.1, size=(len(parts), 3))
    target = np.zeros((8, 8, 8))
    ls = [oracle_loss_f64(cfg, parts, v0, 0.01, 600.0 + k * 0.06, 2, target, (0, 0, 0), (1.2,) * 3, 8)
          for k in range(-2, 3)]
    c = np.polyfit(np.arange(-2, 3) * 0.06, ls, 2)
    resid = np.abs(np.polyval(c, np.arange(-2, 3) * 0.06) - ls).max()
    assert resid <= 1e-9 * max(1.0, abs(ls[2]))
