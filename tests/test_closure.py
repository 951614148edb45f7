"""CPU tests of the closure's checker and dataset builder (no GPU)."""
import numpy as np
import pytest

from cases import config1, laminar_field
from oracle import mlp_oracle as mo


def test_bf16_rounding_nearest_even():
    one = 1.0
    ulp = 2.0 ** -7          # BF16 spacing at 1
    assert mo.bf16(one) == one
    assert mo.bf16(one + ulp / 2) == one               # tie -> even (mantissa 0)
    assert mo.bf16(one + 3 * ulp / 2) == one + 2 * ulp  # tie -> even (mantissa 2)
    assert mo.bf16(one + 0.49 * ulp) == one
    assert mo.bf16(one + 0.51 * ulp) == one + ulp
    assert mo.bf16(-3.0) == -3.0


def test_exact_gradient_matches_finite_differences():
    rng = np.random.default_rng(3)
    p = mo.init_params(5) + 0.01 * rng.standard_normal(mo.NPARAM)  # nonzero biases
    X = rng.standard_normal((12, mo.IN))
    T = rng.standard_normal((12, mo.OUT))
    _, loss, g = mo.exact(p, X, T)
    idx = rng.choice(mo.NPARAM, 60, replace=False)
    h = 1e-6
    for i in idx:
        q = p.copy()
        q[i] += h
        lp = mo.exact(q, X, T)[1]
        q[i] -= 2 * h
        lm = mo.exact(q, X, T)[1]
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_mirrored_equals_exact_on_bf16_representable_data():
    # with every stored value exactly representable the rounding is a no-op
    rng = np.random.default_rng(4)
    p = mo.bf16(mo.init_params(1)) * 0
    W1 = np.zeros((mo.IN, mo.H)); W1[0, 0] = 0.5
    arrs = mo.unpack(p)
    arrs[0] = W1
    arrs[6] = np.zeros((mo.H, mo.OUT)); arrs[6][0, 0] = 2.0
    arrs[2] = np.eye(mo.H); arrs[4] = np.eye(mo.H)
    p = mo.pack(arrs)
    X = np.zeros((4, mo.IN)); X[:, 0] = [1.0, 2.0, -1.0, 4.0]
    y1 = mo.exact(p, X)[0]
    y2 = mo.mirrored(p, X)[0]
    np.testing.assert_array_equal(y1, y2)
    np.testing.assert_array_equal(y1[:, 0], [1.0, 2.0, 0.0, 4.0])


def test_dataset_laminar_closed_form():
    from channel_amd import closure
    from channel_amd.dns import y_grid
    cfg = config1()
    X, T = closure.dataset(cfg, laminar_field(cfg), ratio=4)
    mx, mz = 2 * cfg.nxd // 4, cfg.nzd // 4
    assert X.shape == ((cfg.ny + 1) * mx * mz, 46) and T.shape == (X.shape[0], 6)
    # a laminar profile has no subgrid stress and a single gradient du/dy = 2 - 2y
    assert np.max(np.abs(T)) < 1e-6
    y = y_grid(cfg)[1:cfg.ny + 2]
    dudy = X[:, 1].reshape(cfg.ny + 1, mx, mz)
    np.testing.assert_allclose(dudy[:, 0, 0], 2 - 2 * y, atol=1e-5)
    others = np.delete(X[:, :9], 1, axis=1)
    assert np.max(np.abs(others)) < 1e-6
    np.testing.assert_allclose(X[:, 45].reshape(cfg.ny + 1, mx, mz)[:, 0, 0], np.minimum(y, 2 - y), atol=1e-6)
