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


# ------------------------------------------------------------ LES pieces --
def _les_pair():
    from channel_amd import closure
    from channel_amd.dns import DnsConfig
    dns = DnsConfig(nx=47, ny=24, nz=47, alfa0=1.0, beta0=2.0, ni=1.0 / 3000.0, a=1.5)
    return dns, closure.les_config(dns, 4)


def test_les_config_matches_the_subsampled_grid():
    from channel_amd import closure
    from cases import config2
    c2 = config2()
    les = closure.les_config(c2, 4)
    assert (les.nx, les.nz, 2 * les.nxd, les.nzd) == (31, 32, 96, 96)
    assert (2 * c2.nxd // 4, c2.nzd // 4) == (96, 96)
    dns, les = _les_pair()
    assert (2 * les.nxd, les.nzd) == (2 * dns.nxd // 4, dns.nzd // 4)


def test_filter_truncate_is_the_box_filter_on_the_coarse_points():
    """For a field whose modes all fit the LES grid, the spectral filter and
    truncation reproduce the physical top-hat filter of the dataset sampled
    at every ratio-th point."""
    from channel_amd import closure
    dns, les = _les_pair()
    rng = np.random.default_rng(5)
    V = np.zeros(dns.field_shape, dtype=np.complex128, order="F")
    nz, nzl, nxl = dns.nz, les.nz, les.nx
    V[:, nz - nzl:nz + nzl + 1, :nxl + 1, :] = rng.standard_normal((dns.ny + 3, 2 * nzl + 1, nxl + 1, 3)) + \
        1j * rng.standard_normal((dns.ny + 3, 2 * nzl + 1, nxl + 1, 3))
    V[:, nz, 0, :] = V[:, nz, 0, :].real
    for iz in range(1, nz + 1):
        V[:, nz - iz, 0, :] = np.conj(V[:, nz + iz, 0, :])
    Mx, Mz = 2 * dns.nxd, dns.nzd
    u = closure._physical(dns, V, Mx, Mz)
    uf = closure._box(closure._box(u, 4, -2), 4, -1)[:, :, ::4, ::4]
    Vl = closure.filter_truncate(dns, les, V, 4)
    ul = closure._physical(les, Vl, 2 * les.nxd, les.nzd)
    np.testing.assert_allclose(ul, uf, atol=1e-12 * np.abs(uf).max())


def test_sgs_force_is_minus_the_divergence():
    """tau with x/z harmonics and quadratic y profiles (exact for the
    second-order y differences): -div(tau) in closed form."""
    from channel_amd import closure
    from channel_amd.dns import y_grid
    _, les = _les_pair()
    Mx, Mz = 2 * les.nxd, les.nzd
    y = y_grid(les)[1:les.ny + 2][:, None, None]
    x = (2 * np.pi / les.alfa0 * np.arange(Mx) / Mx)[None, :, None]
    z = (2 * np.pi / les.beta0 * np.arange(Mz) / Mz)[None, None, :]
    a, b = les.alfa0, les.beta0
    q, dq = y * (2 - y), 2 - 2 * y
    full = lambda f: np.broadcast_to(f, (les.ny + 1, Mx, Mz))
    tau = np.stack([full(np.sin(a * x)), full(q), full(np.cos(2 * b * z)),        # xx, yy, zz
                    full(q * np.cos(a * x)), full(np.sin(b * z)), full(y * y)])  # xy, xz, yz
    F = closure.sgs_force(les, tau)
    Fx = full(-(a * np.cos(a * x) + dq * np.cos(a * x) + b * np.cos(b * z)))
    Fy = full(-(-a * q * np.sin(a * x) + dq))
    Fz = full(-(2 * y - 2 * b * np.sin(2 * b * z)))
    for c, ref in enumerate((Fx, Fy, Fz)):
        got = closure._physical(les, F[..., c:c + 1].repeat(3, axis=3), Mx, Mz)[0]
        np.testing.assert_allclose(got, ref, atol=1e-10)


def test_features_les_gradients():
    from channel_amd import closure
    from channel_amd.dns import y_grid
    _, les = _les_pair()
    Mx, Mz = 2 * les.nxd, les.nzd
    nz, a = les.nz, les.alfa0
    y = y_grid(les)
    V = np.zeros(les.field_shape, dtype=np.complex128, order="F")
    V[:, nz + 1, 1, 0] = 0.5 * y * (2 - y)       # u = Re[q e^{i(ax + bz)}]
    X = closure.features_les(les, V).reshape(les.ny + 1, Mx, Mz, 46)
    yy = y[1:-1][:, None, None]
    x = (2 * np.pi / a * np.arange(Mx) / Mx)[None, :, None]
    z = (2 * np.pi / les.beta0 * np.arange(Mz) / Mz)[None, None, :]
    ph = a * x + les.beta0 * z
    q = yy * (2 - yy)
    np.testing.assert_allclose(X[..., 0], -a * q * np.sin(ph), atol=1e-6)          # du/dx
    np.testing.assert_allclose(X[..., 1], (2 - 2 * yy) * np.cos(ph), atol=1e-6)    # du/dy
    np.testing.assert_allclose(X[..., 2], -les.beta0 * q * np.sin(ph), atol=1e-6)  # du/dz
    np.testing.assert_allclose(X[..., 9], np.roll(X[..., 0], -1, axis=1))          # +x neighbour
    np.testing.assert_allclose(X[..., 45], np.minimum(yy, 2 - yy) * np.ones_like(ph), atol=1e-7)
