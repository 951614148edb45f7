"""Pins the oracle's nonlinear buildrhs terms (dnsdata.f90:633-661) against
an independent statement of the Kim-Moin-Moser right-hand sides.

The field holds a few low Fourier modes whose y-profiles are polynomials of
degree <= 2.  Then:

* the dealiased products are exact (their modes reach |k| <= 4 <= nx, nz),
  so they can be formed pointwise on any physical grid that resolves them;
* each product is a polynomial of degree <= 4 in y, on which the compact
  stencils are exact: D1[p] = D0[p'], D2[p] = D0[p''] (dnsdata.f90:246-257).

So with H = -div(u u) formed here in physical space (numpy), x and z
derivatives spectral and y derivatives of the polynomials exact, the oracle's
explicit terms must equal D0 applied to

  D2v equation:  h_v   = -d/dy (dH_x/dx + dH_z/dz) + (d2/dx2 + d2/dz2) H_y
  eta equation:  h_eta = dH_x/dz - dH_z/dx
  (0,0) mode:    (H_x + meanpx, H_z + meanpz)

at every row 1..ny-1 and every mode.  The explicit part is isolated from the
oracle's new right-hand side as buildrhs(ODE=(0,1,0)) - buildrhs(ODE=(0,0,0)),
both from the same field.

The mutation test compiles altered copies of the oracle source, each with one
term of the assembly changed (a swapped product, a flipped sign, a swapped
wavenumber), and requires this check to reject every one of them.
"""
import os
import shutil
import subprocess

import numpy as np
import pytest

from cases import tiny
from channel_amd.dns import y_grid
from oracle import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SRC = os.path.join(HERE, "..", "oracle", "chan_oracle.c")

K = 2  # field modes |kx|, |kz| <= K; products reach 2K
MP = 16  # physical points per direction for the pointwise products (> 4K)
PAIRS = {"uu": (0, 0), "vv": (1, 1), "ww": (2, 2), "uv": (0, 1), "vw": (1, 2), "uw": (0, 2)}


def _case(seed=7):
    cfg = tiny(nx=7, nz=7, ny=24, alfa0=1.3, beta0=2.1, ni=1e-6, meanflowx=0.0, meanpx=0.3, meanpz=-0.2,
               deltat=0.01)
    rng = np.random.default_rng(seed)
    # S[c][r][kz mod MP, kx mod MP]: coefficient of y^r of component c
    S = np.zeros((3, 3, MP, MP), dtype=np.complex128)
    for kx in range(0, K + 1):
        for kz in range(-K, K + 1):
            if kx == 0 and kz < 0:
                continue
            c = rng.standard_normal((3, 3)) + 1j * rng.standard_normal((3, 3))
            if kx == 0 and kz == 0:
                c = c.real.astype(np.complex128)
            S[:, :, kz % MP, kx] = c
            S[:, :, (-kz) % MP, (-kx) % MP] = np.conj(c)
    return cfg, S


def _field(cfg, S):
    """V(iy,iz,ix,c) of the polynomial modes on the oracle's grid."""
    y = y_grid(cfg)
    V = np.zeros(cfg.field_shape, dtype=np.complex128, order="F")
    nz = cfg.nz
    for ix in range(cfg.nx + 1):
        for iz in range(-nz, nz + 1):
            coef = S[:, :, iz % MP, ix]
            if not coef.any():
                continue
            for c in range(3):
                V[:, iz + nz, ix, c] = coef[c, 0] + coef[c, 1] * y + coef[c, 2] * y * y
    return V


def _kmm_terms(cfg, S):
    """h_v, h_eta and the (0,0) mean terms as polynomial coefficient arrays
    [r][kz mod MP, kx mod MP], from H = -div(u u) formed pointwise."""
    k = np.fft.fftfreq(MP, 1.0 / MP)
    KZ, KX = np.meshgrid(k, k, indexing="ij")
    ia, ib = 1j * cfg.alfa0 * KX, 1j * cfg.beta0 * KZ
    # physical fields of every polynomial coefficient: u_c(x,z) = sum S e^{i(a kx x + b kz z)}
    phys = np.fft.ifft2(S, axes=(-2, -1)) * MP * MP
    prod = {}
    for name, (a, b) in PAIRS.items():
        q = np.zeros((5, MP, MP), dtype=np.complex128)
        for r in range(3):
            for s in range(3):
                q[r + s] += phys[a, r] * phys[b, s]
        prod[name] = np.fft.fft2(q, axes=(-2, -1)) / (MP * MP)

    def dy(f):
        out = np.zeros_like(f)
        for r in range(1, f.shape[0]):
            out[r - 1] = r * f[r]
        return out

    Hx = -(ia * prod["uu"] + dy(prod["uv"]) + ib * prod["uw"])
    Hy = -(ia * prod["uv"] + dy(prod["vv"]) + ib * prod["vw"])
    Hz = -(ia * prod["uw"] + dy(prod["vw"]) + ib * prod["ww"])
    hv = -dy(ia * Hx + ib * Hz) + (ia * ia + ib * ib) * Hy
    heta = ib * Hx - ia * Hz
    mean_u = Hx[:, 0, 0].real.copy()
    mean_w = Hz[:, 0, 0].real.copy()
    mean_u[0] += cfg.meanpx
    mean_w[0] += cfg.meanpz
    return hv, heta, mean_u, mean_w


def _expected(cfg, S, d0):
    """D0 of the KMM terms at rows 1..ny-1: [iy-1][iz+nz][ix] for eta, D2v."""
    y = y_grid(cfg)
    hv, heta, mu, mw = _kmm_terms(cfg, S)
    ny, nx, nz = cfg.ny, cfg.nx, cfg.nz
    pw = np.stack([y ** r for r in range(5)])  # [r][row]

    def at_rows(coef):  # coef [r][kz][kx] -> values [row][kz][kx]
        return np.tensordot(pw[: coef.shape[0]].T, coef, axes=(1, 0))

    Ev, Ee = at_rows(hv), at_rows(heta)
    Em = (pw[:5].T @ np.pad(mu, (0, 5 - len(mu)))) + 1j * (pw[:5].T @ np.pad(mw, (0, 5 - len(mw))))
    eta = np.zeros((ny - 1, 2 * nz + 1, nx + 1), dtype=np.complex128)
    d2v = np.zeros_like(eta)
    izs = np.arange(-nz, nz + 1) % MP
    for iy in range(1, ny):
        w = d0[iy - 1]
        rows = slice(iy - 1, iy + 4)  # rows iy-2..iy+2 sit at indices iy-1..iy+3
        sv = np.tensordot(w, Ev[rows], axes=(0, 0))
        se = np.tensordot(w, Ee[rows], axes=(0, 0))
        d2v[iy - 1] = sv[izs][:, : nx + 1]
        eta[iy - 1] = se[izs][:, : nx + 1]
        eta[iy - 1, nz, 0] = w @ Em[rows]
    return eta, d2v


def _oracle_explicit(cfg, S, lib_path=None):
    V0 = _field(cfg, S)
    out = []
    for ode in ((0.0, 1.0, 0.0), (0.0, 0.0, 0.0)):
        o = orc.Oracle(cfg, lib_path=lib_path)
        o.set_field(V0)
        o.buildrhs(ode)
        out.append(o.get_field())
        d0 = o.der()[:, 0, :]
        o.close()
    diff = out[0] - out[1]
    return diff[2: cfg.ny + 1, :, :, 0], diff[2: cfg.ny + 1, :, :, 1], d0


def _deviation(cfg, S, lib_path=None):
    """Largest error of the three groups (eta equation, D2v equation, (0,0)
    mean terms), each relative to its own largest expected value."""
    eta, d2v, d0 = _oracle_explicit(cfg, S, lib_path)
    eta_x, d2v_x = _expected(cfg, S, d0)
    nz = cfg.nz
    mean, mean_x = eta[:, nz, 0].copy(), eta_x[:, nz, 0].copy()
    eta[:, nz, 0] = eta_x[:, nz, 0] = 0.0
    dev = 0.0
    for got, want in ((eta, eta_x), (d2v, d2v_x), (mean, mean_x)):
        dev = max(dev, np.max(np.abs(got - want)) / np.max(np.abs(want)))
    return dev


TOL = 1e-12


@pytest.mark.parametrize("seed", [7, 11])
def test_explicit_terms_match_kmm_form(seed):
    cfg, S = _case(seed)
    assert _deviation(cfg, S) < TOL


def test_case_exercises_every_product():
    # the check can only see a product whose terms are nonzero: every product
    # has modes with kx != 0, kz != 0 and a y-dependence of degree 4
    cfg, S = _case()
    k = np.fft.fftfreq(MP, 1.0 / MP)
    phys = np.fft.ifft2(S, axes=(-2, -1)) * MP * MP
    for name, (a, b) in PAIRS.items():
        q4 = np.fft.fft2(phys[a, 2] * phys[b, 2]) / (MP * MP)
        assert np.max(np.abs(q4[np.ix_(k != 0, k != 0)])) > 0.1, name
        assert abs(q4[0, 0]) > 0.1, name


# one term of the assembly changed per mutant (chan_oracle.c, rhs_rows)
MUTANTS = {
    "expl_ww_to_vv": ("imul(be, DD(d1, 3))", "imul(be, DD(d1, 2))"),
    "rhsw_vw_sign": ("- DD(d1, 5) -", "+ DD(d1, 5) -"),
    "rhsu_uu_to_vv": ("-imul(al, DD(d0, 1))", "-imul(al, DD(d0, 2))"),
    "rhsu_uv_sign": ("- DD(d1, 4) - imul(be, DD0_6)", "+ DD(d1, 4) - imul(be, DD0_6)"),
    "rhsv_vv_sign": ("- DD(d1, 2) -", "+ DD(d1, 2) -"),
    "expl_k2_sign": ("- k2 * rhsv;", "+ k2 * rhsv;"),
    "eta_al_be_swap": ("cplx ex = imul(be, rhsu) - imul(al, rhsw);", "cplx ex = imul(al, rhsu) - imul(be, rhsw);"),
    "mean_meanpz": ("creal(rhsw) + o->meanpz", "creal(rhsw) + o->meanpx"),
    "expl_d2uv_to_d2vw": ("+ DD(d2, 4) +", "+ DD(d2, 5) +"),
}


@pytest.fixture(scope="module")
def mutant_libs(tmp_path_factory):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc is needed to build the mutants")
    src = open(ORACLE_SRC).read()
    d = tmp_path_factory.mktemp("mutants")
    shutil.copy(os.path.join(HERE, "..", "oracle", "chan_oracle.h"), d)
    procs = {}
    for name, (old, new) in MUTANTS.items():
        assert src.count(old) == 1, name
        path = d / (name + ".c")
        path.write_text(src.replace(old, new))
        so = str(d / ("lib" + name + ".so"))
        cmd = [gcc, "-O1", "-std=c11", "-fPIC", "-fopenmp", "-fcx-fortran-rules", "-ffp-contract=off", str(path),
               "-o", so, "-shared", "-fopenmp", "-lm"]
        procs[name] = (subprocess.Popen(cmd, stderr=subprocess.PIPE), so)
    libs = {}
    for name, (p, so) in procs.items():
        _, err = p.communicate()
        assert p.returncode == 0, err.decode()
        libs[name] = so
    return libs


@pytest.mark.parametrize("name", sorted(MUTANTS))
def test_check_rejects_mutant(name, mutant_libs):
    cfg, S = _case()
    dev = _deviation(cfg, S, mutant_libs[name])
    assert dev > 1e3 * TOL, "%s passes the check (deviation %.3e)" % (name, dev)
