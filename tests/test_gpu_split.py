"""The FP32/FP64 split step (CHAN_SPLIT) against the FP64 oracle and the FP64
HIP step: FP32 transforms, products and product intermediates (K1-K3, the
products read by K4); FP64 state, stencil arithmetic, solves, (0,0) mode,
flow rates and reductions.  Tolerances are set by FP32 rounding of the
nonlinear term: relative L2 field error <= 1e-6 after 1 and 10 steps and
<= 1e-5 after 100, Runtimedata columns to 1e-6 of max(|value|, 1)."""

import numpy as np
import pytest

from cases import config1, laminar_field, rel_l2, seeded_field, tiny
from channel_amd import lib
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

SPLIT_TOL = 1e-6


def say(*a):
    print("[case]", *a, flush=True)


def split_channel(cfg, V):
    from channel_amd.solver import Channel
    g = Channel(cfg, lib.CHAN_SPLIT)
    g.set_field(V)
    return g


def rows_close(rg, ro, tol):
    scale = np.maximum(np.abs(ro), 1.0)
    np.testing.assert_array_less(np.abs(rg - ro), tol * scale)


def test_split_config1_vs_oracle():
    cfg = config1()
    say("split config1 10 steps")
    V = seeded_field(cfg)
    g = split_channel(cfg, V)
    o = orc.Oracle(cfg)
    o.set_field(V)
    errs = []
    for k in range(10):
        rg, ro = g.step(), o.step()
        rows_close(rg, ro, SPLIT_TOL)
        if k + 1 in (1, 10):
            e = rel_l2(g.get_field(), o.get_field())
            errs.append((k + 1, e))
            assert e <= SPLIT_TOL, (k + 1, e)
    print("split rel L2:", errs)
    # the error is FP32-sized, not zero: the split path really ran in FP32
    assert errs[0][1] > 1e-14


def test_split_cpi_adaptive_dt():
    cfg = tiny(nx=15, ny=48, nz=15, alfa0=0.5, beta0=1.0, ni=1 / 12431.0, cpi=True, cpi_type=1, gamma=0.161436,
               meanflowx=0.0, cflmax=1.0, deltat=0.0)
    say("split cpi adaptive 10 steps")
    V = seeded_field(cfg)
    g = split_channel(cfg, V)
    o = orc.Oracle(cfg)
    o.set_field(V)
    g.initial()
    o.initial()
    rows_close(g.outstats(), o.outstats(), SPLIT_TOL)
    for _ in range(10):
        rows_close(g.step(), o.step(), SPLIT_TOL)
    assert rel_l2(g.get_field(), o.get_field()) <= SPLIT_TOL


def test_split_100_steps_vs_fp64_gpu():
    from channel_amd.solver import Channel
    cfg = config1()
    say("split vs fp64 100 steps")
    V = seeded_field(cfg)
    g = split_channel(cfg, V)
    d = Channel(cfg)
    d.set_field(V)
    rs = g.steps(100)
    rd = d.steps(100)
    e = rel_l2(g.get_field(), d.get_field())
    print("split vs fp64 after 100 steps:", e)
    assert e <= 1e-5
    rows_close(rs[-1], rd[-1], 1e-5)


def test_split_laminar_fixed_point():
    cfg = tiny(meanflowx=0.0, meanpx=2.0 / 3000.0)
    say("split laminar")
    V0 = laminar_field(cfg)
    g = split_channel(cfg, V0)
    g.steps(10)
    assert np.max(np.abs(g.get_field() - V0)) < 1e-8


@pytest.mark.parametrize("nx", [255, 511])
def test_split_four_step_x_transform_vs_oracle(nx):
    """The FP32 four-step x transforms in the split step against the FP64
    oracle: 768 points (nx = 255, as at 191x384x189) and the two-wave 1536
    points (nx = 511, as at 511x512x511)."""
    cfg = tiny(nx=nx, ny=16, nz=7)
    say("split, nx =", nx, "3 steps")
    V = seeded_field(cfg)
    g = split_channel(cfg, V)
    o = orc.Oracle(cfg)
    o.set_field(V)
    for _ in range(3):
        rows_close(g.step(), o.step(), SPLIT_TOL)
    e = rel_l2(g.get_field(), o.get_field())
    print("split rel L2 after 3 steps:", e)
    assert 1e-14 < e <= SPLIT_TOL


def _field_stats(V, nz):
    """Per-component fields and the statistics of rows 0..ny (plane means,
    rms about them, uv; weight 2 off kx = 0, postpro/history/var_history.f90)."""
    P = V[1:-1]
    m = P[:, nz, 0, :].real
    f = np.full(P.shape[2], 2.0)
    f[0] = 1.0
    Q = P.copy()
    Q[:, nz, 0, :] -= m
    rms = np.sqrt(np.einsum("x,yzxc->cy", f, np.abs(Q) ** 2))
    uv = np.einsum("x,yzx->y", f, (Q[..., 0] * np.conj(Q[..., 1])).real)
    return {"mean_u": m[:, 0], "rms_u": rms[0], "rms_v": rms[1], "rms_w": rms[2], "uv": uv}


def test_split_errors_per_field_and_statistic():
    """The split step's error for every field and statistic: u, v, w
    (relative L2 per component), the mean-velocity, rms and uv profiles
    (largest deviation over the profile's largest value) and the 11
    Runtimedata columns, against the FP64 oracle after 1 and 10 steps and
    against the FP64 HIP step after 100 (31x64x31)."""
    from channel_amd.solver import Channel
    cfg = config1()
    say("split per field and statistic")
    V = seeded_field(cfg)
    g = split_channel(cfg, V)
    o = orc.Oracle(cfg)
    o.set_field(V)
    d = Channel(cfg)
    d.set_field(V)
    report = []
    done = 0
    for n, tol in ((1, SPLIT_TOL), (10, SPLIT_TOL), (100, 1e-5)):
        rg = g.steps(n - done)[-1]
        rref = (o.steps(n - done) if n <= 10 else d.steps(n - done))[-1]
        if n <= 10:
            d.steps(n - done)
        done = n
        Vg = g.get_field()
        Vr = o.get_field() if n <= 10 else d.get_field()
        errs = {"uvw"[c]: rel_l2(Vg[..., c], Vr[..., c]) for c in range(3)}
        sg, sr = _field_stats(Vg, cfg.nz), _field_stats(Vr, cfg.nz)
        for k in sg:
            errs[k] = float(np.max(np.abs(sg[k] - sr[k])) / np.max(np.abs(sr[k])))
        cols = np.abs(rg - rref) / np.maximum(np.abs(rref), 1.0)
        report.append((n, errs, cols))
        print("after %d steps vs %s:" % (n, "oracle" if n <= 10 else "FP64 HIP step"),
              " ".join("%s %.2e" % kv for kv in errs.items()), "| Runtimedata", np.array2string(cols, precision=1))
        for k, e in errs.items():
            assert e <= tol, (n, k, e)
        assert np.all(cols <= tol), (n, cols)
    o.close()
