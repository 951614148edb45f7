"""The learned subgrid-stress closure on a coarse channel grid, measured
against the filtered fine DNS (the closure has no counterpart in the
reference).

Fine DNS: the production grid 191x384x189 at Re 12431, alfa0 0.5, beta0 1,
CPI type 1 (gamma 0.161436), cflmax 1 -- the library's own GPU step, from a
strongly perturbed laminar start.  Coarse (LES) grid: nx = 31, ny = 64,
nz = 31 with the same physics: its dealiased physical grid 96 x 96 is the
fine 768 x 768 divided by 8, and its stretched rows are every 6th fine row
(the same tanh map of iy/ny).

Filter G: a top-hat of width 8 fine cells in x and in z (transfer function
sin(k D/2)/(k D/2), D the width), truncation to the LES modes |kx| <= 31,
|kz| <= 31, and restriction to the LES rows.  The filtered DNS is G(u); the
subgrid stresses are tau_ij = G(u_i u_j) - G(u_i) G(u_j) on the LES points,
G(u_i u_j) formed from the fine products.  The ghost rows of G(u) are
extrapolated quadratically from the wall and the first two rows (the LES's
first solve replaces them).

Features and closure: chan_les_features / chan_les_closure_force
(include/channel_closure.h) on the LES context -- the nine filtered
velocity gradients at the point and at its +-1, +-2 wall-normal neighbours,
and y+ (46 values) -> MLP 46-64-64-64-6 (BF16 weights, FP32 accumulation,
trained with the library's kernels) -> six stresses -> body force -div(tau),
all on the device.  Training features come from the same device operator
applied to G(u) of the DNS snapshots, so training and use see the same
inputs.

Runs are compared on time-averaged mean and rms profiles (chan_stats, the
spectral sums of postpro/history/var_history.f90:36-87) over one window:
the filtered DNS, the LES without a model and the LES with the closure,
each LES started from G(u) at the window's start.  Costs are per simulated
time unit: wall milliseconds of a step divided by its Delta t.
"""
import dataclasses
import json
import os
import time

import numpy as np
import scipy.fft as sfft

from .dns import DnsConfig


def fine_config():
    return DnsConfig(nx=191, ny=384, nz=189, alfa0=0.5, beta0=1.0, ni=1.0 / 12431.0, a=1.5,
                     cpi=True, cpi_type=1, gamma=0.161436, deltat=0.0, cflmax=1.0)


def les_config():
    return DnsConfig(nx=31, ny=64, nz=31, alfa0=0.5, beta0=1.0, ni=1.0 / 12431.0, a=1.5,
                     cpi=True, cpi_type=1, gamma=0.161436, deltat=0.0, cflmax=1.0)


def ratios(cf, cl):
    """Physical-grid ratios in x and z and the row step in y (fine / LES)."""
    rx, rz = (2 * cf.nxd) // (2 * cl.nxd), cf.nzd // cl.nzd
    ys = cf.ny // cl.ny
    assert rx * 2 * cl.nxd == 2 * cf.nxd and rz * cl.nzd == cf.nzd and ys * cl.ny == cf.ny
    return rx, rz, ys


def tophat(m, M, r):
    """Transfer function of the centred top-hat of width r grid cells of an
    M-point grid, at the signed mode indices m."""
    t = np.pi * np.asarray(m, dtype=np.float64) * r / M
    out = np.ones_like(t)
    nz = t != 0
    out[nz] = np.sin(t[nz]) / t[nz]
    return out


def _les_modes(cf, cl, rx, rz):
    Hx = tophat(np.arange(cl.nx + 1), 2 * cf.nxd, rx)
    Hz = tophat(np.arange(-cl.nz, cl.nz + 1), cf.nzd, rz)
    return Hz[:, None] * Hx[None, :]


def _ghosts(cl, y, W):
    """Quadratic extrapolation of rows -1 and ny+1 of W (reference layout)."""
    def ext(y0, y1, y2, f0, f1, f2, yg):
        l0 = (yg - y1) * (yg - y2) / ((y0 - y1) * (y0 - y2))
        l1 = (yg - y0) * (yg - y2) / ((y1 - y0) * (y1 - y2))
        l2 = (yg - y0) * (yg - y1) / ((y2 - y0) * (y2 - y1))
        return l0 * f0 + l1 * f1 + l2 * f2
    ny = cl.ny
    W[0] = ext(y[1], y[2], y[3], W[1], W[2], W[3], y[0])
    W[ny + 2] = ext(y[ny + 1], y[ny], y[ny - 1], W[ny + 1], W[ny], W[ny - 1], y[ny + 2])
    return W


def filtered(cf, cl, V, yl):
    """G(u): the filtered DNS field on the LES grid, reference layout."""
    rx, rz, ys = ratios(cf, cl)
    H = _les_modes(cf, cl, rx, rz)
    W = np.zeros(cl.field_shape, dtype=np.complex128, order="F")
    rows = 1 + ys * np.arange(cl.ny + 1)  # fine row ys*j -> index ys*j + 1
    sub = V[rows][:, cf.nz - cl.nz:cf.nz + cl.nz + 1, :cl.nx + 1, :]
    W[1:cl.ny + 2] = sub * H[None, :, :, None]
    return _ghosts(cl, yl, W)


def _to_phys(S, nz, Mz, Mx, workers):
    """Spectral (..., 2nz+1, nx+1) -> physical (..., Mz, Mx)."""
    A = np.zeros(S.shape[:-2] + (Mz, Mx // 2 + 1), dtype=np.complex128)
    A[..., np.arange(-nz, nz + 1) % Mz, :S.shape[-1]] = S
    return sfft.irfft2(A, s=(Mz, Mx), axes=(-2, -1), norm="forward", workers=workers)


def _to_spec(f, nz, nx, workers):
    """Physical (..., Mz, Mx) -> spectral (..., 2nz+1, nx+1)."""
    F = sfft.rfft2(f, axes=(-2, -1), norm="forward", workers=workers)
    Mz = f.shape[-2]
    return F[..., np.arange(-nz, nz + 1) % Mz, :nx + 1]


PAIRS = ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2))


def targets(cf, cl, V, workers=None):
    """tau (6, ny+1, 96, 96) on the LES points (x-major as the features:
    [row][x][z]) from the fine field V."""
    workers = workers or os.cpu_count() or 1
    rx, rz, ys = ratios(cf, cl)
    H = _les_modes(cf, cl, rx, rz)
    Mxf, Mzf, Mxl, Mzl = 2 * cf.nxd, cf.nzd, 2 * cl.nxd, cl.nzd
    tau = np.empty((6, cl.ny + 1, Mxl, Mzl))
    for j in range(cl.ny + 1):
        row = 1 + ys * j
        u = _to_phys(np.moveaxis(V[row], -1, 0), cf.nz, Mzf, Mxf, workers)  # (3, Mz, Mx)
        prod = np.stack([u[a] * u[b] for a, b in PAIRS])
        P = _to_spec(prod, cl.nz, cl.nx, workers) * H[None]                 # G(u_a u_b) on LES modes
        Ul = np.moveaxis(V[row][cf.nz - cl.nz:cf.nz + cl.nz + 1, :cl.nx + 1, :], -1, 0) * H[None]
        gu = _to_phys(Ul, cl.nz, Mzl, Mxl, workers)
        gp = _to_phys(P, cl.nz, Mzl, Mxl, workers)
        for q, (a, b) in enumerate(PAIRS):
            tau[q, j] = (gp[q] - gu[a] * gu[b]).T  # [z][x] -> [x][z]
    return tau


def _profiles(ch):
    s = ch.stats()
    return np.concatenate([s["mean"][0:1], np.sqrt(np.maximum(s["var"], 0.0)), s["uv"][None]])  # (5, ny+1)


def run(cf=None, cl=None, seed=11, amp=3e-3, kmax=8, warm_time=120.0, ntrain=10, train_every=2.0,
        window=30.0, sample_every=1.0, train_steps=4000, batch=16384, lr=2e-3, max_samples=4_000_000,
        log=print, out=None):
    """The whole study in one process on one GPU; returns a dict (written to
    `out` as JSON when given)."""
    from .closure import Mlp, _corr, _he_init
    from .initial import perturbed_laminar
    from .solver import Channel
    from . import lib
    cf = cf or fine_config()
    cl = cl or les_config()
    res = {"fine": [cf.nx, cf.ny, cf.nz], "les": [cl.nx, cl.ny, cl.nz], "Re": 1.0 / cf.ni}
    t_all = time.time()
    les = Channel(cl)
    yl = les.setup()[0]

    # ---- fine DNS: transition in the split step, then FP64
    V0 = perturbed_laminar(cf, seed=seed, amp=amp, kmax=kmax)
    dns = Channel(cf, lib.CHAN_SPLIT)
    dns.set_field(V0)
    dns.initial()
    dns.outstats()
    hist = []
    t0 = time.time()
    while dns.scalars()["time"] < warm_time:
        rows = dns.steps(100)
        hist.append(rows[-1].tolist())
        log("dns warm-up t %.1f dudy %.4f dp/dx %.5f fr %.4f dt %.5f" % (rows[-1][0], rows[-1][1], rows[-1][6],
                                                                          rows[-1][5], rows[-1][10]))
    res["warm"] = {"wall_s": time.time() - t0, "history": hist}
    V = dns.get_field()
    sc = dns.scalars()
    dns.close()
    dns = Channel(dataclasses.replace(cf, time=sc["time"], deltat=sc["deltat"]))
    dns.set_field(V)
    dns.initial()
    dns.outstats()

    def advance(ch, T):
        """Steps until simulated time T; returns (steps, wall seconds)."""
        n, w0 = 0, time.time()
        while ch.scalars()["time"] < T - 1e-12:
            ch.step()
            n += 1
        ch.sync()
        return n, time.time() - w0

    # ---- training snapshots
    X, Tg = [], []
    tt = sc["time"]
    for k in range(ntrain):
        if k:
            advance(dns, tt + k * train_every)
        Vs = dns.get_field()
        les.set_field(filtered(cf, cl, Vs, yl))
        X.append(les.les_features())
        Tg.append(np.moveaxis(targets(cf, cl, Vs), 0, -1).reshape(-1, 6).astype(np.float32))
        log("training snapshot %d at t %.2f" % (k, dns.scalars()["time"]))
    # ---- window: filtered DNS profiles, and the DNS's cost per time unit
    t_start = dns.scalars()["time"] + train_every
    advance(dns, t_start)
    Vstart = filtered(cf, cl, dns.get_field(), yl)
    prof_dns, nsteps, wall = [], 0, 0.0
    for k in range(1, int(round(window / sample_every)) + 1):
        n, w = advance(dns, t_start + k * sample_every)
        nsteps += n
        wall += w
        les.set_field(filtered(cf, cl, dns.get_field(), yl))
        prof_dns.append(_profiles(les))
    dns.close()
    res["dns"] = {"steps": nsteps, "wall_s": wall, "ms_per_step": 1e3 * wall / max(nsteps, 1),
                  "dt_mean": window / max(nsteps, 1), "ms_per_time_unit": 1e3 * wall / window}
    log("dns window: %d steps, %.3f ms per time unit" % (nsteps, res["dns"]["ms_per_time_unit"]))

    # ---- training (standardised features and targets), a priori check on the last snapshot
    Xtr = np.concatenate(X[:-1])
    Ttr = np.concatenate(Tg[:-1])
    # shuffled (the minibatches run over consecutive samples), at most max_samples
    idx = np.random.default_rng(seed).permutation(len(Xtr))[:max_samples]
    Xtr, Ttr = Xtr[idx], Ttr[idx]
    Xte, Tte = X[-1], Tg[-1]
    xm, xs = Xtr.mean(0), Xtr.std(0) + 1e-6
    tm, ts = Ttr.mean(0), Ttr.std(0) + 1e-12
    norm = np.concatenate([xm, xs, tm, ts]).astype(np.float32)
    m = Mlp(max(len(Xtr), len(Xte)))
    m.set_params(_he_init(seed))
    m.set_data((Xtr - xm) / xs, (Ttr - tm) / ts)
    loss0 = m.grad(0, len(Xtr))[0]
    losses, ms_step = m.train(train_steps, batch, lr)
    loss1 = m.grad(0, len(Xtr))[0]
    Yte = m.forward((Xte - xm) / xs) * ts + tm
    res["train"] = {"samples": int(len(Xtr)), "loss_before": float(loss0), "loss_after": float(loss1),
                    "ms_per_adam_step": ms_step,
                    "corr_test": [_corr(Yte[:, c], Tte[:, c]) for c in range(6)]}
    log("trained: %s" % json.dumps(res["train"]))

    # ---- LES runs from G(u) at the window's start
    runs = {}
    for model in ("none", "mlp"):
        ch = Channel(dataclasses.replace(cl, time=t_start))
        ch.set_field(Vstart)
        ch.initial()
        ch.outstats()
        prof, n, w0, ok = [], 0, time.time(), True
        for k in range(1, int(round(window / sample_every)) + 1):
            while ch.scalars()["time"] < t_start + k * sample_every - 1e-12:
                if model == "mlp":
                    ch.les_closure_force(m, norm)
                row = ch.step()
                n += 1
                if not np.all(np.isfinite(row)):
                    ok = False
                    break
            if not ok:
                break
            prof.append(_profiles(ch))
        ch.sync()
        wall = time.time() - w0
        ch.close()
        runs[model] = {"finite": ok, "steps": n, "ms_per_step": 1e3 * wall / max(n, 1),
                       "ms_per_time_unit": 1e3 * wall / window if ok else None, "profiles": prof}
        log("les %s: %d steps, finite %s, %.3f ms per time unit" % (model, n, ok, 1e3 * wall / window))
    m.close()
    les.close()

    # ---- comparison of time-averaged profiles
    ref = np.mean(prof_dns, axis=0)  # (5, ny+1): mean u, rms u, v, w, uv
    names = ["mean_u", "rms_u", "rms_v", "rms_w", "uv"]
    res["profiles"] = {"y": yl[1:cl.ny + 2].tolist(), "dns_filtered": ref.tolist()}
    res["errors"] = {}
    for model, r in runs.items():
        if not r["finite"] or not r["profiles"]:
            res["errors"][model] = None
            continue
        pm = np.mean(r["profiles"], axis=0)
        res["profiles"]["les_" + model] = pm.tolist()
        res["errors"][model] = {nm: float(np.max(np.abs(pm[i] - ref[i])) / np.max(np.abs(ref[i])))
                                for i, nm in enumerate(names)}
        r.pop("profiles")
    res["runs"] = runs
    res["cost_ratio_dns_over_les"] = {k: res["dns"]["ms_per_time_unit"] / v["ms_per_time_unit"]
                                      for k, v in runs.items() if v["ms_per_time_unit"]}
    res["wall_s"] = time.time() - t_all
    log("errors %s" % json.dumps(res["errors"]))
    if out:
        with open(out, "w") as f:
            json.dump(res, f)
    return res
