"""The learned subgrid-stress closure on top of the channel step.

- ``Mlp``: ctypes binding of include/channel_closure.h (BF16-MFMA MLP
  46 -> 64 -> 64 -> 64 -> 6 with FP32 accumulation; training with Adam).
- ``dataset``: a priori training data from a resolved field: the velocity
  is box-filtered in x and z (top-hat of width ``ratio`` grid cells of the
  dealiased physical grid), subsampled by ``ratio`` in x and z, and each
  sample gets
    features (46): the nine filtered velocity gradients du_i/dx_j at the
                   point and at its four x/z neighbours (45), and the wall
                   distance min(y, 2 - y) (1);
    targets (6):   tau_ij = filt(u_i u_j) - filt(u_i) filt(u_j) for
                   (xx, yy, zz, xy, xz, yz).
  Wall-normal derivatives are second-order differences on the stretched
  grid, x and z derivatives centred differences on the fine grid.
- ``a_posteriori``: the closure inside a coarse step.  An LES grid whose
  dealiased physical grid is the DNS grid subsampled by ``ratio`` runs from
  the filtered DNS field; the modelled stress enters the step as the body
  force -div(tau) (include/channel_hip.h, chan_set_body_force), refreshed
  every step.  Features on the LES side come from the LES field itself
  (spectral x/z derivatives on the coarse grid), and the training data uses
  the same operator on the filtered, truncated DNS field, so that training
  and use see the same inputs.
The reference has no subgrid model: the closure is an addition beside the
DNS step (top-hat filtered GPU DNS data, BF16 MFMA training, a priori and a
posteriori evaluation).
"""
import ctypes
import dataclasses

import numpy as np
from . import lib

IN, OUT, NPARAM = 46, 6, 11718


class Mlp:
    def __init__(self, capacity):
        self._L = lib.load()
        self.h = ctypes.c_void_p()
        lib.check(self._L.chan_mlp_create(int(capacity), ctypes.byref(self.h)))
        self.capacity = int(capacity)

    def close(self):
        if self.h:
            self._L.chan_mlp_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == NPARAM
        lib.check(self._L.chan_mlp_set_params(self.h, lib.fptr(p)))

    def get_params(self):
        p = np.zeros(NPARAM, dtype=np.float32)
        lib.check(self._L.chan_mlp_get_params(self.h, lib.fptr(p)))
        return p

    def set_data(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        assert X.shape[1] == IN and T.shape[1] == OUT and X.shape[0] == T.shape[0]
        lib.check(self._L.chan_mlp_set_data(self.h, lib.fptr(X), lib.fptr(T), X.shape[0]))
        self.n = X.shape[0]

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], OUT), dtype=np.float32)
        lib.check(self._L.chan_mlp_forward(self.h, lib.fptr(X), X.shape[0], lib.fptr(Y)))
        return Y

    def grad(self, first, count):
        g = np.zeros(NPARAM, dtype=np.float32)
        loss = ctypes.c_double()
        lib.check(self._L.chan_mlp_grad(self.h, int(first), int(count), lib.fptr(g), ctypes.byref(loss)))
        return loss.value, g

    def train(self, steps, batch, lr):
        losses = np.zeros(steps, dtype=np.float32)
        ms = ctypes.c_double()
        lib.check(self._L.chan_mlp_train(self.h, int(steps), int(batch), ctypes.c_float(lr), lib.fptr(losses),
                                         ctypes.byref(ms)))
        return losses, ms.value

    def bench_forward(self, reps=10):
        ms = ctypes.c_double()
        lib.check(self._L.chan_mlp_bench_forward(self.h, int(reps), ctypes.byref(ms)))
        return ms.value


# ------------------------------------------------------------- dataset --
def _physical(cfg, V, Mx, Mz):
    """u_c(y, x, z) on the Mx x Mz grid for rows iy = 0..ny: (3, ny+1, Mx, Mz)."""
    nx, nz, ny = cfg.nx, cfg.nz, cfg.ny
    F = np.zeros((3, ny + 1, Mx, Mz), dtype=np.complex128)
    kz = np.arange(-nz, nz + 1) % Mz
    for ix in range(nx + 1):
        F[:, :, ix, kz] = np.moveaxis(V[1:ny + 2, :, ix, :], 2, 0)
        if ix > 0:
            F[:, :, (-ix) % Mx, (-np.arange(-nz, nz + 1)) % Mz] = np.conj(np.moveaxis(V[1:ny + 2, :, ix, :], 2, 0))
    return np.fft.ifft2(F, axes=(2, 3)).real * (Mx * Mz)


def _box(f, ratio, axis):
    """Top-hat average over `ratio` consecutive points (periodic), centred."""
    out = np.zeros_like(f)
    for s in range(ratio):
        out += np.roll(f, s - ratio // 2, axis=axis)
    return out / ratio


def _ddy(cfg, F):
    """Wall-normal derivative of F on rows 0..ny (axis 0): second-order
    differences on the stretched grid, one-sided at the walls (exact on
    quadratics)."""
    from .dns import y_grid
    y = y_grid(cfg)[1:cfg.ny + 2]
    return np.gradient(F, y, axis=0, edge_order=2)


def dataset(cfg, V, ratio=4):
    """(X [n][46], T [n][6]) from the resolved field V (reference layout)."""
    nxd, nzd = cfg.nxd, cfg.nzd
    Mx, Mz = 2 * nxd, nzd
    u = _physical(cfg, V, Mx, Mz)                       # (3, ny+1, Mx, Mz)
    filt = lambda f: _box(_box(f, ratio, -2), ratio, -1)
    uf = filt(u)
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    tau = np.stack([filt(u[i] * u[j]) - uf[i] * uf[j] for i, j in pairs])
    dx = 2 * np.pi / cfg.alfa0 / Mx
    dz = 2 * np.pi / cfg.beta0 / Mz
    # gradients of the filtered field
    g = np.zeros((3, 3) + uf.shape[1:])
    for c in range(3):
        g[c, 0] = (np.roll(uf[c], -1, axis=1) - np.roll(uf[c], 1, axis=1)) / (2 * dx)
        g[c, 1] = _ddy(cfg, uf[c])
        g[c, 2] = (np.roll(uf[c], -1, axis=2) - np.roll(uf[c], 1, axis=2)) / (2 * dz)
    g = g.reshape(9, *uf.shape[1:])
    # subsample to the coarse grid; neighbours one coarse cell away
    sl = (slice(None), slice(None), slice(0, Mx, ratio), slice(0, Mz, ratio))
    gc = g[sl]
    feats = [gc]
    for ax, sh in ((2, 1), (2, -1), (3, 1), (3, -1)):
        feats.append(np.roll(gc, -sh, axis=ax))
    from .dns import y_grid
    y = y_grid(cfg)[1:cfg.ny + 2]
    wd = np.minimum(y - cfg.ymin, cfg.ymax - y)
    W = np.broadcast_to(wd[:, None, None], gc.shape[1:])[None]
    X = np.concatenate(feats + [W], axis=0)             # (46, ny+1, mx, mz)
    T = tau[sl]                                         # (6, ny+1, mx, mz)
    X = X.reshape(46, -1).T
    T = T.reshape(6, -1).T
    return X.astype(np.float32), T.astype(np.float32)


# ------------------------------------------------------- a priori study --
def _smagorinsky_basis(X):
    """-|S| S_ij for the six stress components from the gradient features at
    the point (the eddy-viscosity form; (Cs Delta)^2 is fitted)."""
    g = X[:, :9].astype(np.float64).reshape(-1, 3, 3)  # g[:, i, j] = du_i/dx_j
    S = 0.5 * (g + np.transpose(g, (0, 2, 1)))
    mag = np.sqrt(2.0 * np.sum(S * S, axis=(1, 2)))
    idx = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    B = np.stack([-mag * S[:, i, j] for i, j in idx], axis=1)
    tr = S[:, 0, 0] + S[:, 1, 1] + S[:, 2, 2]
    B[:, :3] += (mag * tr / 3.0)[:, None]  # deviatoric
    return B


def _corr(a, b):
    a = a - a.mean()
    b = b - b.mean()
    return float(np.sum(a * b) / np.sqrt(np.sum(a * a) * np.sum(b * b)))


def _he_init(seed):
    rng = np.random.default_rng(seed)
    shapes = [(IN, 64), (64,), (64, 64), (64,), (64, 64), (64,), (64, OUT), (OUT,)]
    parts = []
    for s in shapes:
        if len(s) == 2:
            lim = np.sqrt(6.0 / s[0])
            parts.append(rng.uniform(-lim, lim, size=s).ravel())
        else:
            parts.append(np.zeros(s))
    return np.concatenate(parts).astype(np.float32)


def a_priori(cfg, V0, steps_between=40, nsnap=3, ratio=4, train_steps=1500, batch=8192, lr=2e-3, seed=3,
             log=print):
    """Runs the DNS from V0 on the GPU, builds filtered datasets from nsnap
    snapshots (training on the first nsnap-1, testing on the last), trains
    the MLP on standardised features/targets and returns held-out
    correlations per stress component for the MLP and for a least-squares
    Smagorinsky fit."""
    from .solver import Channel
    ch = Channel(cfg)
    ch.set_field(V0)
    if cfg.cflmax > 0:  # adaptive time step: the pre-loop CFL pass sets deltat
        ch.initial()
        ch.outstats()
    snaps = []
    for _ in range(nsnap):
        ch.steps(steps_between)
        snaps.append(ch.get_field())
    ch.close()
    data = [dataset(cfg, V, ratio) for V in snaps]
    Xtr = np.concatenate([d[0] for d in data[:-1]])
    Ttr = np.concatenate([d[1] for d in data[:-1]])
    Xte, Tte = data[-1]
    xm, xs = Xtr.mean(0), Xtr.std(0) + 1e-12
    tm, ts = Ttr.mean(0), Ttr.std(0) + 1e-30
    m = Mlp(len(Xtr))
    m.set_params(_he_init(seed))
    m.set_data((Xtr - xm) / xs, (Ttr - tm) / ts)
    losses, ms = m.train(train_steps, batch, lr)
    Yte = m.forward((Xte - xm) / xs) * ts + tm
    m.close()
    corr_mlp = [_corr(Yte[:, c], Tte[:, c]) for c in range(OUT)]
    # Smagorinsky: tau_ij = C B_ij with one C for the deviatoric part,
    # least squares on the training set
    Btr, Bte = _smagorinsky_basis(Xtr), _smagorinsky_basis(Xte)
    C = float(np.sum(Btr * (Ttr - Ttr.mean(0))) / np.sum(Btr * Btr))
    corr_smag = [_corr(C * Bte[:, c], Tte[:, c]) for c in range(OUT)]
    out = dict(n_train=len(Xtr), n_test=len(Xte), loss_first=float(losses[0]), loss_last=float(losses[-20:].mean()),
               ms_per_step=ms, corr_mlp=corr_mlp, corr_smagorinsky=corr_smag)
    log(out)
    return out


# ------------------------------------------------------ a posteriori LES --
def les_config(cfg, ratio=4):
    """The LES grid: same y grid and flow parameters, and nx, nz such that
    its dealiased physical grid (2 nxd, nzd) is the DNS's divided by ratio."""
    from .dns import dealiased_sizes
    Mx, Mz = 2 * cfg.nxd // ratio, cfg.nzd // ratio
    nx, nz = Mx // 3 - 1, Mz // 3
    while nx > 0 and 2 * dealiased_sizes(nx, nz)[0] != Mx:
        nx -= 1
    while nz > 0 and dealiased_sizes(nx, nz)[1] != Mz:
        nz -= 1
    if nx < 1 or nz < 1:
        raise ValueError("no LES grid for ratio %d" % ratio)
    return dataclasses.replace(cfg, nx=nx, nz=nz)


def box_transfer(k, M, ratio):
    """Transfer function of the top-hat filter of _box (offsets ratio//2 - s,
    s < ratio) for the signed wavenumber indices k on an M-point grid."""
    k = np.asarray(k, dtype=np.float64)
    off = ratio // 2 - np.arange(ratio)
    return np.exp(2j * np.pi * np.outer(k, off) / M).mean(axis=1)


def filter_truncate(cfg, cfg_les, V, ratio=4):
    """The box-filtered DNS field V on the LES modes (all rows, ghost rows
    included): H(kx) H(kz) V for ix <= nx_les, |iz| <= nz_les."""
    nxl, nzl, nz = cfg_les.nx, cfg_les.nz, cfg.nz
    Hx = box_transfer(np.arange(nxl + 1), 2 * cfg.nxd, ratio)
    Hz = box_transfer(np.arange(-nzl, nzl + 1), cfg.nzd, ratio)
    W = V[:, nz - nzl:nz + nzl + 1, :nxl + 1, :]
    return np.asfortranarray(W * Hz[None, :, None, None] * Hx[None, None, :, None])


def _spectral(cfg, f):
    """Physical rows (..., ny+1, Mx, Mz) -> spectral (..., ny+1, 2nz+1, nx+1)."""
    Mx, Mz = f.shape[-2], f.shape[-1]
    F = np.fft.fft2(f, axes=(-2, -1)) / (Mx * Mz)
    iz = np.arange(-cfg.nz, cfg.nz + 1) % Mz
    return np.swapaxes(F[..., :cfg.nx + 1, :][..., iz], -1, -2)


def features_les(cfg_les, V):
    """Closure features (n, 46) of an LES field: velocity gradients on the
    LES's dealiased physical grid (spectral in x and z, the stretched-grid
    differences of _ddy in y) at the point and its four x/z neighbours, and
    the wall distance."""
    from .dns import y_grid
    Mx, Mz = 2 * cfg_les.nxd, cfg_les.nzd
    ix = np.arange(cfg_les.nx + 1)
    iz = np.arange(-cfg_les.nz, cfg_les.nz + 1)
    u = _physical(cfg_les, V, Mx, Mz)
    ux = _physical(cfg_les, V * (1j * cfg_les.alfa0 * ix)[None, None, :, None], Mx, Mz)
    uz = _physical(cfg_les, V * (1j * cfg_les.beta0 * iz)[None, :, None, None], Mx, Mz)
    ny1 = cfg_les.ny + 1
    # samples in (row, x, z) order with the 46 features contiguous per sample
    X = np.empty((ny1, Mx, Mz, 46), dtype=np.float32)
    for c in range(3):
        X[..., 3 * c + 0] = ux[c]
        X[..., 3 * c + 1] = _ddy(cfg_les, u[c])
        X[..., 3 * c + 2] = uz[c]
    for n, (ax, sh) in enumerate(((1, 1), (1, -1), (2, 1), (2, -1))):
        X[..., 9 * (n + 1):9 * (n + 2)] = np.roll(X[..., 0:9], -sh, axis=ax)
    y = y_grid(cfg_les)[1:cfg_les.ny + 2]
    X[..., 45] = np.minimum(y - cfg_les.ymin, cfg_les.ymax - y)[:, None, None]
    return X.reshape(-1, 46)


def dataset_les(cfg, V, ratio=4):
    """Training pairs for use inside the LES: features_les of the filtered,
    truncated DNS field and the subgrid stresses of dataset() at the same
    points."""
    cfg_les = les_config(cfg, ratio)
    X = features_les(cfg_les, filter_truncate(cfg, cfg_les, V, ratio))
    _, T = dataset(cfg, V, ratio)
    return X, T


def sgs_force(cfg_les, tau):
    """Body force -div(tau) in the layout of V from the six stresses
    tau (6, ny+1, Mx, Mz) = (xx, yy, zz, xy, xz, yz) on the LES's physical
    grid; x and z derivatives spectral, y derivatives as in _ddy."""
    T = _spectral(cfg_les, tau)  # (6, ny+1, 2nz+1, nx+1)
    ikx = (1j * cfg_les.alfa0 * np.arange(cfg_les.nx + 1))[None, None, :]
    ikz = (1j * cfg_les.beta0 * np.arange(-cfg_les.nz, cfg_les.nz + 1))[None, :, None]
    dy = lambda f: _ddy(cfg_les, f)
    xx, yy, zz, xy, xz, yz = T
    F = np.zeros(cfg_les.field_shape, dtype=np.complex128, order="F")
    F[1:-1, :, :, 0] = -(ikx * xx + dy(xy) + ikz * xz)
    F[1:-1, :, :, 1] = -(ikx * xy + dy(yy) + ikz * yz)
    F[1:-1, :, :, 2] = -(ikx * xz + dy(yz) + ikz * zz)
    return F


def _fluct_err(V, R):
    """Relative L2 distance of the fluctuations (every mode but (0,0)), and
    of the mean profiles, on rows 0..ny."""
    nz = (V.shape[1] - 1) // 2
    d = (V - R)[1:-1].copy()
    r = R[1:-1].copy()
    dm, rm = d[:, nz, 0, :].copy(), r[:, nz, 0, :].copy()
    d[:, nz, 0, :] = 0
    r[:, nz, 0, :] = 0
    return float(np.linalg.norm(d) / np.linalg.norm(r)), float(np.abs(dm).max() / np.abs(rm).max())


def a_posteriori(cfg, V0, ratio=4, warm_steps=100, snap_every=25, nsnap=4, horizon=100, check_every=20,
                 train_steps=2000, batch=8192, lr=2e-3, seed=3, models=("none", "mlp", "smagorinsky"), log=print):
    """DNS on the GPU from V0 (adaptive time step during the warm-up, then a
    fixed one), nsnap snapshots snap_every steps apart for training (the
    last one is also the LES start), a reference DNS run of `horizon` steps
    from there, and three LES runs from the filtered start with the same
    time step: no model, the MLP closure, and a Smagorinsky model with its
    constant fitted to the training stresses.  Returns, at every check_every
    steps, each LES's relative L2 distance from the filtered DNS in the
    fluctuations and in the mean profiles."""
    from .solver import Channel
    cfg_les = les_config(cfg, ratio)
    ch = Channel(cfg)
    ch.set_field(V0)
    if cfg.cflmax > 0:
        ch.initial()
        ch.outstats()
    ch.steps(warm_steps)
    dt = ch.scalars()["deltat"]
    V = ch.get_field()
    ch.close()
    fixed = dataclasses.replace(cfg, cflmax=0.0, deltat=dt)
    ch = Channel(fixed)
    ch.set_field(V)
    snaps = [V]
    for _ in range(nsnap - 1):
        ch.steps(snap_every)
        snaps.append(ch.get_field())
    # reference trajectory from the last snapshot
    ref = [filter_truncate(cfg, cfg_les, snaps[-1], ratio)]
    for _ in range(horizon // check_every):
        ch.steps(check_every)
        ref.append(filter_truncate(cfg, cfg_les, ch.get_field(), ratio))
    ch.close()
    log("a posteriori: DNS done, dt %.3g; building %d datasets" % (dt, len(snaps)))
    # training on all snapshots but the last, which is the a priori test set
    data = [dataset_les(cfg, Vs, ratio) for Vs in snaps]
    Xtr = np.concatenate([d[0] for d in data[:-1]])
    Ttr = np.concatenate([d[1] for d in data[:-1]])
    Xte, Tte = data[-1]
    xm, xs = Xtr.mean(0), Xtr.std(0) + 1e-12
    tm, ts = Ttr.mean(0), Ttr.std(0) + 1e-30
    m = Mlp(max(len(Xtr), len(Xte)))
    m.set_params(_he_init(seed))
    m.set_data((Xtr - xm) / xs, (Ttr - tm) / ts)
    losses, _ = m.train(train_steps, batch, lr)
    Yte = m.forward((Xte - xm) / xs) * ts + tm
    corr = [_corr(Yte[:, c], Tte[:, c]) for c in range(OUT)]
    Btr = _smagorinsky_basis(Xtr)
    C = float(np.sum(Btr * (Ttr - Ttr.mean(0))) / np.sum(Btr * Btr))
    log("a posteriori: trained, loss %.3g -> %.3g, held-out correlations %s, Smagorinsky C %.3g"
        % (losses[0], losses[-20:].mean(), np.round(corr, 3).tolist(), C))
    Mx, Mz = 2 * cfg_les.nxd, cfg_les.nzd
    shape = (cfg_les.ny + 1, Mx, Mz)

    def stress(model, Vl):
        X = features_les(cfg_les, Vl)
        if model == "mlp":
            Y = m.forward((X - xm) / xs) * ts + tm
        else:
            Y = C * _smagorinsky_basis(X)
        return np.moveaxis(Y.astype(np.float64).reshape(shape + (OUT,)), -1, 0)

    runs = {}
    for model in models:
        les = Channel(dataclasses.replace(fixed, nx=cfg_les.nx, nz=cfg_les.nz))
        les.set_field(ref[0])
        errs = [_fluct_err(ref[0], ref[0])]
        for k in range(1, len(ref)):
            for _ in range(check_every):
                if model != "none":
                    les.set_body_force(sgs_force(cfg_les, stress(model, les.get_field())))
                les.step()
            Vl = les.get_field()
            if not np.all(np.isfinite(Vl)):
                errs.append((float("inf"), float("inf")))
                break
            errs.append(_fluct_err(Vl, ref[k]))
            log("a posteriori: %s step %d fluct/mean error %.4g %.4g" % (model, k * check_every, *errs[-1]))
        les.close()
        runs[model] = errs
    m.close()
    out = dict(dt=dt, les=(cfg_les.nx, cfg_les.ny, cfg_les.nz), steps=[k * check_every for k in range(len(ref))],
               loss_first=float(losses[0]), loss_last=float(losses[-20:].mean()), corr_mlp=corr, smagorinsky_c=C,
               fluct_err={k: [e[0] for e in v] for k, v in runs.items()},
               mean_err={k: [e[1] for e in v] for k, v in runs.items()})
    log(out)
    return out
