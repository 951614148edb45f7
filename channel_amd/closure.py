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
The reference has no subgrid model; this follows the north star's closure
row (box filter, a priori dataset from a GPU DNS, BF16 MFMA training).
"""
import ctypes

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
