"""Feedback tables of the reference's body-force modules.

The reference selects one module at compile time (header.h:28-31) and calls
its ``set_body_force`` before every ``buildrhs`` (channel.f90:129-157); each
one writes ``F(iy,iz,ix,pc) = g * V(iy,iz,ix,c)`` where a fixed mask holds.
All three are linear per mode and per row, so here each becomes a table for
``chan_set_body_force_feedback``: ``src[d]`` is the velocity component that
force component ``d`` copies and ``gain[d, iy+1, col]`` the factor, with
``col = iz+nz`` for every mode except (ix,iz) = (0,0), which reads column
``2nz+1``.  The device forms the force from the field inside the rhs kernel.

Constants keep the reference's kinds: ``PI`` and the literals ``2.3`` and
``1E10`` are default-real (single precision) values (dnsdata.f90:26,
am_f1.inc:18-22).
"""
import math

import numpy as np

from .dns import y_grid

PI_REF = float(np.float32(math.pi))  # dnsdata.f90:26, a default-real literal
_R23 = float(np.float32(2.3))
_R1E10 = float(np.float32(1e10))


def _nint(x):
    """Fortran NINT: nearest integer, halves away from zero."""
    return int(math.floor(x + 0.5)) if x >= 0 else -int(math.floor(-x + 0.5))


def _table(cfg):
    return np.zeros((3, cfg.ny + 3, 2 * cfg.nz + 2))


def _yplus(cfg, y):
    # MERGE(ymax-y(iy), y(iy), y(iy)>1)*1000: wall distance in units of
    # 1/1000 of the half height (Re_tau 1000)
    return np.where(y > 1, cfg.ymax - y, y) * 1000


def iz_f(cfg, lambdaz_f=500.0):
    """NINT((2 PI / lambdaz_f) / (beta0 / 1000)) (am_f1.inc:7)."""
    return _nint((2.0 * PI_REF / lambdaz_f) / (cfg.beta0 / 1000.0))


def am_f1(cfg, amp=1000.0, lambdaz_f=500.0, y=None):
    """body_forces/am_f1/am_f1.inc:11-27: F = -amp V on the modes with
    |iz| <= iz_f whose spanwise wavelength in wall units exceeds 2.3 y+^2,
    except the mean (0,0).  The reference's loop runs over -iz_f..iz_f; here
    iz_f is capped at nz, the last mode that exists."""
    y = y_grid(cfg) if y is None else np.asarray(y, dtype=np.float64)
    nz = cfg.nz
    g = _table(cfg)
    yp = _yplus(cfg, y)
    k = min(iz_f(cfg, lambdaz_f), nz)
    for iz in range(-k, k + 1):
        lzp = _R1E10 if iz == 0 else 2 * PI_REF / (cfg.beta0 * abs(iz)) * 1000
        on = lzp > _R23 * yp ** 2
        g[:, :, iz + nz] = np.where(on, -amp, 0.0)
    # (ix,iz) = (0,0) is excluded: column 2nz+1 stays zero
    return (0, 1, 2), g


def am_butterfly(cfg, amp=1000.0, lambdaz_f=500.0, y=None):
    """body_forces/am_butterfly/am_butterfly.inc:11-27: F = -amp V for
    |iz| <= iz_f within y+ <= 60 of a wall (mean excluded) and for
    |iz| > iz_f farther out."""
    y = y_grid(cfg) if y is None else np.asarray(y, dtype=np.float64)
    nz = cfg.nz
    g = _table(cfg)
    yp = _yplus(cfg, y)
    k = iz_f(cfg, lambdaz_f)
    for iz in range(-nz, nz + 1):
        on = (yp <= 60) if abs(iz) <= k else (yp > 60)
        g[:, :, iz + nz] = np.where(on, -amp, 0.0)
    # the mean mode falls in the first branch, which excludes it
    return (0, 1, 2), g


def coriolis(cfg, omega2, kz_cutoff=9999999.0, y_threshold_bot=1.0, y=None):
    """body_forces/coriolis/coriolis.inc:4-39: spanwise rotation,
    F_v = +omega2 u and F_u = -omega2 v on |iz| <= min(nz, floor(kz_cutoff /
    beta0)) and on the rows with y <= y_threshold_bot or y >= ymax -
    y_threshold_bot; every ix, the mean mode included."""
    y = y_grid(cfg) if y is None else np.asarray(y, dtype=np.float64)
    nz = cfg.nz
    g = _table(cfg)
    top = cfg.ymax - y_threshold_bot
    rows = (y <= y_threshold_bot) | (y >= top)
    thr = min(nz, int(math.floor(kz_cutoff / cfg.beta0)))
    for iz in range(-thr, thr + 1):
        g[0, :, iz + nz] = np.where(rows, -1 * omega2, 0.0)
        g[1, :, iz + nz] = np.where(rows, 1 * omega2, 0.0)
    if thr >= 0:
        g[:, :, 2 * nz + 1] = g[:, :, nz]
    return (1, 0, -1), g


def read_coriolis_in(path):
    """coriolis.in: 2*Ro, kz_cutoff, y_threshold (one value per line, list
    directed; anything after the value is ignored)."""
    vals = []
    with open(path) as f:
        for line in f:
            tok = line.split()
            if tok:
                vals.append(float(tok[0].replace("d", "e").replace("D", "e")))
            if len(vals) == 3:
                break
    if len(vals) != 3:
        raise ValueError("%s: expected 3 values, got %d" % (path, len(vals)))
    return dict(omega2=vals[0], kz_cutoff=vals[1], y_threshold_bot=vals[2])
