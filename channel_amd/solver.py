"""Host-side mirror of the reference's per-step interface on top of the HIP
library: the routines of MODULE dnsdata that PROGRAM channel calls
(buildrhs, linsolve, outstats, read/save_restart_file) and the program's time
loop itself (channel.f90:95-181).
"""
import ctypes
import os

import numpy as np

from . import lib
from .dns import DnsConfig, read_restart, runtimedata_line, save_restart

RK1 = (120.0 / 32.0, 2.0, 0.0)
RK2 = (120.0 / 8.0, 50.0 / 8.0, 34.0 / 8.0)
RK3 = (120.0 / 20.0, 90.0 / 20.0, 50.0 / 20.0)


class Channel:
    """Device-resident state of one channel DNS (the module globals of
    dnsdata.f90:25-90).  Fields in and out use the reference layout
    V(iy,iz,ix,ic): numpy complex128, shape (ny+3, 2nz+1, nx+1, 3), order F."""

    def __init__(self, cfg: DnsConfig, precision=lib.CHAN_FP64):
        self.cfg = cfg
        self._L = lib.load()
        self._p = lib.params_from(cfg, precision)
        h = ctypes.c_void_p()
        lib.check(self._L.chan_create(ctypes.byref(self._p), ctypes.byref(h)))
        self.h = h
        s = (ctypes.c_int32 * 5)()
        lib.check(self._L.chan_sizes(self.h, s))
        self.nx, self.ny, self.nz, self.nxd, self.nzd = list(s)
        self.shape = cfg.field_shape

    def close(self):
        if getattr(self, "h", None):
            self._L.chan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ----------------------------------------------------------- state --
    def set_field(self, V):
        V = np.asfortranarray(V, dtype=np.complex128)
        assert V.shape == self.shape, (V.shape, self.shape)
        lib.check(self._L.chan_set_field(self.h, lib.dptr(V)))

    def get_field(self):
        V = np.zeros(self.shape, dtype=np.complex128, order="F")
        lib.check(self._L.chan_get_field(self.h, lib.dptr(V)))
        return V

    def get_rhs(self):
        R = np.zeros((self.ny - 1, 2 * self.nz + 1, self.nx + 1, 2), dtype=np.complex128, order="F")
        lib.check(self._L.chan_get_rhs(self.h, lib.dptr(R)))
        return R

    def scalars(self):
        s = lib.ChanScalars()
        lib.check(self._L.chan_get_scalars(self.h, ctypes.byref(s)))
        return dict(time=s.time, deltat=s.deltat, meanpx=s.meanpx, meanpz=s.meanpz, corrpx=s.corrpx,
                    corrpz=s.corrpz, fr1=s.fr[0], fr2=s.fr[1], fr3=s.fr[2], cfl=s.cfl)

    def set_scalars(self, **kw):
        d = self.scalars()
        d.update(kw)
        s = lib.ChanScalars()
        s.time, s.deltat, s.meanpx, s.meanpz = d["time"], d["deltat"], d["meanpx"], d["meanpz"]
        s.corrpx, s.corrpz, s.cfl = d["corrpx"], d["corrpz"], d["cfl"]
        s.fr[0], s.fr[1], s.fr[2] = d["fr1"], d["fr2"], d["fr3"]
        lib.check(self._L.chan_set_scalars(self.h, ctypes.byref(s)))

    def setup(self):
        y = np.zeros(self.ny + 3)
        der = np.zeros((self.ny - 1, 4, 5))
        bnd = np.zeros((10, 5))
        lib.check(self._L.chan_get_setup(self.h, lib.dptr(y), lib.dptr(der), lib.dptr(bnd)))
        return y, der, bnd

    # ------------------------------------------------------ the step --
    def buildrhs(self, ode, compute_cfl=False):
        o = np.ascontiguousarray(ode, dtype=np.float64)
        lib.check(self._L.chan_buildrhs(self.h, lib.dptr(o), int(bool(compute_cfl))))

    def linsolve(self, lam):
        lib.check(self._L.chan_linsolve(self.h, float(lam)))

    def initial(self):
        lib.check(self._L.chan_initial(self.h))

    def outstats(self):
        r = np.zeros(11)
        lib.check(self._L.chan_outstats(self.h, lib.dptr(r)))
        return r

    def step(self):
        r = np.zeros(11)
        lib.check(self._L.chan_step(self.h, lib.dptr(r)))
        return r

    def stats(self, avg=None):
        """Statistics profiles of the current field for rows 0..ny: dict of
        mean (3, ny+1), var (3, ny+1) and uv (ny+1,); avg (3, ny+1) is the mean
        subtracted from the (0,0) mode (default: the field's own)."""
        n1 = self.cfg.ny + 1
        out = np.zeros((7, n1))
        a = None
        if avg is not None:
            a = np.ascontiguousarray(avg, dtype=np.float64).reshape(3, n1)
        lib.check(self._L.chan_stats(self.h, lib.dptr(a) if a is not None else None, lib.dptr(out)))
        return {"mean": out[0:3], "var": out[3:6], "uv": out[6]}

    # ---------------------------------------------------- body force --
    def set_body_force(self, F):
        """Static body force in the layout of V (None removes it)."""
        if F is None:
            lib.check(self._L.chan_set_body_force(self.h, None))
            return
        F = np.asfortranarray(F, dtype=np.complex128)
        assert F.shape == self.shape, (F.shape, self.shape)
        lib.check(self._L.chan_set_body_force(self.h, lib.dptr(F)))

    def set_body_force_feedback(self, src, gain):
        """F_d += gain[d, iy+1, col] * V[..., src[d]] at every buildrhs; see
        channel_amd.body_force for the tables of the reference's modules.
        src=None removes it."""
        if src is None:
            lib.check(self._L.chan_set_body_force_feedback(self.h, None, None))
            return
        s = (ctypes.c_int32 * 3)(*[int(v) for v in src])
        g = np.ascontiguousarray(gain, dtype=np.float64)
        assert g.shape == (3, self.ny + 3, 2 * self.nz + 2), g.shape
        lib.check(self._L.chan_set_body_force_feedback(self.h, s, lib.dptr(g)))

    def get_body_force(self):
        """The force the next buildrhs applies (wall and ghost rows included)."""
        F = np.zeros(self.shape, dtype=np.complex128, order="F")
        lib.check(self._L.chan_get_body_force(self.h, lib.dptr(F)))
        return F

    def keep_body_force(self, on=True):
        """Keep a copy of the force each step applies (its third substep's)."""
        lib.check(self._L.chan_keep_body_force(self.h, 1 if on else 0))

    def get_applied_body_force(self):
        """The force the last step applied, as the reference's F holds it
        after a time step (Force.cart.N.out); needs keep_body_force()."""
        F = np.zeros(self.shape, dtype=np.complex128, order="F")
        lib.check(self._L.chan_get_applied_body_force(self.h, lib.dptr(F)))
        return F

    # ------------------------------------------------ learned closure --
    def les_samples(self):
        n = ctypes.c_int64()
        lib.check(self._L.chan_les_samples(self.h, ctypes.byref(n)))
        return n.value

    def les_features(self):
        """Closure features (n, 46) float32 of the resident field
        (include/channel_closure.h)."""
        X = np.zeros((self.les_samples(), 46), dtype=np.float32)
        lib.check(self._L.chan_les_features(self.h, lib.fptr(X)))
        return X

    def les_closure_force(self, mlp, norm):
        """Static body force := -div(tau) of the MLP's stresses on the
        resident field, on the device; norm = [46 xmean, 46 xscale, 6 tmean,
        6 tscale]."""
        nv = np.ascontiguousarray(norm, dtype=np.float32)
        assert nv.size == 104
        lib.check(self._L.chan_les_closure_force(self.h, mlp.h, lib.fptr(nv)))

    def steps(self, n):
        r = np.zeros((n, 11))
        if n:
            lib.check(self._L.chan_steps(self.h, int(n), lib.dptr(r)))
        return r

    def sync(self):
        lib.check(self._L.chan_sync(self.h))

    # ----------------------------------------------------- profiling --
    def timing(self, enable):
        lib.check(self._L.chan_timing(self.h, int(bool(enable))))

    def kernel_times(self):
        cap = 64
        names = ctypes.create_string_buffer(32 * cap)
        counts = (ctypes.c_int64 * cap)()
        ms = (ctypes.c_double * cap)()
        n = ctypes.c_int32()
        lib.check(self._L.chan_kernel_times(self.h, cap, names, counts, ms, ctypes.byref(n)))
        out = {}
        for i in range(min(n.value, cap)):
            nm = names.raw[32 * i:32 * i + 32].split(b"\0")[0].decode()
            out[nm] = (int(counts[i]), float(ms[i]))
        return out


def run(workdir=".", log=print, precision=lib.CHAN_FP64):
    """PROGRAM channel (channel.f90:16-196) on one GPU: reads dns.in and
    Dati.cart.out (generating u = 1.5 y(2-y) when it is missing,
    dnsdata.f90:704-718), appends to Runtimedata, saves Dati.cart.out at the
    end and Dati.cart.N.out / Dati.cart.out on dt_field / dt_save."""
    cfg = DnsConfig.read(os.path.join(workdir, "dns.in"))
    deltat_from_dnsin = cfg.deltat
    fname = os.path.join(workdir, "Dati.cart.out")
    if os.path.exists(fname):
        V, time, _ = read_restart(fname, cfg)
    else:
        from .dns import y_grid
        V = np.zeros(cfg.field_shape, dtype=np.complex128, order="F")
        y = y_grid(cfg)
        V[:, cfg.nz, 0, 0] = 3 * 0.5 * y * (2 - y)
        time = cfg.time
    rtd = os.path.join(workdir, "Runtimedata")
    rtd_exists = os.path.exists(rtd)
    rows_keep = []
    if cfg.time_from_restart and rtd_exists:
        # get_record (dnsdata.f90:181-218): keep lines up to the restart time
        with open(rtd) as f:
            for ln in f:
                t = ln.split()
                if len(t) != 11:
                    continue
                rows_keep.append(ln)
                sel, dt = float(t[0].replace("D", "E")), float(t[10].replace("D", "E"))
                if abs(sel - time) < 0.5 * dt:
                    cfg.deltat = dt
                    rows_keep.pop()
                    break
    elif not cfg.time_from_restart:
        time = cfg.time
    if cfg.cflmax == 0:
        cfg.deltat = deltat_from_dnsin
    cfg.time = time
    ch = Channel(cfg, precision)
    ch.set_field(V)
    ch.set_scalars(time=time, deltat=cfg.deltat)
    ifield = int(np.floor((time + 0.5 * cfg.deltat) / cfg.dt_field)) if cfg.dt_field > 0 else 0
    time0 = time
    ch.initial()
    out = open(rtd, "w" if not (cfg.time_from_restart and rtd_exists) else "w")
    out.writelines(rows_keep)
    row = ch.outstats()
    out.write(runtimedata_line(row))
    istep, prev_close = 0, False
    deltat = row[10]
    while time < cfg.t_max - deltat / 2.0 and istep < cfg.nstep:
        istep += 1
        row = ch.step()
        time, deltat = row[0], row[10]
        out.write(runtimedata_line(row))
        out.flush()
        if cfg.dt_save > 0 and np.floor((time + 0.5 * deltat) / cfg.dt_save) > np.floor((time - 0.5 * deltat) / cfg.dt_save) and istep > 1:
            save_restart(fname, cfg, ch.get_field(), time)
        if cfg.dt_field > 0 and time + deltat >= (ifield + 1) * cfg.dt_field:
            if prev_close or (np.floor((time + 0.5 * deltat) / cfg.dt_field) > np.floor((time - 0.5 * deltat) / cfg.dt_field) and time > time0):
                ifield += 1
                save_restart(os.path.join(workdir, "Dati.cart.%d.out" % ifield), cfg, ch.get_field(), time)
                prev_close = False
            else:
                prev_close = True
    out.close()
    save_restart(fname, cfg, ch.get_field(), time)
    ch.close()
    return time, istep
