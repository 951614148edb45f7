"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the parity checker.  The product path never
imports this module.

Fields are numpy complex128 arrays in the reference's Fortran layout
V(iy,iz,ix,ic): shape (ny+3, 2nz+1, nx+1, 3) with order='F'.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("alfa0", ctypes.c_double), ("beta0", ctypes.c_double),
        ("ni", ctypes.c_double),
        ("a", ctypes.c_double), ("ymin", ctypes.c_double), ("ymax", ctypes.c_double),
        ("cpi", ctypes.c_int32), ("cpi_type", ctypes.c_int32),
        ("gamma", ctypes.c_double),
        ("meanpx", ctypes.c_double), ("meanpz", ctypes.c_double),
        ("meanflowx", ctypes.c_double), ("meanflowz", ctypes.c_double),
        ("u0", ctypes.c_double), ("uN", ctypes.c_double),
        ("deltat", ctypes.c_double), ("cflmax", ctypes.c_double), ("time", ctypes.c_double),
    ]


def build():
    """Compile liboracle.so from oracle/chan_oracle.c (make)."""
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib(path=None):
    """The oracle library; `path` names another build of the same source
    (the mutation tests load deliberately altered copies this way)."""
    key = path or os.path.join(_HERE, "liboracle.so")
    if key not in _LIBS:
        if path is None and not os.path.exists(key):
            build()
        L = ctypes.CDLL(key)
        P = ctypes.c_void_p
        D = ctypes.POINTER(ctypes.c_double)
        I = ctypes.c_int
        L.orc_create.restype = P
        L.orc_create.argtypes = [ctypes.POINTER(OrcParams)]
        L.orc_destroy.argtypes = [P]
        L.orc_set_threads.argtypes = [I]
        L.orc_get_threads.restype = I
        L.orc_sizes.argtypes = [P, ctypes.POINTER(ctypes.c_int32)]
        for n in ("orc_set_field", "orc_get_field", "orc_get_scalars", "orc_set_scalars",
                  "orc_get_der", "orc_get_bnd", "orc_get_bc", "orc_get_d0mat"):
            getattr(L, n).argtypes = [P, D]
        L.orc_get_grid.argtypes = [P, D, D]
        L.orc_convolutions.argtypes = [P, I, I, D]
        L.orc_buildrhs.argtypes = [P, D, I]
        L.orc_linsolve.argtypes = [P, ctypes.c_double]
        L.orc_initial.argtypes = [P]
        L.orc_outstats.argtypes = [P, D]
        L.orc_step.argtypes = [P, D]
        L.orc_steps.argtypes = [P, I, D]
        L.orc_time_sample.argtypes = [P, I, I, D]
        L.orc_fft.argtypes = [I, I, D]
        L.orc_c2r.argtypes = [I, D, D]
        L.orc_r2c.argtypes = [I, D, D]
        L.orc_ludecomp5.argtypes = [D]
        L.orc_lludiv5.argtypes = [D, D, D]
        L.orc_lu5decomp.argtypes = [D, I, I, I]
        L.orc_lu5div1.argtypes = [D, I, D]
        L.orc_lu5div2.argtypes = [D, I, D]
        L.orc_yintegr.restype = ctypes.c_double
        L.orc_yintegr.argtypes = [P, D]
        L.orc_complexderiv.argtypes = [P, D, D]
        L.orc_set_body_force.argtypes = [P, D]
        L.orc_set_body_force_module.argtypes = [P, I, D]
        L.orc_apply_body_force.argtypes = [P]
        L.orc_body_force_walls.argtypes = [P]
        L.orc_get_body_force.argtypes = [P, D]
        _LIBS[key] = L
    return _LIBS[key]


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def set_threads(n):
    lib().orc_set_threads(int(n))


class Oracle:
    """The reference's FP64 step on the CPU (dnsdata.f90, linsolve_blocking.inc,
    rbparmat_blocking.f90, channel.f90).  `cfg` is a channel_amd.dns.DnsConfig
    or any object with the dns.in fields (ni already inverted to 1/Re)."""

    SCALARS = ("time", "deltat", "meanpx", "meanpz", "corrpx", "corrpz", "fr1", "fr2", "fr3", "cfl")

    def __init__(self, cfg, lib_path=None):
        p = OrcParams()
        for name, _ in OrcParams._fields_:
            v = getattr(cfg, name)
            setattr(p, name, int(v) if name in ("nx", "ny", "nz", "cpi", "cpi_type") else float(v))
        self._p = p
        self._L = lib(lib_path)
        self.h = self._L.orc_create(ctypes.byref(p))
        s = (ctypes.c_int32 * 5)()
        self._L.orc_sizes(self.h, s)
        self.nx, self.ny, self.nz, self.nxd, self.nzd = list(s)
        self.shape = (self.ny + 3, 2 * self.nz + 1, self.nx + 1, 3)

    def close(self):
        if self.h:
            self._L.orc_destroy(self.h)
            self.h = None

    __del__ = close

    # field and scalars
    def set_field(self, V):
        V = np.asfortranarray(V, dtype=np.complex128)
        assert V.shape == self.shape
        self._L.orc_set_field(self.h, _dp(V))

    def get_field(self):
        V = np.zeros(self.shape, dtype=np.complex128, order="F")
        self._L.orc_get_field(self.h, _dp(V))
        return V

    def scalars(self):
        s = np.zeros(10)
        self._L.orc_get_scalars(self.h, _dp(s))
        return dict(zip(self.SCALARS, s))

    def set_scalars(self, **kw):
        d = self.scalars()
        d.update(kw)
        s = np.array([d[k] for k in self.SCALARS], dtype=np.float64)
        self._L.orc_set_scalars(self.h, _dp(s))

    # set-up
    def grid(self):
        y = np.zeros(self.ny + 3)
        dy = np.zeros(self.ny - 1)
        self._L.orc_get_grid(self.h, _dp(y), _dp(dy))
        return y, dy

    def der(self):
        d = np.zeros((self.ny - 1, 4, 5))
        self._L.orc_get_der(self.h, _dp(d))
        return d  # [iy-1][d0,d1,d2,d4][-2..2]

    def bnd(self):
        b = np.zeros((10, 5))
        self._L.orc_get_bnd(self.h, _dp(b))
        return dict(zip(("d040", "d140", "d240", "d14m1", "d24m1", "d04n", "d14n", "d24n", "d14np1", "d24np1"), b))

    def bc(self):
        b = np.zeros((8, 5))
        self._L.orc_get_bc(self.h, _dp(b))
        return dict(zip(("v0bc", "v0m1bc", "eta0bc", "eta0m1bc", "vnbc", "vnp1bc", "etanbc", "etanp1bc"), b))

    def d0mat(self):
        d = np.zeros((self.ny + 1, 5))
        self._L.orc_get_d0mat(self.h, _dp(d))
        return d

    # the step
    def convolutions(self, iy, compute_cfl=False):
        out = np.zeros((6, self.nx + 1, 2 * self.nz + 1), dtype=np.complex128)
        self._L.orc_convolutions(self.h, int(iy), int(compute_cfl), _dp(out))
        return out

    def buildrhs(self, ode, compute_cfl=False):
        o = np.ascontiguousarray(ode, dtype=np.float64)
        self._L.orc_buildrhs(self.h, _dp(o), int(compute_cfl))

    def linsolve(self, lam):
        self._L.orc_linsolve(self.h, float(lam))

    def initial(self):
        self._L.orc_initial(self.h)

    def outstats(self):
        r = np.zeros(11)
        self._L.orc_outstats(self.h, _dp(r))
        return r

    def step(self):
        r = np.zeros(11)
        self._L.orc_step(self.h, _dp(r))
        return r

    def steps(self, n):
        r = np.zeros((n, 11))
        self._L.orc_steps(self.h, int(n), _dp(r))
        return r

    # body force
    MODULES = {"am_f1": 1, "am_butterfly": 2, "coriolis": 3}

    def set_body_force(self, F):
        if F is None:
            self._L.orc_set_body_force(self.h, None)
            return
        F = np.asfortranarray(F, dtype=np.complex128)
        assert F.shape == self.shape
        self._L.orc_set_body_force(self.h, _dp(F))

    def set_body_force_module(self, name, *par):
        """am_f1/am_butterfly(amp, lambdaz_f) or coriolis(omega2, kz_cutoff,
        y_threshold_bot); name None switches the module off."""
        p = np.zeros(3)
        p[:len(par)] = par
        self._L.orc_set_body_force_module(self.h, self.MODULES[name] if name else 0, _dp(p))

    def apply_body_force(self):
        self._L.orc_apply_body_force(self.h)

    def body_force_walls(self):
        self._L.orc_body_force_walls(self.h)

    def get_body_force(self):
        F = np.zeros(self.shape, dtype=np.complex128, order="F")
        self._L.orc_get_body_force(self.h, _dp(F))
        return F

    def time_sample(self, nplanes, nix):
        """Seconds per plane of products, per row of rhs, per x-mode column
        of the solve, measured on a bounded sample (modifies the state)."""
        out = np.zeros(3)
        self._L.orc_time_sample(self.h, int(nplanes), int(nix), _dp(out))
        return out

    def yintegr(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64)
        return self._L.orc_yintegr(self.h, _dp(f))

    def complexderiv(self, f0):
        f0 = np.ascontiguousarray(f0, dtype=np.complex128)
        f1 = np.zeros_like(f0)
        self._L.orc_complexderiv(self.h, _dp(f0), _dp(f1))
        return f1


def fft(x, sign):
    x = np.array(x, dtype=np.complex128)
    lib().orc_fft(len(x), int(sign), _dp(x))
    return x


def c2r(X, m):
    X = np.ascontiguousarray(X, dtype=np.complex128)
    assert len(X) == m // 2 + 1
    x = np.zeros(m)
    lib().orc_c2r(int(m), _dp(X), _dp(x))
    return x


def r2c(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    X = np.zeros(len(x) // 2 + 1, dtype=np.complex128)
    lib().orc_r2c(len(x), _dp(x), _dp(X))
    return X


def ludecomp5(A):
    A = np.ascontiguousarray(A, dtype=np.float64).copy()
    lib().orc_ludecomp5(_dp(A))
    return A


def lludiv5(A, b):
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(5)
    lib().orc_lludiv5(_dp(A), _dp(b), _dp(x))
    return x


def lu5decomp(A):
    A = np.ascontiguousarray(A, dtype=np.float64).copy()
    lib().orc_lu5decomp(_dp(A), A.shape[0], 1, 1)
    return A


def lu5div1(A, x):
    A = np.ascontiguousarray(A, dtype=np.float64)
    x = np.ascontiguousarray(x, dtype=np.complex128).copy()
    lib().orc_lu5div1(_dp(A), A.shape[0], _dp(x))
    return x


def lu5div2(A, b):
    A = np.ascontiguousarray(A, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.complex128).copy()
    lib().orc_lu5div2(_dp(A), A.shape[0], _dp(b))
    return b
