"""ctypes binding of the C-ABI in include/channel_hip.h.

Loads channel_amd/libchannel_hip.so (built by channel_amd.build); there is no
fallback: a missing or unloadable library raises.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libchannel_hip.so")

CHAN_FP64 = 0
CHAN_SPLIT = 1
# status codes below zero (include/channel_hip.h)
CHAN_E_ARG = -1
CHAN_E_SIZE = -2
CHAN_E_NOTFINITE = -3
CHAN_E_CPITYPE = -4
CHAN_E_STATE = -5
CHAN_E_NODEV = -6
# rows that chan_mlp_forward_noalloc stages (include/channel_closure.h)
CHAN_MLP_NOALLOC_MAX = 65536

_lib = None


class ChanParams(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int32), ("ny", ctypes.c_int32), ("nz", ctypes.c_int32),
        ("precision", ctypes.c_int32),
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


class ChanScalars(ctypes.Structure):
    _fields_ = [
        ("time", ctypes.c_double), ("deltat", ctypes.c_double),
        ("meanpx", ctypes.c_double), ("meanpz", ctypes.c_double),
        ("corrpx", ctypes.c_double), ("corrpz", ctypes.c_double),
        ("fr", ctypes.c_double * 3),
        ("cfl", ctypes.c_double),
    ]


# every symbol the header declares, with (restype, argtypes)
_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_I32 = ctypes.c_int32
SIGNATURES = {
    "chan_abi_version": (ctypes.c_int, []),
    "chan_last_error": (ctypes.c_char_p, []),
    "chan_create": (ctypes.c_int, [ctypes.POINTER(ChanParams), ctypes.POINTER(_P)]),
    "chan_destroy": (ctypes.c_int, [_P]),
    "chan_sizes": (ctypes.c_int, [_P, ctypes.POINTER(_I32)]),
    "chan_get_setup": (ctypes.c_int, [_P, _D, _D, _D]),
    "chan_setup_host": (ctypes.c_int, [ctypes.POINTER(ChanParams), ctypes.POINTER(_I32), _D, _D, _D, _D, _D]),
    "chan_set_field": (ctypes.c_int, [_P, _D]),
    "chan_get_rhs": (ctypes.c_int, [_P, _D]),
    "chan_get_field": (ctypes.c_int, [_P, _D]),
    "chan_get_scalars": (ctypes.c_int, [_P, ctypes.POINTER(ChanScalars)]),
    "chan_set_scalars": (ctypes.c_int, [_P, ctypes.POINTER(ChanScalars)]),
    "chan_initial": (ctypes.c_int, [_P]),
    "chan_buildrhs": (ctypes.c_int, [_P, _D, _I32]),
    "chan_linsolve": (ctypes.c_int, [_P, ctypes.c_double]),
    "chan_outstats": (ctypes.c_int, [_P, _D]),
    "chan_stats": (ctypes.c_int, [_P, _D, _D]),
    "chan_set_body_force": (ctypes.c_int, [_P, _D]),
    "chan_set_body_force_feedback": (ctypes.c_int, [_P, ctypes.POINTER(_I32), _D]),
    "chan_get_body_force": (ctypes.c_int, [_P, _D]),
    "chan_keep_body_force": (ctypes.c_int, [_P, _I32]),
    "chan_get_applied_body_force": (ctypes.c_int, [_P, _D]),
    "chan_step": (ctypes.c_int, [_P, _D]),
    "chan_steps": (ctypes.c_int, [_P, _I32, _D]),
    "chan_timing": (ctypes.c_int, [_P, _I32]),
    "chan_kernel_times": (ctypes.c_int, [_P, _I32, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int64), _D,
                                         ctypes.POINTER(_I32)]),
    "chan_sync": (ctypes.c_int, [_P]),
    "chan_fft_selftest": (ctypes.c_int, [_I32, _I32, _I32, _D]),
    "chan_copy_selftest": (ctypes.c_int, [ctypes.c_int64, _I32, _D]),
    "chan_fft_bench": (ctypes.c_int, [_I32, _I32, _I32, _I32, _I32, _D, _D]),
    "chan_fft768_fs_test": (ctypes.c_int, [_I32, _I32, _I32, _I32, _D, _D]),
    # include/channel_closure.h
    "chan_mlp_create": (ctypes.c_int, [ctypes.c_int64, ctypes.POINTER(_P)]),
    "chan_mlp_destroy": (ctypes.c_int, [_P]),
    "chan_mlp_set_params": (ctypes.c_int, [_P, _F]),
    "chan_mlp_get_params": (ctypes.c_int, [_P, _F]),
    "chan_mlp_set_data": (ctypes.c_int, [_P, _F, _F, ctypes.c_int64]),
    "chan_mlp_forward": (ctypes.c_int, [_P, _F, ctypes.c_int64, _F]),
    "chan_mlp_forward_noalloc": (ctypes.c_int, [_P, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p]),
    "chan_mlp_grad": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int64, _F, _D]),
    "chan_mlp_train": (ctypes.c_int, [_P, _I32, _I32, ctypes.c_float, _F, _D]),
    "chan_mlp_bench_forward": (ctypes.c_int, [_P, _I32, _D]),
    "chan_les_samples": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "chan_les_features": (ctypes.c_int, [_P, _F]),
    "chan_les_closure_force": (ctypes.c_int, [_P, _P, _F]),
}


def fft768_fs(x, sign, inner=1, reps=1):
    """The x pass's four-step 768-point transform of the rows of x (natural
    order in and out); returns (result, ms per timed launch)."""
    x = np.ascontiguousarray(x, dtype=np.complex128).copy()
    b, n = x.shape
    assert n == 768
    ms = ctypes.c_double()
    check(load().chan_fft768_fs_test(int(sign), b, int(inner), int(reps), dptr(x), ctypes.byref(ms)))
    return x, ms.value


def fft_bench(x, sign, variant, reps=10):
    """Timed batched transform of the rows of x with a build variant of the
    device FFT; returns (result, ms per launch)."""
    x = np.ascontiguousarray(x, dtype=np.complex128).copy()
    b, n = x.shape
    ms = ctypes.c_double()
    check(load().chan_fft_bench(n, int(sign), b, int(variant), int(reps), dptr(x), ctypes.byref(ms)))
    return x, ms.value


class ChanError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("channel_hip error %d: %s" % (code, msg))
        self.code = code


def load():
    """Load the HIP library (no fallback: raises when it is missing)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIBPATH):
            raise ChanError(-6, "%s is not built (python -m channel_amd.build)" % LIBPATH)
        L = ctypes.CDLL(LIBPATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise ChanError(rc, load().chan_last_error().decode())
    return rc


def dptr(a):
    return a.ctypes.data_as(_D)


def fptr(a):
    return a.ctypes.data_as(_F)


def params_from(cfg, precision=CHAN_FP64):
    p = ChanParams()
    for name, _ in ChanParams._fields_:
        if name == "precision":
            p.precision = precision
        elif name in ("nx", "ny", "nz", "cpi", "cpi_type"):
            setattr(p, name, int(getattr(cfg, name)))
        else:
            setattr(p, name, float(getattr(cfg, name)))
    return p


def setup_host(cfg):
    """Set-up products of chan_setup_host (no device needed)."""
    p = params_from(cfg)
    ny = cfg.ny
    sizes = (ctypes.c_int32 * 5)()
    y = np.zeros(ny + 3)
    der = np.zeros((ny - 1, 4, 5))
    bnd = np.zeros((10, 5))
    bc = np.zeros((8, 5))
    d0 = np.zeros((ny + 1, 5))
    check(load().chan_setup_host(ctypes.byref(p), sizes, dptr(y), dptr(der), dptr(bnd), dptr(bc), dptr(d0)))
    return dict(sizes=list(sizes), y=y, der=der, bnd=bnd, bc=bc, d0fac=d0)


def fft_selftest(x, sign):
    """Batched c2c FFT of the rows of x on the device (diagnostic)."""
    x = np.ascontiguousarray(x, dtype=np.complex128).copy()
    b, n = x.shape
    check(load().chan_fft_selftest(n, int(sign), b, dptr(x)))
    return x
