"""Failure behaviour of the C-ABI on the device: calls out of order, a field
that holds NaN (the CFL maximum then is not finite, which the reference
would carry into its time step; the library stops with CHAN_E_NOTFINITE),
and transform sizes outside the supported fftFIT set.  Every case ends with
an error code and a message, never a fault."""
import ctypes

import numpy as np
import pytest

from cases import seeded_field, tiny
from channel_amd import lib

pytestmark = pytest.mark.gpu


def channel(cfg, **kw):
    from channel_amd.solver import Channel
    return Channel(cfg, **kw)


def test_calls_before_set_field():
    g = channel(tiny())
    for call, args in ((g._L.chan_buildrhs, (g.h, lib.dptr(np.array([3.75, 2.0, 0.0])), 0)),
                       (g._L.chan_linsolve, (g.h, 1.0)),
                       (g._L.chan_initial, (g.h,))):
        rc = call(*args)
        assert rc == lib.CHAN_E_STATE, rc
        assert b"before set_field" in g._L.chan_last_error()
    g.close()


def test_nan_field_stops_with_notfinite():
    cfg = tiny()
    V = seeded_field(cfg)
    V[10, cfg.nz + 1, 2, 0] = np.nan
    g = channel(cfg)
    g.set_field(V)
    with pytest.raises(lib.ChanError) as e:
        g.step()
    assert e.value.code == lib.CHAN_E_NOTFINITE
    assert "not finite" in str(e.value)
    g.close()


def test_unsupported_transform_size():
    # nz = 1023 needs nzd = 3072, beyond the 1536-point transforms
    p = lib.params_from(tiny(nz=1023, nx=3, ny=8))
    h = ctypes.c_void_p()
    rc = lib.load().chan_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == lib.CHAN_E_SIZE and not h.value
    assert b"not supported" in lib.load().chan_last_error()


def test_nx_beyond_the_x_pass():
    # nx = 2: fftFIT keeps nxd = 4, and 3(nx+1) = 9 > 2nxd = 8
    p = lib.params_from(tiny(nx=2, ny=8, nz=3))
    h = ctypes.c_void_p()
    rc = lib.load().chan_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == lib.CHAN_E_SIZE and not h.value
    msg = lib.load().chan_last_error()
    assert b"3(nx+1) <= 2nxd" in msg and b"nx=1 or 3" in msg
