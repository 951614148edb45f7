"""The C-ABI library without a GPU: it loads, exports every entry point that
include/channel_hip.h declares, and its host-side set-up (grid, compact
stencils, boundary rows, factorised D0 band) matches the oracle's."""
import ctypes
import os
import re

import numpy as np
import pytest

from cases import config1, config3, tiny
from channel_amd import lib
from channel_amd.dns import DnsConfig
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    import glob
    txt = "".join(open(f).read() for f in sorted(glob.glob(os.path.join(ROOT, "include", "*.h"))))
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(chan_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = lib.load()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(lib.SIGNATURES), set(syms) ^ set(lib.SIGNATURES)
    assert L.chan_abi_version() == 1


def test_struct_layouts_match_header(tmp_path):
    # the ctypes mirrors must have the offsets the C compiler gives the header
    # structs (the Fortran host's bind(C) types follow the same order)
    import shutil
    import subprocess
    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no C compiler")
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "channel_hip.h"', "int main(void){"]
    for st, cls in (("chan_params", lib.ChanParams), ("chan_scalars", lib.ChanScalars)):
        lines.append('printf("%%s %%zu\\n", "%s", sizeof(%s));' % (st, st))
        for name, _ in cls._fields_:
            lines.append('printf("%%s.%%s %%zu\\n", "%s", "%s", offsetof(%s, %s));' % (st, name, st, name))
    lines.append("return 0;}")
    src = tmp_path / "probe.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "probe"
    subprocess.check_call([cc, "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)])
    got = dict(ln.split() for ln in subprocess.check_output([str(exe)]).decode().splitlines())
    for st, cls in (("chan_params", lib.ChanParams), ("chan_scalars", lib.ChanScalars)):
        assert int(got[st]) == ctypes.sizeof(cls)
        for name, _ in cls._fields_:
            assert int(got["%s.%s" % (st, name)]) == getattr(cls, name).offset, (st, name)


@pytest.mark.parametrize("mk", [tiny, config1, config3])
def test_host_setup_matches_oracle(mk):
    cfg = mk()
    s = lib.setup_host(cfg)
    o = orc.Oracle(cfg)
    assert s["sizes"] == [cfg.nx, cfg.ny, cfg.nz, o.nxd, o.nzd]
    y, _ = o.grid()
    np.testing.assert_array_equal(s["y"], y)
    np.testing.assert_array_equal(s["der"], o.der())
    b = o.bnd()
    np.testing.assert_array_equal(s["bnd"], np.array([b[k] for k in ("d040", "d140", "d240", "d14m1", "d24m1", "d04n", "d14n", "d24n", "d14np1", "d24np1")]))
    c = o.bc()
    np.testing.assert_array_equal(s["bc"], np.array([c[k] for k in ("v0bc", "v0m1bc", "eta0bc", "eta0m1bc", "vnbc", "vnp1bc", "etanbc", "etanp1bc")]))
    np.testing.assert_array_equal(s["d0fac"], o.d0mat())


def test_invalid_arguments_fail_loudly():
    cfg = tiny(ny=4)
    with pytest.raises(lib.ChanError) as e:
        lib.setup_host(cfg)
    assert e.value.code == -1


def test_unknown_precision_mode_is_rejected():
    import ctypes
    p = lib.params_from(tiny(), precision=5)
    h = ctypes.c_void_p()
    rc = lib.load().chan_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == -1 and not h.value
    assert b"precision" in lib.load().chan_last_error()


def test_wrong_cpi_type_is_the_reference_stop():
    """CPI with CPI_type outside {0, 1} stops the reference ("Wrong selection
    of CPI_Type", linsolve_blocking.inc:93-95); chan_create refuses it."""
    import ctypes
    p = lib.params_from(tiny(cpi=True, cpi_type=2))
    h = ctypes.c_void_p()
    rc = lib.load().chan_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == lib.CHAN_E_CPITYPE and not h.value
    assert b"CPI_Type" in lib.load().chan_last_error()
