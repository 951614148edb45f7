"""Sanitizer build of the library's host set-up: channel_amd/csrc/setup.cpp
compiled by g++ with AddressSanitizer and UndefinedBehaviorSanitizer (every
report fatal) and run over the workloads' sizes and the edge sizes; its
outputs must equal those of the shipped library's chan_setup_host, and its
fftFIT those of a brute-force search (ffts.f90:78-86: whether i is 2^k or
3*2^k)."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from cases import config1, config2, config3, tiny
from channel_amd import lib

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "channel_amd", "csrc")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is not installed")
    out = str(tmp_path_factory.mktemp("asan") / "sanitize_setup")
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-I" + CSRC, os.path.join(CSRC, "setup.cpp"),
           os.path.join(HERE, "sanitize_setup.cpp"), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def run_driver(driver, tmp_path, cfg):
    f = str(tmp_path / "out.bin")
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:halt_on_error=1", UBSAN_OPTIONS="halt_on_error=1")
    # the sanitizer runtime must come first in the process: drop any preload
    env.pop("LD_PRELOAD", None)
    args = [cfg.nx, cfg.ny, cfg.nz, cfg.alfa0, cfg.beta0, cfg.a, cfg.ymin, cfg.ymax]
    r = subprocess.run([driver, f] + [repr(float(v)) if isinstance(v, float) else str(v) for v in args],
                       capture_output=True, text=True, env=env)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr and "ERROR: AddressSanitizer" not in r.stderr, r.stderr[-4000:]
    return np.fromfile(f, dtype=np.float64)


def fftfit_brute(i):
    forms = {1 << k for k in range(14)} | {3 << k for k in range(13)}
    return int(i in forms)


CASES = {
    "tiny": tiny(),
    "c1": config1(),
    "c2": config2(),
    "c3": config3(),
    "c4": config1(nx=511, ny=512, nz=511, alfa0=0.5, beta0=1.0),
    "ny8": tiny(nx=1, ny=8, nz=1),
    "odd_ny": tiny(ny=33, a=2.5),
    "shifted_walls": tiny(ymin=-1.0, ymax=1.0, a=0.8),
}


@pytest.mark.parametrize("name", sorted(CASES))
def test_sanitized_setup_matches_library(driver, tmp_path, name):
    cfg = CASES[name]
    got = run_driver(driver, tmp_path, cfg)
    ny = cfg.ny
    want = lib.setup_host(cfg)
    parts = [("nxd", 1), ("nzd", 1), ("y", ny + 3), ("der", 20 * (ny - 1)), ("bnd", 50), ("bc", 40),
             ("d0fac", 5 * (ny + 1)), ("fftfit", 4097)]
    assert got.size == sum(n for _, n in parts)
    o = 0
    seg = {}
    for k, n in parts:
        seg[k] = got[o:o + n]
        o += n
    assert int(seg["nxd"][0]) == want["sizes"][3] and int(seg["nzd"][0]) == want["sizes"][4]
    for k in ("y", "der", "bnd", "bc", "d0fac"):
        w = want[k].ravel()
        # the same source through another compiler: equal to the last bit
        # unless the compilers order a sum differently
        np.testing.assert_allclose(seg[k], w, rtol=2e-15, atol=2e-15 * np.abs(w).max(), err_msg=k)
    assert [int(v) for v in seg["fftfit"]] == [fftfit_brute(i) for i in range(4097)]
