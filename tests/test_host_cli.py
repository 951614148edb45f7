"""The Fortran host on the CPU: its run-time selections (CHANNEL_BODYFORCE,
CHANNEL_PRECISION) are checked before any GPU call, so an unknown value stops
the program with an error that names the variable; its body-force module
tables (host_io.f90) equal those of channel_amd.body_force."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from cases import config1, config2, tiny
from channel_amd import body_force as bf
from channel_amd.dns import y_grid

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "channel_amd", "host", "channel")


@pytest.mark.skipif(not os.path.exists(EXE), reason="Fortran host not built (__graft_entry__.build())")
@pytest.mark.parametrize("var,val,msg", [
    ("CHANNEL_BODYFORCE", "am_f2", "must be am_f1, am_butterfly, coriolis or none"),
    ("CHANNEL_PRECISION", "fp16", "must be fp64 or split"),
])
def test_host_rejects_unknown_selection(tmp_path, var, val, msg):
    cfg = config1()
    cfg.t_max = cfg.deltat
    cfg.nstep = 1
    cfg.write(str(tmp_path / "dns.in"))
    env = dict(os.environ, **{var: val})
    r = subprocess.run([EXE], cwd=str(tmp_path), capture_output=True, text=True, timeout=60,
                       stdin=subprocess.DEVNULL, env=env)
    assert r.returncode != 0
    assert "ERROR: %s %s, not %s" % (var, msg, val) in r.stdout, r.stdout
    assert not (tmp_path / "Runtimedata").exists()


FLANG = shutil.which("flang") or ("/opt/rocm/lib/llvm/bin/flang" if os.path.exists("/opt/rocm/lib/llvm/bin/flang") else None)
LIBSO = os.path.join(ROOT, "channel_amd", "libchannel_hip.so")
HOST = os.path.join(ROOT, "channel_amd", "host")

# writes the host's table of the module named on the command line for the
# grid in in.bin (nx, ny, nz as int32; beta0, ymax, y(-1:ny+1) as float64)
DRIVER = """
PROGRAM tables
  USE, intrinsic :: iso_c_binding
  USE host_io
  IMPLICIT NONE
  real(C_DOUBLE), allocatable :: y(:), gain(:,:,:)
  integer(C_INT32_T) :: src(3)
  character(len=32) :: arg
  CALL GET_COMMAND_ARGUMENT(1, arg)
  bf_name = TRIM(arg)
  OPEN(20, file='in.bin', access='stream', status='old')
  READ(20) nx, ny, nz
  READ(20) beta0, ymax
  ALLOCATE(y(-1:ny+1), gain(0:2*nz+1, -1:ny+1, 0:2))
  READ(20) y
  CLOSE(20)
  CALL body_force_table(y, src, gain)
  OPEN(21, file='out.bin', access='stream', status='replace')
  WRITE(21) src, gain
  CLOSE(21)
END PROGRAM
"""


@pytest.mark.skipif(FLANG is None or not os.path.exists(LIBSO), reason="flang or the library is missing")
def test_host_force_tables_match_python(tmp_path):
    """The host's Fortran tables of am_f1, am_butterfly and coriolis equal
    channel_amd.body_force's (pinned against the oracle's literal module
    restatements) bit for bit, on four grids."""
    (tmp_path / "tables.f90").write_text(DRIVER)
    subprocess.check_call([FLANG, "-O2", "-J", str(tmp_path), os.path.join(HOST, "chanlib.f90"),
                           os.path.join(HOST, "host_io.f90"), str(tmp_path / "tables.f90"), "-o",
                           str(tmp_path / "tables"), "-L" + os.path.dirname(LIBSO), "-lchannel_hip",
                           "-Wl,-rpath," + os.path.dirname(LIBSO)], cwd=str(tmp_path))
    (tmp_path / "coriolis.in").write_text("0.3     ! 2*Ro\n5.0\n0.4\n")
    cor = bf.read_coriolis_in(str(tmp_path / "coriolis.in"))
    python = {"am_f1": lambda c: bf.am_f1(c), "am_butterfly": lambda c: bf.am_butterfly(c),
              "coriolis": lambda c: bf.coriolis(c, **cor)}
    for cfg in (tiny(nz=7, beta0=2.0), config1(), config2(), tiny(nz=15, beta0=0.5, ny=64)):
        y = y_grid(cfg)
        with open(tmp_path / "in.bin", "wb") as f:
            np.array([cfg.nx, cfg.ny, cfg.nz], np.int32).tofile(f)
            np.array([cfg.beta0, cfg.ymax]).tofile(f)
            y.tofile(f)
        for name, table in python.items():
            subprocess.run([str(tmp_path / "tables"), name], cwd=str(tmp_path), check=True, capture_output=True)
            raw = (tmp_path / "out.bin").read_bytes()
            src = np.frombuffer(raw[:12], np.int32)
            gain = np.frombuffer(raw[12:], np.float64).reshape(3, cfg.ny + 3, 2 * cfg.nz + 2)
            s2, g2 = table(cfg)
            assert tuple(src) == tuple(s2), (name, src, s2)
            np.testing.assert_array_equal(gain, g2, err_msg="%s %s" % (name, (cfg.nx, cfg.ny, cfg.nz, cfg.beta0)))
            assert np.count_nonzero(gain) > 0
