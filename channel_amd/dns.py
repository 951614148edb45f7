"""Host-side files of channel's time step: dns.in, Dati.cart*.out, Runtimedata.

The formats are those of the reference and are kept unchanged:

* ``dns.in`` -- 12 list-directed lines read by ``read_dnsin``
  (dnsdata.f90:98-125); ``ni`` on line 3 is the Reynolds number and is
  inverted to the viscosity on read (dnsdata.f90:115).
* ``Dati.cart.out`` -- stream file: int32 nx, ny, nz; float64 alfa0, beta0,
  ni (viscosity), a, ymin, ymax, time (68 bytes), then complex128
  V(-1:ny+1, -nz:nz, 0:nx, 1:3) in Fortran order (dnsdata.f90:677-720,
  821-848; README.md:186-216).
* ``Runtimedata`` -- one line per step with 11 columns: time, dudy bottom,
  -dudy top, dwdy bottom, -dwdy top, fr_x, dp/dx, fr_z, dp/dz, CFL, deltat
  (dnsdata.f90:876-879).
"""
import dataclasses
import math
import os
import re
import struct

import numpy as np

HEADER = struct.Struct("<3i7d")  # 68 bytes


def fftfit(i):
    """fftFIT (ffts.f90:78-86): True when i is 2^k or 3*2^k."""
    j = int(i)
    while j % 2 == 0 and j > 0:
        j >>= 1
    return j in (1, 3)


def dealiased_sizes(nx, nz):
    """nxd, nzd of read_dnsin (dnsdata.f90:110-114) with useFFTfit on."""
    nxd = 3 * (nx + 1) // 2
    nzd = 3 * nz
    while not fftfit(nxd):
        nxd += 1
    while not fftfit(nzd):
        nzd += 1
    return nxd, nzd


_NUM = re.compile(r"[dD]")


def _fnum(tok):
    return float(_NUM.sub("e", tok))


def _flog(tok):
    t = tok.strip().strip(".").upper()
    if t.startswith("T"):
        return True
    if t.startswith("F"):
        return False
    raise ValueError("not a Fortran logical: %r" % tok)


@dataclasses.dataclass
class DnsConfig:
    """Contents of dns.in.  ``ni`` holds the viscosity 1/Re as in the
    reference after dnsdata.f90:115."""
    nx: int = 31
    ny: int = 64
    nz: int = 31
    alfa0: float = 1.0
    beta0: float = 2.0
    ni: float = 1.0 / 3000.0
    a: float = 1.5
    ymin: float = 0.0
    ymax: float = 2.0
    cpi: bool = False
    cpi_type: int = 1
    gamma: float = 0.0
    meanpx: float = 0.0
    meanpz: float = 0.0
    meanflowx: float = 0.0
    meanflowz: float = 0.0
    u0: float = 0.0
    uN: float = 0.0
    deltat: float = 0.0
    cflmax: float = 0.0
    time: float = 0.0
    dt_field: float = 0.0
    dt_save: float = 0.0
    t_max: float = 0.0
    time_from_restart: bool = False
    nstep: int = 0
    npy: int = 1

    @property
    def nxd(self):
        return dealiased_sizes(self.nx, self.nz)[0]

    @property
    def nzd(self):
        return dealiased_sizes(self.nx, self.nz)[1]

    @property
    def field_shape(self):
        return (self.ny + 3, 2 * self.nz + 1, self.nx + 1, 3)

    @classmethod
    def read(cls, path="dns.in"):
        """read_dnsin (dnsdata.f90:98-125): each READ consumes one line and
        ignores what follows the values on it (the '!' comments)."""
        with open(path) as f:
            lines = [ln for ln in f.read().splitlines()]
        toks = []
        for ln in lines:
            ln = ln.split("!")[0]
            toks.append(ln.replace(",", " ").split())
        it = iter(toks)

        def take(n):
            t = next(it)
            while len(t) == 0:
                t = next(it)
            if len(t) < n:
                raise ValueError("dns.in line has %d values, expected %d: %r" % (len(t), n, t))
            return t[:n]

        c = cls()
        t = take(3)
        c.nx, c.ny, c.nz = (int(x) for x in t)
        c.alfa0, c.beta0 = (_fnum(x) for x in take(2))
        c.ni = 1.0 / _fnum(take(1)[0])
        c.a, c.ymin, c.ymax = (_fnum(x) for x in take(3))
        t = take(3)
        c.cpi, c.cpi_type, c.gamma = _flog(t[0]), int(t[1]), _fnum(t[2])
        c.meanpx, c.meanpz = (_fnum(x) for x in take(2))
        c.meanflowx, c.meanflowz = (_fnum(x) for x in take(2))
        c.u0, c.uN = (_fnum(x) for x in take(2))
        c.deltat, c.cflmax, c.time = (_fnum(x) for x in take(3))
        t = take(4)
        c.dt_field, c.dt_save, c.t_max, c.time_from_restart = _fnum(t[0]), _fnum(t[1]), _fnum(t[2]), _flog(t[3])
        c.nstep = int(_fnum(take(1)[0]))
        c.npy = int(take(1)[0])
        return c

    def write(self, path="dns.in"):
        """Writes a dns.in that read() (and the reference) read back to the
        same values; Re is written as 1/ni with 17 significant digits."""
        L = lambda b: ".TRUE." if b else ".FALSE."
        g = lambda x: repr(float(x))
        with open(path, "w") as f:
            f.write("%d %d %d      ! nx, ny, nz\n" % (self.nx, self.ny, self.nz))
            f.write("%s %s      ! alfa0 beta0\n" % (g(self.alfa0), g(self.beta0)))
            f.write("%s      ! ni\n" % g(1.0 / self.ni))
            f.write("%s %s %s      ! a, ymin, ymax\n" % (g(self.a), g(self.ymin), g(self.ymax)))
            f.write("%s %d %s      ! CPI, CPItype, gamma\n" % (L(self.cpi), self.cpi_type, g(self.gamma)))
            f.write("%s %s      ! meanpx, meanpz\n" % (g(self.meanpx), g(self.meanpz)))
            f.write("%s %s      ! meanflowx meanflowz\n" % (g(self.meanflowx), g(self.meanflowz)))
            f.write("%s %s      ! u0 uN\n" % (g(self.u0), g(self.uN)))
            f.write("%s %s %s      ! deltat, cflmax, time\n" % (g(self.deltat), g(self.cflmax), g(self.time)))
            f.write("%s %s %s %s      ! dt_field, dt_save, t_max, time_from_restart\n"
                    % (g(self.dt_field), g(self.dt_save), g(self.t_max), L(self.time_from_restart)))
            f.write("%d      ! nstep\n" % self.nstep)
            f.write("%d      ! npy\n" % self.npy)


# ----------------------------------------------------------------- fields --
def read_restart(path, cfg=None):
    """read_restart_file (dnsdata.f90:677-720).  Returns (V, time, header).
    With ``cfg`` given, a header that differs from it raises ValueError (the
    reference STOPs, dnsdata.f90:696-703); ``ni`` is compared exactly."""
    with open(path, "rb") as f:
        hdr = HEADER.unpack(f.read(HEADER.size))
        nx, ny, nz = hdr[:3]
        alfa0, beta0, ni, a, ymin, ymax, time = hdr[3:]
        if cfg is not None:
            mine = (cfg.nx, cfg.ny, cfg.nz, cfg.alfa0, cfg.beta0, cfg.ni, cfg.a, cfg.ymin, cfg.ymax)
            theirs = (nx, ny, nz, alfa0, beta0, ni, a, ymin, ymax)
            if mine != theirs:
                raise ValueError("mismatch in metadata between restart file and dns.in: %r vs %r" % (theirs, mine))
        n = (ny + 3) * (2 * nz + 1) * (nx + 1) * 3
        V = np.fromfile(f, dtype=np.complex128, count=n)
        if V.size != n:
            raise ValueError("restart file %s is truncated" % path)
    V = V.reshape((ny + 3, 2 * nz + 1, nx + 1, 3), order="F")
    return V, time, dict(nx=nx, ny=ny, nz=nz, alfa0=alfa0, beta0=beta0, ni=ni, a=a, ymin=ymin, ymax=ymax)


def save_restart(path, cfg, V, time):
    """save_restart_file (dnsdata.f90:821-848): 68-byte header + V."""
    V = np.asarray(V)
    assert V.shape == cfg.field_shape, (V.shape, cfg.field_shape)
    tmp = path + ".tmp"
    with open(tmp, "wb") as f:
        f.write(HEADER.pack(cfg.nx, cfg.ny, cfg.nz, cfg.alfa0, cfg.beta0, cfg.ni, cfg.a, cfg.ymin, cfg.ymax, time))
        np.asfortranarray(V, dtype=np.complex128).ravel(order="F").tofile(f)
    os.replace(tmp, path)


def runtimedata_line(row):
    """One Runtimedata line (dnsdata.f90:878): 11 list-directed doubles."""
    return " " + " ".join("%25.16E" % v for v in row) + "\n"


def parse_runtimedata(path):
    rows = []
    with open(path) as f:
        for ln in f:
            t = ln.split()
            if len(t) == 11:
                rows.append([_fnum(x) for x in t])
    return np.array(rows)


def y_grid(cfg):
    """y(iy), iy=-1..ny+1 (dnsdata.f90:153).  2.0d0*real(iy)/real(ny) is
    evaluated left to right in double precision."""
    iy = np.arange(-1, cfg.ny + 2, dtype=np.float64)
    r = 2.0 * iy / float(cfg.ny)
    return cfg.ymin + 0.5 * (cfg.ymax - cfg.ymin) * (np.tanh(cfg.a * (r - 1)) / math.tanh(cfg.a) + 1)
