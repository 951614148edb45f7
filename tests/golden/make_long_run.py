"""Writes long_run_c1.npz (or long_run_c2.npz): the oracle's 31x64x31 (or
95x128x95) FP64 run from the seeded perturbed-laminar field (seed
20261202), 2000 RK3 steps of deltat 0.004.

Contents (float64): every 10th Runtimedata row and the last one ('rows',
'row_steps'), and at steps 200 and 2000 the statistics profiles of rows
0..ny ('mean_200', 'var_200', 'uv_200', 'mean_2000', ...) as the spectral
sums of postpro/history/var_history.f90:36-87 (weight 1 at kx = 0, 2
otherwise; the field's own plane mean subtracted from the (0,0) mode), and
the per-row energy of each component sum |V|^2 ('energy_2000').

Run from the repository root: python tests/golden/make_long_run.py [c1|c2]
[threads] (c1: about 10 minutes on 8 cores; c2: about 3 hours on 6).  The c2
record also holds the profiles after 1000 steps, and the file is rewritten
at every profile step, so that a run cut short keeps what it reached
('last_step').
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from cases import config1, config2, seeded_field  # noqa: E402
from oracle import oracle as orc  # noqa: E402


def profiles(V, nz):
    """mean (3, ny+1), var (3, ny+1), uv (ny+1,) of rows 0..ny."""
    P = V[1:-1]  # rows 0..ny
    m = P[:, nz, 0, :].real  # (ny+1, 3)
    f = np.full(P.shape[2], 2.0)
    f[0] = 1.0
    Q = P.copy()
    Q[:, nz, 0, :] -= m
    var = np.einsum("x,yzxc->cy", f, np.abs(Q) ** 2)
    uv = np.einsum("x,yzx->y", f, (Q[..., 0] * np.conj(Q[..., 1])).real)
    return m.T.copy(), var, uv


def main():
    name = sys.argv[1] if len(sys.argv) > 1 else "c1"
    threads = int(sys.argv[2]) if len(sys.argv) > 2 else (os.cpu_count() or 1)
    orc.set_threads(threads)
    cfg = config1() if name == "c1" else config2()
    at = (200, 2000) if name == "c1" else (200, 1000, 2000)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "long_run_%s.npz" % name)
    o = orc.Oracle(cfg)
    o.set_field(seeded_field(cfg))
    rows, steps, out = [], [], {}
    t0 = time.time()
    for s in range(1, 2001):
        r = o.step()
        if s % 10 == 0 or s == 1:
            rows.append(r)
            steps.append(s)
        if s in at:
            V = o.get_field()
            m, v, uv = profiles(V, cfg.nz)
            out["mean_%d" % s], out["var_%d" % s], out["uv_%d" % s] = m, v, uv
            if s == 2000:
                out["energy_2000"] = np.sum(np.abs(V) ** 2, axis=(1, 2)).T.copy()
            print("step", s, "%.0f s" % (time.time() - t0), flush=True)
            out["rows"] = np.array(rows)
            out["row_steps"] = np.array(steps)
            if name != "c1":
                out["last_step"] = np.array(s)
            np.savez_compressed(path, **out)


if __name__ == "__main__":
    main()
