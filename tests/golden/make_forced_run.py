"""Writes forced_run_c1.npz: the oracle's 31x64x31 FP64 run from the seeded
perturbed-laminar field (seed 20261202) with the Coriolis body-force module
(body_forces/coriolis: 2 Ro = 0.5, kz_cutoff = 9, y_threshold = 0.4, so the
force acts on |iz| <= 4 and on the rows within 0.4 of a wall), 500 RK3 steps
of deltat 0.004.

Contents (float64): every 10th Runtimedata row and the first ('rows',
'row_steps'), and at steps 100 and 500 the statistics profiles of rows
0..ny ('mean_100', 'var_100', 'uv_100', 'mean_500', ...) computed as in
make_long_run.py.

Run from the repository root: python tests/golden/make_forced_run.py [threads]
(about 3 minutes on 8 cores).
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))
sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))

from cases import config1, seeded_field  # noqa: E402
from make_long_run import profiles  # noqa: E402
from oracle import oracle as orc  # noqa: E402

CORIOLIS = (0.5, 9.0, 0.4)  # omega2 = 2 Ro, kz_cutoff, y_threshold
NSTEPS = 500


def main():
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1)
    orc.set_threads(threads)
    cfg = config1()
    o = orc.Oracle(cfg)
    o.set_field(seeded_field(cfg))
    o.set_body_force_module("coriolis", *CORIOLIS)
    rows, steps, out = [], [], {}
    t0 = time.time()
    for s in range(1, NSTEPS + 1):
        r = o.step()
        if s % 10 == 0 or s == 1:
            rows.append(r)
            steps.append(s)
        if s in (100, NSTEPS):
            m, v, uv = profiles(o.get_field(), cfg.nz)
            out["mean_%d" % s], out["var_%d" % s], out["uv_%d" % s] = m, v, uv
            print("step", s, "%.0f s" % (time.time() - t0), flush=True)
    out["rows"] = np.array(rows)
    out["row_steps"] = np.array(steps)
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "forced_run_c1.npz"), **out)


if __name__ == "__main__":
    main()
