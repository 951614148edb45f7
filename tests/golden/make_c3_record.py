"""Writes c3_record.npz: a fingerprint of the oracle's 191x384x189 FP64 run
from the benchmark's seeded field (seed 20261202): the initial CFL pass
(which sets Delta t = cflmax/CFL), then 10 RK3 steps with CPI type 1.

Contents: 'rows' -- the Runtimedata row of the initial pass and of steps
1..10 (11 x 11); 'idx' -- (n, 4) entries (row, iz + nz, ix, component) of
the field in the reference layout: every row of the modes |kx|, |kz| <= 2
(index ix <= 2, |iz| <= 2) and 4096 entries drawn uniformly (numpy seed 5);
'v1', 'v10' -- those entries after steps 1 and 10 (complex128).

With a step count above 10 (c3_record_100.npz) the run goes on to that
step, 'rows' holds every row reached, 'v20', 'v50', 'v100' the entries after
those steps, and the file is rewritten at each of them with 'last_step', so
that a run cut short keeps what it reached.

The oracle needs about 27 s per step on 16 cores; run from the repository
root: python tests/golden/make_c3_record.py [threads] [out.npz] [steps]
[checkpoint-dir].  With a checkpoint directory the field and the scalars are
written there every 10 steps, and a run given the same directory resumes
from its last checkpoint (a restart from the field and the scalars
continues the run bit for bit), with any thread count.
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from cases import config3, seeded_field  # noqa: E402


def sample_index(cfg):
    nz = cfg.nz
    Ny, Nz, Nx = cfg.ny + 3, 2 * cfg.nz + 1, cfg.nx + 1
    rows = np.arange(Ny)
    low = [(r, nz + iz, ix, c) for r in rows for ix in range(3) for iz in range(-2, 3) for c in range(3)]
    rng = np.random.default_rng(5)
    rnd = np.stack([rng.integers(0, Ny, 4096), rng.integers(0, Nz, 4096), rng.integers(0, Nx, 4096),
                    rng.integers(0, 3, 4096)], axis=1)
    return np.concatenate([np.array(low), rnd]).astype(np.int64)


def take(V, idx):
    return V[idx[:, 0], idx[:, 1], idx[:, 2], idx[:, 3]]


def main():
    from oracle import oracle as orc
    threads = int(sys.argv[1]) if len(sys.argv) > 1 else (os.cpu_count() or 1)
    out = sys.argv[2] if len(sys.argv) > 2 else os.path.join(os.path.dirname(os.path.abspath(__file__)), "c3_record.npz")
    nsteps = int(sys.argv[3]) if len(sys.argv) > 3 else 10
    ckpt = sys.argv[4] if len(sys.argv) > 4 else None
    orc.set_threads(threads)
    cfg = config3()
    idx = sample_index(cfg)
    o = orc.Oracle(cfg)
    first = 1
    if ckpt and os.path.exists(os.path.join(ckpt, "state.npz")):
        st = np.load(os.path.join(ckpt, "state.npz"))
        o.set_field(np.load(os.path.join(ckpt, "field_%d.npy" % int(st["step"]))))
        o.set_scalars(**dict(zip(orc.Oracle.SCALARS, st["scalars"])))
        rows = list(st["rows"])
        res = {k: st[k] for k in st.files if k.startswith("v")}
        first = int(st["step"]) + 1
        print("resumed after step", first - 1, flush=True)
    else:
        o.set_field(seeded_field(cfg))
        o.initial()
        rows = [o.outstats()]
        res = {}
    t0 = time.time()
    marks = (1, 10) if nsteps <= 10 else (1, 10, 20, 50, 100, nsteps)
    for s in range(first, nsteps + 1):
        rows.append(o.step())
        if s in marks:
            res["v%d" % s] = take(o.get_field(), idx)
            if nsteps > 10:
                np.savez_compressed(out, rows=np.array(rows), idx=idx.astype(np.int16), last_step=np.array(s), **res)
        if ckpt and s % 10 == 0:
            # the field goes to a file of its own step and the state, which
            # names it, is renamed into place last: a run stopped mid-write
            # keeps the previous checkpoint whole
            os.makedirs(ckpt, exist_ok=True)
            np.save(os.path.join(ckpt, "field_%d.npy" % s), o.get_field())
            sc = o.scalars()
            np.savez(os.path.join(ckpt, "state.next.npz"), scalars=np.array([sc[k] for k in orc.Oracle.SCALARS]),
                     rows=np.array(rows), step=np.array(s), **res)
            os.replace(os.path.join(ckpt, "state.next.npz"), os.path.join(ckpt, "state.npz"))
            if os.path.exists(os.path.join(ckpt, "field_%d.npy" % (s - 10))):
                os.remove(os.path.join(ckpt, "field_%d.npy" % (s - 10)))
        print("step", s, "%.0f s" % (time.time() - t0), flush=True)
    if nsteps <= 10:
        np.savez_compressed(out, rows=np.array(rows), idx=idx.astype(np.int16), **res)


if __name__ == "__main__":
    main()
