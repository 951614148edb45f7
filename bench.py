"""RK3 steps/s of the channel DNS time step on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp64]
                    [--config c3] [--no-cpu] [--no-split] [--body-force coriolis]

Workload: the production 384-class grid (nx=191, ny=384, nz=189, a=1.5,
alfa0=0.5, beta0=1, Re=12431, CPI type 1 with gamma=0.161436, cflmax=1),
seeded perturbed-laminar start, state resident in HBM.  Times exactly K RK3
steps after W warm-up steps, bracketed by device synchronisation, and prints
one JSON line.  Per-kernel durations come from HIP events that the library
records on its own stream around every launch during the timed steps.
--body-force coriolis adds the reference's Coriolis body force with the
defaults of body_forces/coriolis/coriolis.in (2 Ro = 0.02, every mode and
row), formed from the field at every buildrhs.
With --precision fp64 (the default) the line also carries a "split" object:
the same workload timed the same way with the FP32/FP64 split step, and its
rate over the FP64 one (--no-split leaves it out).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X


def workload(name):
    from channel_amd.dns import DnsConfig
    if name == "c3":
        return "channel 191x384x189 (nx, ny, nz), Re 12431, CPI gamma 0.161436, cflmax 1", DnsConfig(
            nx=191, ny=384, nz=189, alfa0=0.5, beta0=1.0, ni=1.0 / 12431.0, a=1.5, cpi=True, cpi_type=1,
            gamma=0.161436, deltat=0.0, cflmax=1.0)
    if name == "c2":
        return "channel 95x128x95 (nx, ny, nz), Re 3000, constant flow rate, dt 0.004", DnsConfig(
            nx=95, ny=128, nz=95, alfa0=1.0, beta0=2.0, ni=1.0 / 3000.0, a=1.5, meanflowx=4.0 / 3.0, deltat=0.004)
    if name == "c1":
        return "channel 31x64x31 (nx, ny, nz), Re 3000, constant flow rate, dt 0.004", DnsConfig(
            nx=31, ny=64, nz=31, alfa0=1.0, beta0=2.0, ni=1.0 / 3000.0, a=1.5, meanflowx=4.0 / 3.0, deltat=0.004)
    if name == "c4":
        return "channel 511x512x511 (nx, ny, nz), Re 20000, constant flow rate, dt 0.002", DnsConfig(
            nx=511, ny=512, nz=511, alfa0=0.5, beta0=1.0, ni=1.0 / 20000.0, a=1.5, meanflowx=4.0 / 3.0, deltat=0.002)
    raise SystemExit("unknown workload %s" % name)


def algorithmic_bytes(cfg, nxd, nzd, c=16):
    """Algorithmic HBM bytes of one substep per kernel (each launch covers the
    whole volume).  S64: one FP64 spectral plane of one component; Sc, Pc:
    spectral / z-physical planes of the transform intermediates."""
    Nx, Nz, Ny, Nyi = cfg.nx + 1, 2 * cfg.nz + 1, cfg.ny + 3, cfg.ny - 1
    S64 = Nx * Nz * 16
    Sc = Nx * Nz * c
    Pc = Nx * nzd * c
    per = {
        "zinv": Ny * (3 * S64 + 3 * Pc),
        "xpass": Ny * (3 * Pc + 6 * Pc),
        "zfwd": Ny * (6 * Pc + 6 * Sc),
        "rhs": Ny * (6 * Sc + 3 * S64) + Nyi * (4 * S64 + 2 * S64),
        "solve": Nyi * 2 * S64 + Ny * 3 * S64,
    }
    step = 3 * (Ny * (9 * S64 + 12 * Sc + 18 * Pc) + 8 * Nyi * S64)
    assert abs(3 * sum(per.values()) - step) < 1e-6 * step
    return per, step


def cpu_threads():
    """Threads for the oracle: the CPUs this process may run on, capped by
    OMP_NUM_THREADS when that is set (the GPU box sets it to the host share
    that goes with one GPU)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        return min(avail, int(env)), avail
    return avail, avail


def cpu_baseline(cfg, V):
    """The oracle (plain-C restatement of the reference step: OpenMP threads
    over y planes for the products and the rhs and over (kx, kz) pairs for
    the solves, its own mixed-radix FFT in place of FFTW) timed on the host
    cores for one whole RK3 step of the same workload: the initial CFL pass
    sets Delta t, then one step is timed by the wall clock (about 30 s on 16
    cores at 191x384x189)."""
    from oracle import oracle as orc
    threads, avail = cpu_threads()
    orc.set_threads(threads)
    o = orc.Oracle(cfg)
    o.set_field(V)
    o.initial()
    o.outstats()
    t0 = time.perf_counter()
    row = o.step()
    wall = time.perf_counter() - t0
    o.close()
    if not np.all(np.isfinite(row)):
        raise SystemExit("non-finite Runtimedata in the CPU baseline step")
    return {
        "value": 1.0 / wall,
        "unit": "steps/s",
        "cores": threads,
        "kind": "port",
        "sample": "one whole RK3 step (3 substeps, %.1f s wall) of the oracle: the plain-C restatement of the "
                  "reference step with its own mixed-radix FFT, not FFTW; %d OpenMP threads over y planes (products, "
                  "rhs) and (kx,kz) pairs (solves) on %d usable host CPUs (nproc %d)"
                  % (wall, threads, avail, os.cpu_count() or 0),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--precision", default="fp64", choices=["fp64", "split"])
    ap.add_argument("--config", default="c3")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-split", action="store_true", help="no split-step leg beside the FP64 line")
    ap.add_argument("--no-timing", action="store_true", help="skip per-kernel events")
    ap.add_argument("--body-force", default="none", choices=["none", "coriolis"])
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="BYTES",
                    help="only run the library's plain copy kernel on BYTES (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU only: the library has no multi-GPU path (run with --gpus 1)", file=sys.stderr)
        sys.exit(2)

    from channel_amd import lib
    if args.copy_calibrate:
        import ctypes
        ms = ctypes.c_double()
        lib.check(lib.load().chan_copy_selftest(args.copy_calibrate, 3, ctypes.byref(ms)))
        print(json.dumps({"copy_bytes": args.copy_calibrate, "avg_ms": ms.value,
                          "GBps_read_plus_write": 2 * args.copy_calibrate / ms.value / 1e6}), flush=True)
        return
    from channel_amd.initial import perturbed_laminar
    from channel_amd.solver import Channel

    wname, cfg = workload(args.config)
    prec = lib.CHAN_FP64 if args.precision == "fp64" else lib.CHAN_SPLIT
    V = perturbed_laminar(cfg)

    def timed_run(prec):
        """W warm-up steps, then K timed steps between device synchronisations;
        returns (elapsed s, per-kernel times, the context, still open)."""
        ch = Channel(cfg, prec)
        ch.set_field(V)
        if args.body_force == "coriolis":
            from channel_amd import body_force
            ch.set_body_force_feedback(*body_force.coriolis(cfg, 0.02))
        ch.initial()
        ch.outstats()
        for _ in range(args.warmup):
            ch.step()
        if not args.no_timing:
            ch.timing(True)
        ch.sync()
        t0 = time.perf_counter()
        rows = ch.steps(args.steps)
        ch.sync()
        t1 = time.perf_counter()
        ktimes = ch.kernel_times() if not args.no_timing else {}
        ch.timing(False)
        if not np.all(np.isfinite(rows)):
            raise SystemExit("non-finite Runtimedata in the timed steps")
        return t1 - t0, ktimes, ch

    elapsed, ktimes, ch = timed_run(prec)
    sps = args.steps / elapsed

    per, step_bytes = algorithmic_bytes(cfg, ch.nxd, ch.nzd, 16 if prec == lib.CHAN_FP64 else 8)
    # kernel-level roofline of the dominant kernel
    roof = None
    if ktimes:
        groups = {"zinv": ["zinv"], "xpass": ["xpass"], "zfwd": ["zfwd"], "rhs": ["rhs"],
                  "solve": ["solve_ve", "mean", "uw"]}
        best = None
        for g, ks in groups.items():
            tot = sum(ktimes.get(k, (0, 0.0))[1] for k in ks)
            cnt = max(ktimes.get(k, (0, 0.0))[0] for k in ks)
            if cnt and (best is None or tot > best[1]):
                best = (g, tot, cnt)
        g, tot, cnt = best
        avg_s = tot / cnt / 1e3
        ach = per[g] / avg_s
        traffic = None
        tfile = os.path.join(ROOT, "profiles", "traffic_%s_%s.json" % (g, args.config))
        if os.path.exists(tfile):
            traffic = json.load(open(tfile)).get("hbm_bytes_per_launch")
        roof = {"kernel": g, "bound": "hbm", "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                "frac": ach / HBM_PEAK, "traffic": traffic, "algorithmic_bytes": per[g], "avg_ms": avg_s * 1e3,
                "launches": cnt}

    out = {
        "metric": "RK3 steps/s at 191x384x189, FP64 vs FP32/FP64 split; % HBM roofline",
        "value": sps,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if prec == lib.CHAN_FP64 else "f32+f64",
        "data": "synthetic: seeded perturbed-laminar field (splitmix64 seed 20261202)",
        "config": {"workload": wname, "precision": args.precision, "nxd": ch.nxd, "nzd": ch.nzd,
                   "body_force": args.body_force},
        "step_roofline": {"bytes_per_step": step_bytes, "achieved_GBps": step_bytes * sps / 1e9,
                          "frac": step_bytes * sps / HBM_PEAK},
        "roofline": roof,
        "kernel_ms_per_step": {k: v[1] / args.steps for k, v in sorted(ktimes.items())},
    }
    ch.close()
    if prec == lib.CHAN_FP64 and not args.no_split:
        el2, kt2, ch2 = timed_run(lib.CHAN_SPLIT)
        ch2.close()
        sps2 = args.steps / el2
        _, split_bytes = algorithmic_bytes(cfg, ch.nxd, ch.nzd, 8)
        out["split"] = {"value": sps2, "unit": "steps/s", "ms_per_step": 1e3 * el2 / args.steps,
                        "ratio_to_fp64": sps2 / sps, "dtype": "f32+f64",
                        "step_roofline": {"bytes_per_step": split_bytes, "frac": split_bytes * sps2 / HBM_PEAK},
                        "kernel_ms_per_step": {k: v[1] / args.steps for k, v in sorted(kt2.items())}}
    if not args.no_cpu and args.gpus == 1:
        out["cpu_baseline"] = cpu_baseline(cfg, V)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
