"""Builds channel_amd/libchannel_hip.so from channel_amd/csrc with hipcc for
gfx950 (cross-compiles without a GPU).

    python -m channel_amd.build [-v]
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libchannel_hip.so")
ARCH = os.environ.get("CHAN_OFFLOAD_ARCH", "gfx950")

SOURCES = ["setup.cpp", "k_fft.hip", "k_rhs.hip", "k_solve.hip", "k_stats.hip", "k_mlp.hip", "k_les.hip", "api.hip"]
HEADERS = ["fft.hpp", "setup.hpp", "chan_internal.hpp", "ctx_internal.hpp", "mlp_internal.hpp"]


def hipcc():
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _flags():
    return [
        "-O3", "-std=c++17", "-fPIC", "--offload-arch=" + ARCH,
        # no packed FP32 instructions anywhere in the device code
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-munsafe-fp-atomics", "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
        "-I" + os.path.join(HERE, "..", "include"),
    ]


def _compile(src, verbose):
    obj = os.path.join(OUT, os.path.splitext(src)[0] + ".o")
    lang = ["-x", "hip"] if src.endswith(".hip") else []
    cmd = [hipcc()] + _flags() + lang + ["-c", os.path.join(SRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile of %s failed:\n%s%s" % (src, r.stdout, r.stderr))
    noise = [ln for ln in (r.stderr or "").splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
    if verbose and noise:
        print("\n".join(noise), flush=True)
    return obj


def build(verbose=False):
    os.makedirs(OUT, exist_ok=True)
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n%s%s" % (r.stdout, r.stderr))
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
