"""Build libmplb.so (HIP kernels for gfx950 + the C ABI) in-tree.

    python -m multiphaselb_amd.build            # compile everything
    python -m multiphaselb_amd.build --check    # compile, then load and list exports

The library lands next to this file (multiphaselb_amd/libmplb.so) so it travels with
the source tree.  Each translation unit is compiled separately, in parallel.
"""
import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libmplb.so")
ARCH = os.environ.get("MPLB_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
CXX = shutil.which("g++") or "g++"

COMMON = ["-std=c++17", "-fPIC", "-I" + CSRC, "-I" + os.path.join(ROOT, "include")]
# Packed FP32 instructions are disabled for all device code.
HIP_DEV = ["--offload-arch=" + ARCH, "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
           "-Wno-unused-command-line-argument"]

# (source, flags): the parity kernels, the batched-ensemble kernels (whose FP64 mode is a
# parity mode) and all host code keep plain IEEE arithmetic (no fused multiply-add
# contraction); the fused FP32/FP16 kernels may contract.
UNITS = [
    ("kernels_generic.hip", ["-O3", "-ffp-contract=off"] + HIP_DEV),
    ("kernels_fused.hip", ["-O3", "-ffp-contract=fast"] + HIP_DEV),
    ("kernels_batch.hip", ["-O3", "-ffp-contract=off"] + HIP_DEV),
    ("mlp.hip", ["-O3", "-ffp-contract=off"] + HIP_DEV),
    ("warm.hip", ["-O3", "-ffp-contract=off"] + HIP_DEV),
    ("fieldio.hip", ["-O2", "-ffp-contract=off"] + HIP_DEV),
    ("capi.hip", ["-O2", "-ffp-contract=off"] + HIP_DEV),
    ("dict.cpp", ["-O2", "-ffp-contract=off"]),
    ("mesh.cpp", ["-O2", "-ffp-contract=off"]),
]


def _compile(unit):
    src, flags = unit
    obj = os.path.join(BUILD, src + ".o")
    path = os.path.join(CSRC, src)
    if src.endswith(".hip"):
        cmd = [HIPCC, "-c", "-x", "hip", path, "-o", obj] + COMMON + flags
    else:
        cmd = [CXX, "-c", path, "-o", obj] + COMMON + flags
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = (r.stdout + r.stderr).strip()
    # hipcc's host pass does not know the device-only feature flag; that note is harmless
    lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l]
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), "\n".join(lines)))
    return obj, "\n".join(lines)


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    jobs = min(len(UNITS), max(1, int(os.environ.get("MAX_JOBS", "8"))))
    with cf.ThreadPoolExecutor(jobs) as ex:
        results = list(ex.map(_compile, UNITS))
    for obj, msg in results:
        if verbose and msg:
            print(msg)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stdout + r.stderr))
    os.replace(tmp, LIB)
    _build_probe()
    return LIB


# The memory-pattern probe (csrc/stream_probe.hip): a stand-alone program that bench.py
# runs for the streaming ceilings of the single pass's access pattern.
PROBE = os.path.join(HERE, "stream_probe")


def _build_probe():
    cmd = [HIPCC, "-x", "hip", os.path.join(CSRC, "stream_probe.hip"), "-o", PROBE + ".tmp", "-O3", "-std=c++17"] + HIP_DEV
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("probe build failed: %s\n%s" % (" ".join(cmd), r.stdout + r.stderr))
    os.replace(PROBE + ".tmp", PROBE)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--check", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    lib = build(verbose=a.verbose)
    print("built", lib)
    if a.check:
        out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True).stdout
        print("\n".join(l for l in out.splitlines() if " mplb_" in l))


if __name__ == "__main__":
    sys.exit(main())
