"""Build the gfx950 library in-tree: vren_amd/libvren_amd.so from vren_amd/csrc/*.hip.

Every translation unit is compiled by hipcc for gfx950 only. The bit-exact stages depend on the compiler not
contracting a*b+c into fma, so contraction is off everywhere except in the shading kernels (shade.hip), whose
floating-point results are checked against an FP64 oracle with a tolerance instead of bit for bit. Packed FP32
instructions are disabled for all device code.
"""
import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libvren_amd.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

COMMON = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
    "-Wno-pass-failed",
]
# translation units whose float results are tolerance-checked rather than bit-exact
CONTRACT_OK = {"shade.hip"}


def sources():
    return sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))


def _compile(src, verbose):
    obj = os.path.join(OBJ, src[:-4] + ".o")
    # fast-honor-pragmas: fused multiply-adds everywhere except in the `#pragma clang fp contract(off)` regions
    # (plain -ffp-contract=fast lets the backend fuse across those regions as well)
    fp = ["-ffp-contract=fast-honor-pragmas"] if src in CONTRACT_OK else ["-ffp-contract=off"]
    cmd = [HIPCC] + COMMON + fp + ["-c", os.path.join(CSRC, src), "-o", obj]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed on %s:\n%s" % (src, r.stderr[-8000:]))
    warn = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    return obj, warn


def build(verbose=False, jobs=None):
    os.makedirs(OBJ, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(8, os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda s: _compile(s, verbose), srcs))
    for obj, warn in results:
        if warn and verbose:
            print(warn, flush=True)
    objs = [o for o, _ in results]
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs
    if verbose:
        print(" ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
