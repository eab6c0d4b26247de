#!/usr/bin/env python3
"""Build tests/cpp/drop_in_test (the C++ drop-in classes' test driver) against vren_amd/libvren_amd.so."""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")


def build(verbose=True):
    out = os.path.join(HERE, "drop_in_test")
    cmd = ["g++", "-std=c++17", "-O2", "-Wall", "-Wextra", "-D__HIP_PLATFORM_AMD__",
           "-I" + os.path.join(ROOT, "include", "vren"), "-I" + os.path.join(ROCM, "include"),
           os.path.join(HERE, "drop_in_test.cpp"), "-o", out,
           "-L" + os.path.join(ROOT, "vren_amd"), "-lvren_amd", "-L" + os.path.join(ROCM, "lib"), "-lamdhip64",
           "-Wl,-rpath,$ORIGIN/../../vren_amd", "-Wl,-rpath," + os.path.join(ROCM, "lib")]
    if verbose:
        print(" ".join(cmd))
    subprocess.run(cmd, check=True)
    return out


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
