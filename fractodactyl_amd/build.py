"""Builds libfdg.so (HIP kernels for gfx950 + the C++ host + the C ABI) in-tree.

    python -m fractodactyl_amd.build

-ffp-contract=off keeps every floating-point expression in the reference's operation
order on host and device; the packed-fp32 target feature is switched off for device code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libfdg.so")
SOURCES = ["host.cpp", "jetaa.cpp", "atlas.cpp", "certify.cpp", "api.cpp", "film.cpp", "perturb_f64.hip", "perturb_fe32.hip", "encode.hip", "noalloc.hip"]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-I" + os.path.join(os.path.dirname(HERE), "include"), "-O3", "-std=c++17", "-fPIC", "-shared",
         "-ffp-contract=off", "-fno-fast-math",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-Wno-unused-command-line-argument"]


def build(verbose=False):
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = OUT + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-o", tmp] + srcs + ["-lpthread"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libfdg.so")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
