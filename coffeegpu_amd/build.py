"""Build recipe for libcoffee_hip.so (gfx950) — explicit hipcc, in-tree."""

from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcoffee_hip.so")

HIP_SOURCES = ["coffee_hip.hip", "coffee_sph.hip", "coffee_gr.hip", "unet.hip"]
HOST_SOURCES = []  # C++ host surface, added when present

# -packed-fp32-ops: keep the compiler from emitting packed FP32 instructions.
# -ffp-contract=off: every fused multiply-add is written out as fma() in the
# source, so each kernel schedule computes bitwise the same result.
HIPCC_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
               "-I" + os.path.join(ROOT, "include"), "-I" + CSRC]


def hipcc() -> str:
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", "hipcc"):
        if c and (os.path.isabs(c) and os.path.exists(c) or not os.path.isabs(c)):
            return c
    return "hipcc"


DRIVERS = [("coffee_drive_f32", "drivers/coffee_drive.cpp", []),
           ("coffee_drive_f64", "drivers/coffee_drive.cpp", ["-DUSE_DOUBLE"])]


def build_drivers(verbose: bool = False):
    """C++ drivers over the host surface (include/coffee/*.h), linked to the
    library with an rpath to this directory."""
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    for name, src, extra in DRIVERS:
        out = os.path.join(bindir, name)
        cmd = [hipcc(), "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include")] + extra + [
            "-o", out + ".tmp", os.path.join(CSRC, src), "-L" + HERE, "-lcoffee_hip",
            "-Wl,-rpath,$ORIGIN/.."]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError(f"hipcc failed for {name}")
        os.replace(out + ".tmp", out)


def build(verbose: bool = False) -> str:
    srcs = [os.path.join(CSRC, s) for s in HIP_SOURCES]
    srcs += [os.path.join(CSRC, s) for s in HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    tmp = OUT + ".tmp"
    cmd = [hipcc()] + HIPCC_FLAGS + ["-shared", "-o", tmp] + srcs
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed for libcoffee_hip.so")
    os.replace(tmp, OUT)
    build_drivers(verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
