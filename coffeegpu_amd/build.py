"""Build recipe for libcoffee_hip.so (gfx950) — explicit hipcc, in-tree.

Each HIP translation unit compiles to its own object (in parallel, under
_build/), then one hipcc link makes the shared library; the C++ drivers link
against it.
"""

from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libcoffee_hip.so")
OBJDIR = os.path.join(ROOT, "_build")

HIP_SOURCES = ["coffee_hip.hip", "coffee_sph.hip", "coffee_gr.hip", "unet.hip"]

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


# (name, source, extra flags): C++ drivers over the host surface
DRIVERS = [("coffee_drive_f32", "drivers/coffee_drive.cpp", []),
           ("coffee_drive_f64", "drivers/coffee_drive.cpp", ["-DUSE_DOUBLE"])]


def _run(cmd, what: str, verbose: bool):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed for {what}")


def _jobs() -> int:
    n = os.cpu_count() or 4
    return max(1, min(16, n))


def build_drivers(verbose: bool = False):
    """C++ drivers over the host surface (include/coffee/*.h), linked to the
    library with an rpath to this directory."""
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)

    def one(d):
        name, src, extra = d
        out = os.path.join(bindir, name)
        cmd = [hipcc(), "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include")] + extra + [
            "-o", out + ".tmp", os.path.join(CSRC, src), "-L" + HERE, "-lcoffee_hip",
            "-Wl,-rpath,$ORIGIN/.."]
        _run(cmd, name, verbose)
        os.replace(out + ".tmp", out)

    todo = [d for d in DRIVERS if os.path.exists(os.path.join(CSRC, d[1]))]
    with ThreadPoolExecutor(_jobs()) as ex:
        list(ex.map(one, todo))


def build(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)

    def compile_one(src):
        obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
        _run([hipcc()] + HIPCC_FLAGS + ["-c", "-o", obj + ".tmp", os.path.join(CSRC, src)], src,
             verbose)
        os.replace(obj + ".tmp", obj)
        return obj

    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    # the slowest translation units first
    srcs.sort(key=lambda s: -os.path.getsize(os.path.join(CSRC, s)))
    with ThreadPoolExecutor(_jobs()) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = OUT + ".tmp"
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, "link", verbose)
    os.replace(tmp, OUT)
    build_drivers(verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
