"""Build recipe for libcoffee_hip.so (gfx950) — explicit hipcc, in-tree.

Each HIP translation unit compiles to its own object (in parallel, under
_build/), then one hipcc link makes the shared library; libcoffee_h5.so (the
HDF5 file layer, host code) is built beside it; the C++ drivers link against
both.
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

HIP_SOURCES = ["coffee_hip.hip", "coffee_sph.hip", "coffee_gr.hip", "unet.hip", "coffee_io.hip"]

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
           ("coffee_drive_f64", "drivers/coffee_drive.cpp", ["-DUSE_DOUBLE"]),
           ("coffee_restart_f32", "drivers/coffee_restart.cpp", ["-lcoffee_h5"]),
           ("coffee_restart_f64", "drivers/coffee_restart.cpp", ["-DUSE_DOUBLE", "-lcoffee_h5"])]


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
        flags = [e for e in extra if not e.startswith("-l")]
        libs = [e for e in extra if e.startswith("-l")]
        cmd = [hipcc(), "-std=c++17", "-O2", "-I" + os.path.join(ROOT, "include")] + flags + [
            "-o", out + ".tmp", os.path.join(CSRC, src), "-L" + HERE, "-lcoffee_hip"] + libs + [
            "-Wl,-rpath,$ORIGIN/.."]
        _run(cmd, name, verbose)
        os.replace(out + ".tmp", out)

    todo = [d for d in DRIVERS if os.path.exists(os.path.join(CSRC, d[1]))
            and ("-lcoffee_h5" not in d[2] or os.path.exists(H5_OUT))]
    with ThreadPoolExecutor(_jobs()) as ex:
        list(ex.map(one, todo))


# HDF5 C library for the snapshot / field-output files (libcoffee_h5.so)
HDF5_ROOT = os.environ.get("HDF5_ROOT", "/opt/conda")
H5_OUT = os.path.join(HERE, "libcoffee_h5.so")


def build_h5(verbose: bool = False):
    """libcoffee_h5.so: the HDF5 file layer (host code, g++), linked to the
    serial HDF5 C library under HDF5_ROOT with a run-time path to it (the C++
    runtime linked in statically and not exported, so that path cannot hand
    an older libstdc++ from the same directory to a C++ host).  When
    the HDF5 headers are absent it is not built, and the exporter raises when
    used."""
    inc = os.path.join(HDF5_ROOT, "include")
    lib = os.path.join(HDF5_ROOT, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")):
        print(f"[build] no hdf5.h under {inc}: libcoffee_h5.so not built", file=sys.stderr)
        return None
    tmp = H5_OUT + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-Wall", "-I" + os.path.join(ROOT, "include"),
          "-I" + inc, "-o", tmp, os.path.join(CSRC, "io", "coffee_h5.cpp"), "-L" + lib, "-lhdf5",
          "-static-libstdc++", "-static-libgcc", "-Wl,--exclude-libs,ALL",
          "-Wl,--enable-new-dtags,-rpath," + lib], "libcoffee_h5.so", verbose)
    os.replace(tmp, H5_OUT)
    return H5_OUT


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
    build_h5(verbose)
    build_drivers(verbose)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
