#!/usr/bin/env python3
"""Benchmark of the flat EZ field-update hot path on one MI355X.

Workload: the 3D Alfven-wave collision — 512^3 periodic cells (guard 3), FP32,
background B0 = z, two counter-propagating Gaussian Alfven packets with seeded
amplitudes, divergence cleaning, E.B / E<B projections and 6th-order
Kreiss-Oliger on.  A step is one full time step (5 LSRK stages + KO +
projections + guard fill) of the whole grid, inputs resident in HBM.

Prints one JSON line: cell-updates/s (value), the roofline of the dominant
kernel (ez_stage: per-launch duration from HIP events recorded on its launch
stream inside the timed region) and a CPU baseline timed on this host.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 512]
                       [--dtype fp32|fp64] [--no-cpu-baseline]
"""

from __future__ import annotations

import argparse
import json
import math
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# Algorithmic HBM bytes per cell, in scalars, of the 2N-storage schedule with
# stencil halos served on chip:
#   stage 1: read U (7), write U' (7), write dU (7)            = 21
#   stages 2-4: read U, dU; write U', dU                       = 28 each
#   stage 5: read U, dU; write U' (dU is dead afterwards)      = 21
#   Kreiss-Oliger: read E, B (6), write E', B' (6); P, which it
#   leaves unchanged, is written by stage 5 into its final buffer = 12
SCALARS_PER_CELL_UPDATE = 21 + 3 * 28 + 21 + 12  # = 138
SCALARS_PER_CELL = {"ez_stage": 28, "ez_stage_first": 21, "ez_stage_last": 21, "ez_ko": 12}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--dtype", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="default workload only: skip the extra keys (attainable bandwidth, "
                         "batched sweep, retiring sweep, Kerr-Schild Wald)")
    ap.add_argument("--extra-steps", type=int, default=10,
                    help="timed steps of the sweep and Kerr-Schild extras")
    ap.add_argument("--retire-scale", type=float, default=5e-4,
                    help="retiring sweep: fraction of the steps of 5 rotation periods")
    ap.add_argument("--variant", type=int, default=0,
                    help="0 = LDS-tiled (default tile), 1 = simple, 2-4 = 64-wide tiles of "
                         "4/8/16 rows, 5/6 = 128/256-wide tiles, 7 = 128 x 4")
    ap.add_argument("--zchunk", type=int, default=0, help="z planes per workgroup (0 = auto)")
    ap.add_argument("--cpu-sample-n", type=int, default=0,
                    help="CPU baseline sample size (0: 256 with several MPI ranks, else 128)")
    ap.add_argument("--cpu-sample-steps", type=int, default=10)
    ap.add_argument("--workload", choices=["alfven3d", "sph-sweep", "gr-wald", "unet-conv"],
                    default="alfven3d",
                    help="alfven3d: the headline flat 512^3 run (default); sph-sweep: batched 2D "
                         "pulsar runs; gr-wald: Kerr-Schild Wald problem; unet-conv: the U-Net's "
                         "BF16 MFMA convolution and weight-gradient kernels")
    ap.add_argument("--runs", type=int, default=256, help="sph-sweep: number of runs")
    ap.add_argument("--tile", type=int, default=-1,
                    help="sph-sweep: tile index (default: autotuned)")
    ap.add_argument("--spin", type=float, default=0.9, help="gr-wald: black-hole spin a")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="run only a plain device copy of this many MiB (3 launches), for "
                         "calibrating rocprofv3 FETCH_SIZE/WRITE_SIZE against known bytes")
    return ap.parse_args()


def copy_calibration(mib: int):
    import torch

    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_read": n * 4, "bytes_written": n * 4}))


def alfven_params(n: int, dtype):
    from coffeegpu_amd.params import SimParams

    return SimParams(dt=0.4 / n, N=[n, n, n], guard=[3, 3, 3], lower=[0.0, 0.0, 0.0],
                     size=[1.0, 1.0, 1.0], nodes=[1, 1, 1], periodic_boundary=[True] * 3,
                     pml=[0, 0, 0], divB_clean=True, ch2=1.0, tau=2.0, KOeps=0.1,
                     use_edotb_damping=False, problem=0, calc_current=True, clean_ep=True,
                     check_egb=True, KO_geometry=False)


def load_traffic(name: str):
    """HBM bytes per launch measured by rocprofv3 --pmc (FETCH_SIZE+WRITE_SIZE,
    copy-kernel calibrated), committed under profiles/ — or None."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f)
    except Exception:
        return None


def cpu_baseline(args, dtype):
    """Reference CPU build (oracle/_ref) on a bounded sample of the same
    workload.  The reference is single-threaded per MPI rank and scales with
    ranks, so it runs under mpiexec with one rank per core (up to 16, the z
    axis split over the ranks, nodes = [1, 1, ranks]); with no mpiexec it runs
    as one rank; with no reference build, the oracle port (OpenMP) is timed."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    from coffeegpu_amd import initial_data as ic
    from coffeegpu_amd.params import Grid

    ranks = min(16, os.cpu_count() or 1) if O.mpiexec_available() else 1
    n = args.cpu_sample_n if args.cpu_sample_n > 0 else (256 if ranks > 1 else 128)
    while ranks > 1 and n // ranks < 3:  # every rank needs at least `guard` planes
        ranks //= 2
    sp = alfven_params(n, dtype)
    g = Grid.from_params(sp, dtype)
    fields, _ = ic.alfven_packets_seeded(g, args.seed, dtype=dtype)
    U0 = ic.stack(fields)
    del fields
    steps = args.cpu_sample_steps
    sample = f"{n}^3 cells, {steps} steps after 2 warm-up steps, same physics and dtype"
    if O.reference_available(dtype):
        tries = [ranks, 1] if ranks > 1 else [1]
        for nr in tries:
            try:
                _, _, err = O.run_reference(sp, U0, steps + 2, timing_warmup=2, ranks=nr)
                for line in err.splitlines():
                    if line.startswith("TIMING"):
                        kv = dict(t.split("=") for t in line.split()[1:])
                        return {"value": float(kv["cell_updates_per_s"]),
                                "unit": "cell-updates/s", "cores": nr, "kind": "reference",
                                "sample": sample + f"; reference CPU build (use_cuda=0), {nr} MPI "
                                                   f"rank(s), one per core, z split over ranks",
                                "host_cpus": os.cpu_count()}
            except Exception as e:  # reference binary or mpiexec unusable on this host
                print(f"[bench] reference CPU build with {nr} rank(s) unavailable here: {e}",
                      file=sys.stderr)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    st = O.FlatState(U0, dtype)
    O.flat_evolve(sp, st, 2, mask_lanes=16 if dtype == np.float32 else 8)
    t0 = time.perf_counter()
    O.flat_evolve(sp, st, steps, time=2 * sp.dt, mask_lanes=16 if dtype == np.float32 else 8)
    dt = time.perf_counter() - t0
    return {"value": n ** 3 * steps / dt, "unit": "cell-updates/s", "cores": threads,
            "kind": "port", "sample": sample + f"; project CPU oracle, OpenMP {threads} threads"}


def _kernel_table(timing, scalars, esz, cells):
    return {k: {"launches": c, "avg_ms": m / c,
                "algorithmic_GBps": scalars[k] * esz * cells / (m / c * 1e-3) / 1e9
                if k in scalars else None}
            for k, (c, m) in timing.items()}


def sph_sweep_params(rng, nr: int, nth: int):
    """One run of the batched 2D pulsar sweep: star R = 1, b0 = 1, x = ln r from
    -0.1 to ln 60, theta in [0, pi], absorbing layer r > 48, R_LC ~ U[3, 20]."""
    import math

    from coffeegpu_amd.params import SimParams

    r_out, r_abs = 60.0, 48.0
    x0 = -0.1
    size_x = math.log(r_out) - x0
    dx, dth = size_x / nr, math.pi / nth
    rlc = float(rng.uniform(3.0, 20.0))
    pml = int(math.ceil((math.log(r_out) - math.log(r_abs)) / dx))
    return SimParams(dt=0.4 * min(dx, dth), N=[nr, nth, 1], guard=[3, 4, 0], lower=[x0, 0.0, 0.0],
                     size=[size_x, math.pi, 0.0], nodes=[1, 1, 1],
                     periodic_boundary=[False, False, False], pml=[pml, 0, 0], pmllen=10,
                     sigpml=1.0, divB_clean=True, ch2=1.0, tau=2.0, KOeps=0.1, KO_geometry=True,
                     use_edotb_damping=False, calc_current=True, clean_ep=True, check_egb=True,
                     problem=0, radius=1.0, omega=1.0 / rlc, b0=1.0, dw0=0.0, dw1=0.0), rlc


# sph stage kernels move the same scalars per cell as the flat ones
SPH_SCALARS = {"sph_stage": 28, "sph_stage_first": 21, "sph_stage_last": 21, "sph_ko": 14}


def _sph_batch(args, runs: int, retire_scale: float = 0.0):
    """The batched sweep's runs (grids and R_LC drawn from the seed).  With
    retire_scale > 0 each run stops after retire_scale x the steps of 5
    rotation periods (5 * 2 pi R_LC / dt), so runs retire at different
    times."""
    import math

    from coffeegpu_amd import initial_data as ic
    from coffeegpu_amd.solver import SimData, SimEnvironment
    from coffeegpu_amd.spherical import SphBatch

    rng = np.random.default_rng(args.seed)
    shapes = [(512, 256), (1024, 512), (2048, 1024)]
    envs, datas, cells, rlcs, nsteps = [], [], [], [], []
    layer_cells = 0  # cells beyond the absorbing layer's inner edge (Un kept)
    for r in range(runs):
        nr, nth = shapes[int(rng.integers(0, len(shapes)))]
        sp, rlc = sph_sweep_params(rng, nr, nth)
        dx = sp.size[0] / nr
        xh = sp.lower[0] + sp.size[0] - sp.pml[0] * dx
        layer_cells += int(np.count_nonzero(sp.lower[0] + dx * np.arange(nr) > xh)) * nth
        env = SimEnvironment(sp, np.float32, "cuda:0")
        d = SimData(env)
        f = ic.pulsar_dipole_sph(env.grid(), sp, dtype=np.float32)
        for name, arr in zip(ic.FIELDS, d.state_arrays()):
            arr.set(f[name])
        envs.append(env)
        datas.append(d)
        cells.append(nr * nth)
        rlcs.append(rlc)
        nsteps.append(max(1, int(math.ceil(retire_scale * 5 * 2 * math.pi * rlc / sp.dt)))
                      if retire_scale > 0 else 0)
    b = SphBatch(envs, datas, nsteps=nsteps)
    b.layer_fraction = layer_cells / float(sum(cells))
    return b, shapes, np.array(cells), np.array(rlcs), np.array(nsteps), datas


def sph_sweep_result(args, runs: int, steps: int, warmup: int) -> dict:
    """Batched 2D pulsar sweep: `runs` runs advanced by one graph-replayed
    launch sequence per step (runs never retire during the timing)."""
    import torch

    t0 = time.perf_counter()
    b, shapes, cells_r, rlcs, _, datas = _sph_batch(args, runs)
    cells = int(cells_r.sum())
    if args.tile >= 0:
        b.set_tile(args.tile)
        best, tile_ms = args.tile, []
    else:
        best, tile_ms = b.autotune(trials=3)
    b.use_graph(True)
    print(f"[bench] sph-sweep {runs} runs, {cells / 1e6:.1f} M cells, set-up "
          f"{time.perf_counter() - t0:.1f}s, tile {best}", file=sys.stderr, flush=True)
    b.step(warmup)
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    b.step(steps)
    stop.record()
    torch.cuda.synchronize()
    ms_per_step = start.elapsed_time(stop) / steps
    value = cells / (ms_per_step * 1e-3)
    # per-kernel durations: direct launches with events (graphs are off while timing)
    b.set_timing(True)
    b.step(3)
    torch.cuda.synchronize()
    timing = b.get_timing()
    b.set_timing(False)
    cnt, ms = timing["sph_stage"]
    bytes_per_launch = SPH_SCALARS["sph_stage"] * 4 * cells
    achieved = bytes_per_launch / (ms / cnt * 1e-3) / 1e9
    # the first stage writes, the last reads the step-start copy Un beyond the
    # absorbing layer's inner edge
    f_un = b.layer_fraction
    sph_bpc = {k: 4.0 * v for k, v in SPH_SCALARS.items()}
    sph_bpc["sph_stage_first"] += 4.0 * 7 * f_un
    sph_bpc["sph_stage_last"] += 4.0 * 7 * f_un
    step_scalars = sum(SPH_SCALARS[k] * (3 if k == "sph_stage" else 1) for k in SPH_SCALARS) + 14 * f_un
    lum = b.luminosity(1.5 * rlcs)
    res = b.residuals()
    # HBM bytes per middle-stage launch, measured for one tile and batch size
    tr = load_traffic("sph_sweep_fp32_traffic.json") or {}
    traffic = tr.get("bytes_per_launch") if (tr.get("tile") == best and
                                             tr.get("runs") == runs and
                                             tr.get("seed") == args.seed) else None
    b.close()
    del b, datas
    return {
        "value": value, "unit": "cell-updates/s", "ms_per_step": ms_per_step, "steps": steps,
        "warmup": warmup,
        "config": {"workload": f"sph-sweep {runs} runs, grids from {shapes}, R_LC~U[3,20], "
                               "r_out 60, absorbing r>48, hipGraph replay, autotuned tile",
                   "cells": cells, "tile": best, "tile_ms": tile_ms},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "kernel": "sph_stage",
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "un_fraction": f_un,
        "step_algorithmic_GBps": step_scalars * 4 * value / 1e9,
        "frac_step_616B": (616.0 + 56.0 * f_un) * value / (HBM_PEAK_GBS * 1e9),
        "kernels": {k: {"launches": c, "avg_ms": m / c,
                        "algorithmic_GBps": sph_bpc[k] * cells / (m / c * 1e-3) / 1e9
                        if k in sph_bpc else None} for k, (c, m) in timing.items()},
        "luminosity_1p5RLC_median": float(np.median(lum)),
        "residuals_max_divB": max(r["max_divB"] for r in res),
        "residuals_nonfinite": sum(r["nonfinite"] for r in res),
    }


def sph_retiring_result(args, runs: int, scale: float) -> dict:
    """The sweep run to its (scaled) end: run r stops after scale x the steps
    of 5 rotation periods, so the batch thins out as short runs retire.
    value = sum of the runs' cell-updates / wall time until the last run
    retires (graph replay, inactive runs' work items skipped)."""
    import torch

    t0 = time.perf_counter()
    b, _, cells_r, rlcs, nsteps, datas = _sph_batch(args, runs, retire_scale=scale)
    if args.tile >= 0:
        b.set_tile(args.tile)
    else:
        b.autotune(trials=3)
    b.use_graph(True)
    total = int((cells_r * nsteps).sum())
    nmax = int(nsteps.max())
    print(f"[bench] sph-sweep retiring: {runs} runs, {nsteps.min()}..{nmax} steps, set-up "
          f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    done = 0
    while done < nmax:  # chunks, so a long run prints progress
        k = min(200, nmax - done)
        b.step(k)
        done += k
        torch.cuda.synchronize()
        print(f"[bench] sph-sweep retiring: {done}/{nmax} steps", file=sys.stderr, flush=True)
    stop.record()
    torch.cuda.synchronize()
    ms = start.elapsed_time(stop)
    act, _, st = b.status()
    ok = (sum(act) == 0) and all(int(a) == int(e) for a, e in zip(st, nsteps))
    b.close()
    del b, datas
    return {"value": total / (ms * 1e-3), "unit": "cell-updates/s", "wall_ms": ms,
            "cell_updates": total, "steps_min": int(nsteps.min()), "steps_max": nmax,
            "retire_scale": scale, "all_runs_retired_at_target": bool(ok),
            "config": f"{runs} runs, each to {scale} x the steps of 5 rotation periods"}


def run_sph_sweep(args):
    r = sph_sweep_result(args, args.runs, args.steps, args.warmup)
    out = {"metric": "cell-updates/s (batched 2D log-spherical pulsar sweep, 1 GPU)",
           "value": r["value"], "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (rotating dipoles, seeded grids and R_LC)",
           "config": dict(r["config"], parallelism="single GPU"), "roofline": r["roofline"],
           "cpu_baseline": None}
    out.update({k: v for k, v in r.items() if k not in out and k not in ("config",)})
    print(json.dumps(out))


GR_SCALARS = {"gr_stage": 28, "gr_stage_first": 21, "gr_stage_last": 21, "gr_ko": 14,
              "gr_point": 18}
# Kerr-Schild algorithmic bytes per cell add, to the field scalars, the
# minimal FP64 metric (H, l_x, l_y, l_z = 32 B) read by each of the 5 stages
# and, in the absorbing layer and the excision region, the step-start copy Un
# (7 scalars written by stage 1, read by stage 5).  The tiled kernels compute
# the metric instead of reading it; the count is the work's, not the schedule's.
GR_METRIC_BYTES = 32


def gr_un_fraction(sp) -> float:
    """Fraction of interior cells where the Kerr-Schild stages keep the
    step-start copy Un: outside the absorbing layer's inner bounds on any axis,
    or inside r1 = 0.7 r_H (the excision damping), at the cell positions
    lower + i delta the kernels use."""
    from coffeegpu_amd.initial_data import cks_metric_np

    n = sp.N
    delta = [sp.size[a] / n[a] for a in range(3)]
    pos = [sp.lower[a] + delta[a] * np.arange(n[a]) for a in range(3)]
    lo = [sp.lower[a] + sp.pml[a] * delta[a] for a in range(3)]
    hi = [sp.lower[a] + sp.size[a] - sp.pml[a] * delta[a] for a in range(3)]
    pml_any = any(v > 0 for v in sp.pml)
    outx = (pos[0] > hi[0]) | (pos[0] < lo[0])
    outy = (pos[1] > hi[1]) | (pos[1] < lo[1])
    r1 = 0.7 * (1.0 + math.sqrt(1.0 - sp.a * sp.a))
    x, y = np.meshgrid(pos[0], pos[1], indexing="xy")
    oxy = outx[None, :] | outy[:, None]
    count = 0
    for zk in pos[2]:
        outer = (oxy | ((zk > hi[2]) or (zk < lo[2]))) if pml_any else np.zeros_like(oxy)
        r = cks_metric_np(sp.a, x, y, np.full_like(x, zk))[0]
        count += int(np.count_nonzero(outer | (r < r1)))
    return count / float(n[0] * n[1] * n[2])


def gr_bytes_per_cell(esz: int, f_un: float) -> dict:
    return {"gr_stage": 28 * esz + GR_METRIC_BYTES,
            "gr_stage_first": 21 * esz + GR_METRIC_BYTES + 7 * esz * f_un,
            "gr_stage_last": 21 * esz + GR_METRIC_BYTES + 7 * esz * f_un,
            "gr_ko": 14 * esz}


UNET_SHAPES = [(8, 128, 64, 32, 32), (8, 256, 128, 64, 64), (8, 128, 64, 128, 128),
               (8, 64, 32, 256, 256), (8, 16, 8, 256, 256)]


def run_unet_conv(args):
    """Forward 3x3 convolution and weight + bias gradient of the surrogate's
    kernels at U-Net layer shapes (N images x H x W, Cin -> Cout), HIP events
    around `--steps` launches each; FLOP = 2 N H W Cin Cout 9 per launch."""
    import ctypes

    import torch

    from coffeegpu_amd import _lib

    L = _lib.load()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    p = lambda t: ctypes.c_void_p(t.data_ptr())  # noqa: E731
    reps = max(1, args.steps)

    def timed(fn):
        for _ in range(max(1, args.warmup)):
            _lib.check(fn())
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    rows = []
    for (n, h, w, cin, cout) in UNET_SHAPES:
        x = torch.randn(n * h * w * cin, device="cuda").to(torch.bfloat16)
        wt = torch.randn(9 * cout * cin, device="cuda").to(torch.bfloat16)
        b = torch.zeros(cout, device="cuda")
        out = torch.empty(n * h * w * cout, device="cuda", dtype=torch.bfloat16)
        dz = torch.randn(n * h * w * cout, device="cuda").to(torch.bfloat16)
        nsplit = max(1, min(256, n * h * w // 512))
        part = torch.empty(nsplit * 9 * cout * cin, device="cuda")
        gw = torch.empty(9 * cout * cin, device="cuda")
        gb = torch.empty(cout, device="cuda")
        ms_f = timed(lambda: L.coffee_unet_conv3x3(p(x), p(wt), p(b), p(out), None, n, h, w, cin,
                                                   cout, 0, 1, s))
        ms_w = timed(lambda: L.coffee_unet_conv3x3_wgrad(p(x), p(dz), p(part), nsplit, p(gw), p(gb),
                                                         n, h, w, cin, cout, s))
        fl = 2.0 * n * h * w * cin * cout * 9
        rows.append({"shape": [n, h, w, cin, cout], "conv_us": ms_f * 1e3,
                     "conv_TFLOPs": fl / ms_f / 1e9, "wgrad_us": ms_w * 1e3,
                     "wgrad_TFLOPs": fl / ms_w / 1e9})
    peak = 2500.0
    best = max(r["conv_TFLOPs"] for r in rows)
    out = {"metric": "TFLOP/s (U-Net BF16 3x3 convolution, forward, 1 GPU)", "value": best,
           "unit": "TFLOP/s", "n_gpus": 1, "steps": reps, "warmup": args.warmup,
           "higher_is_better": True, "dtype": "bf16 (fp32 accumulate)", "data": "synthetic",
           "config": {"workload": "unet-conv layer shapes N x H x W, Cin -> Cout"},
           "roofline": {"bound": "mfma", "achieved": best, "peak": peak, "unit": "TFLOP/s",
                        "frac": best / peak, "traffic": None, "kernel": "conv3x3"},
           "shapes": rows}
    print(json.dumps(out))


def gr_wald_params(n: int, spin: float):
    """Kerr-Schild Wald problem: box [-20, 20]^3 (M = 1), guard 3, absorbing
    layer of 32 cells (sigpml 10, pmllen 1), dt = 0.4 delta."""
    from coffeegpu_amd.params import SimParams

    delta = 40.0 / n
    return SimParams(dt=0.4 * delta, N=[n, n, n], guard=[3, 3, 3], lower=[-20.0] * 3,
                     size=[40.0] * 3, nodes=[1, 1, 1], periodic_boundary=[False] * 3,
                     pml=[32, 32, 32], pmllen=1, sigpml=10.0, divB_clean=True, ch2=1.0, tau=2.0,
                     KOeps=0.1, use_edotb_damping=False, calc_current=True, clean_ep=True,
                     check_egb=True, problem=0, a=spin)


def gr_result(args, n: int, spin: float, dtype, steps: int, warmup: int,
              variant: int = 0, zchunk: int = 0) -> dict:
    import torch

    from coffeegpu_amd import initial_data as ic
    from coffeegpu_amd.gr import FieldSolverGREZ
    from coffeegpu_amd.solver import SimData, SimEnvironment

    sp = gr_wald_params(n, spin)
    env = SimEnvironment(sp, dtype, "cuda:0")
    t0 = time.perf_counter()
    f = ic.wald_uniform(env.grid(), sp.a, noise=1e-3, seed=args.seed, dtype=dtype)
    data = SimData(env)
    for name, arr in zip(ic.FIELDS, data.state_arrays()):
        arr.set(f[name])
    del f
    solver = FieldSolverGREZ(data, env)
    solver.set_schedule(variant, zchunk)
    print(f"[bench] gr-wald {n}^3 a={sp.a} variant {variant} set-up "
          f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    for _ in range(warmup):
        solver.evolve_fields(0.0)
    torch.cuda.synchronize()
    solver.set_timing(True)
    solver.reset_timing()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        solver.evolve_fields(0.0)
    stop.record()
    torch.cuda.synchronize()
    ms_per_step = start.elapsed_time(stop) / steps
    timing = solver.get_timing()
    solver.set_timing(False)
    cells = n ** 3
    value = cells / (ms_per_step * 1e-3)
    esz = np.dtype(dtype).itemsize
    cnt, ms = timing["gr_stage"]
    f_un = gr_un_fraction(sp)
    bpc = gr_bytes_per_cell(esz, f_un)
    bytes_per_launch = bpc["gr_stage"] * cells
    achieved = bytes_per_launch / (ms / cnt * 1e-3) / 1e9
    fields_only = GR_SCALARS["gr_stage"] * esz * cells / (ms / cnt * 1e-3) / 1e9
    # per cell-update: 5 stages of 28 scalars + 14 for KO, the metric read per
    # stage, Un written and read in the layer
    step_bytes = (5 * 28 + 14) * esz + 5 * GR_METRIC_BYTES + 14 * esz * f_un
    res = solver.residuals()
    dname = "fp32" if dtype == np.float32 else "fp64"
    traffic = load_traffic(f"gr_stage_{n}_{dname}_traffic.json")
    if traffic and (traffic.get("variant", 0) != variant or traffic.get("spin", spin) != spin):
        traffic = None  # measured for another schedule or spin
    solver.close()
    del solver, data
    torch.cuda.empty_cache()
    return {
        "value": value, "unit": "cell-updates/s", "ms_per_step": ms_per_step, "steps": steps,
        "warmup": warmup, "dtype": "f32" if dtype == np.float32 else "f64",
        "config": {"workload": f"gr-wald {n}^3 box [-20,20]^3 a={sp.a}, pml 32 (sigpml 10, "
                               "pmllen 1), horizon excision, FP64 metric on the fly"
                               + (" (halo points of stages 2-5 from a precomputed table)"
                                  if variant == 2 else ""),
                   "cells": cells, "schedule": "tiled" if variant != 1 else "point+stage"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": traffic.get("bytes_per_launch") if traffic else None,
                     "kernel": "gr_stage", "algorithmic_bytes_per_launch": bytes_per_launch,
                     "bytes_per_cell": bpc["gr_stage"],
                     "frac_fields_only": fields_only / HBM_PEAK_GBS},
        "un_fraction": f_un,
        "step_bytes_per_cell_update": step_bytes,
        "frac_step": value * step_bytes / (HBM_PEAK_GBS * 1e9),
        "kernels": {k: {"launches": c, "avg_ms": m / c,
                        "algorithmic_GBps": bpc[k] * cells / (m / c * 1e-3) / 1e9 if k in bpc else None}
                    for k, (c, m) in timing.items()},
        "residuals_after": res,
    }


def run_gr(args):
    n = args.n if args.n != 512 else 384
    dtype = np.float32 if args.dtype == "fp32" else np.float64
    r = gr_result(args, n, args.spin, dtype, args.steps, args.warmup, args.variant, args.zchunk)
    out = {"metric": "cell-updates/s (Kerr-Schild GR Wald problem, 1 GPU)",
           "value": r["value"], "unit": "cell-updates/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": r["dtype"],
           "data": "synthetic (Wald start B = z/sqrt(gamma) + 1e-3 seeded noise)",
           "config": dict(r["config"], parallelism="single GPU"), "roofline": r["roofline"],
           "kernels": r["kernels"], "residuals_after": r["residuals_after"],
           "un_fraction": r["un_fraction"], "step_bytes_per_cell_update": r["step_bytes_per_cell_update"],
           "frac_step": r["frac_step"], "cpu_baseline": None}
    print(json.dumps(out))


def attainable_bandwidth(total_mib: int = 8192) -> dict:
    """HBM bandwidth of a plain read + write stream (coffee_stream_copy,
    16 bytes per lane, a contiguous chunk per workgroup, nontemporal), best
    over loads in flight per lane and grid sizes;
    bytes counted = bytes read + bytes written."""
    import ctypes

    import torch

    from coffeegpu_amd import _lib

    L = _lib.load()
    n = total_mib * (1 << 20) // 2
    src = torch.ones(n // 4, dtype=torch.float32, device="cuda:0")
    dst = torch.empty_like(src)
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    best = {"GBps": 0.0}
    rows = []
    for unroll in (1, 2, 4, 8):
        for blocks in (1024, 4096, 16384):
            call = lambda: L.coffee_stream_copy(ctypes.c_void_p(src.data_ptr()),  # noqa: E731
                                                ctypes.c_void_p(dst.data_ptr()), n, unroll,
                                                blocks, s)
            for _ in range(2):
                _lib.check(call())
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(5):
                call()
            e1.record()
            torch.cuda.synchronize()
            gbs = 2.0 * n * 5 / (e0.elapsed_time(e1) * 1e-3) / 1e9
            rows.append([unroll, blocks, round(gbs, 1)])
            if gbs > best["GBps"]:
                best = {"GBps": gbs, "unroll": unroll, "blocks": blocks}
    del src, dst
    torch.cuda.empty_cache()
    best.update({"bytes_per_copy_read": n, "kernel": "coffee_stream_copy (16 B per lane, chunk per workgroup, nontemporal)",
                 "sweep_unroll_blocks_GBps": rows})
    return best


def main():
    args = parse()
    if args.copy_calibration_mib:
        copy_calibration(args.copy_calibration_mib)
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; multi-GPU runs are not "
                                   "part of it (the halo exchange is the single-GPU periodic "
                                   "fill)", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "sph-sweep":
        return run_sph_sweep(args)
    if args.workload == "unet-conv":
        return run_unet_conv(args)
    if args.workload == "gr-wald":
        return run_gr(args)
    import torch

    from coffeegpu_amd import initial_data as ic
    from coffeegpu_amd.solver import FieldSolverEZ, SimData, SimEnvironment

    dtype = np.float32 if args.dtype == "fp32" else np.float64
    n = args.n
    sp = alfven_params(n, dtype)
    env = SimEnvironment(sp, dtype, "cuda:0")
    t0 = time.perf_counter()
    fields, amps = ic.alfven_packets_seeded(env.grid(), args.seed, dtype=dtype)
    data = SimData(env)
    for name, arr in zip(ic.FIELDS, data.state_arrays()):
        arr.set(fields[name])
    del fields
    solver = FieldSolverEZ(data, env)
    solver.set_schedule(args.variant, args.zchunk)
    print(f"[bench] {n}^3 {args.dtype} set-up {time.perf_counter() - t0:.1f}s, amplitudes "
          f"{amps[0]:.3f} {amps[1]:.3f}", file=sys.stderr, flush=True)

    t = 0.0
    for _ in range(args.warmup):
        solver.evolve_fields(t)
        t += sp.dt
    torch.cuda.synchronize()
    solver.set_timing(True)
    solver.reset_timing()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    wall0 = time.perf_counter()
    start.record()
    for _ in range(args.steps):
        solver.evolve_fields(t)
        t += sp.dt
    stop.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms_total = start.elapsed_time(stop)
    timing = solver.get_timing()
    solver.set_timing(False)
    res = solver.residuals()
    ms_per_step = ms_total / args.steps
    cells = n ** 3
    value = cells / (ms_per_step * 1e-3)

    solver.close()
    del solver, data
    torch.cuda.empty_cache()

    esz = np.dtype(dtype).itemsize
    cnt, ms = timing["ez_stage"]
    avg_ms = ms / cnt
    bytes_per_launch = SCALARS_PER_CELL["ez_stage"] * esz * cells
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    traffic = load_traffic(f"ez_stage_{n}_{args.dtype}_traffic.json")
    if traffic and traffic.get("variant", 0) != args.variant:
        traffic = None  # measured for another schedule
    kernels = _kernel_table(timing, SCALARS_PER_CELL, esz, cells)
    step_gbs = SCALARS_PER_CELL_UPDATE * esz * value / 1e9

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args, dtype)

    out = {
        "metric": "cell-updates/s (flat EZ 3D Alfven collision, 1 GPU)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if dtype == np.float32 else "f64",
        "data": "synthetic (seeded Alfven packets)",
        "config": {"workload": f"alfven3d {n}^3 periodic, guard 3, {args.dtype}, LSRK(5,4)+KO6, "
                               "cleaning + E.B/E<B projections",
                   "cells": cells, "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": (traffic or {}).get("bytes_per_launch"),
                     "kernel": "ez_stage",
                     "algorithmic_bytes_per_launch": bytes_per_launch},
        "step_algorithmic_GBps": step_gbs,
        # whole-step fraction of peak with the two byte counts in use: 552 B
        # (this schedule's compulsory transfers) and 616 B (5 x 28 + 14 scalars)
        "frac_step_552B": SCALARS_PER_CELL_UPDATE * esz * value / (HBM_PEAK_GBS * 1e9),
        "frac_step_616B": 154 * esz * value / (HBM_PEAK_GBS * 1e9),
        "kernels": kernels,
        "wall_s_timed": wall,
        "residuals_after": res,
        "cpu_baseline": cpu,
    }
    if not args.no_extras:
        att = attainable_bandwidth()
        out["attainable_GBps"] = att["GBps"]
        out["attainable"] = att
        out["frac_of_attainable"] = achieved / att["GBps"]
        print(f"[bench] attainable {att['GBps']:.0f} GB/s", file=sys.stderr, flush=True)
        out["sph_sweep"] = sph_sweep_result(args, 256, args.extra_steps, 3)
        out["sph_sweep_retiring"] = sph_retiring_result(args, 256, args.retire_scale)
        out["gr_wald"] = gr_result(args, 384, 0.9, np.float32, args.extra_steps, 2)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
