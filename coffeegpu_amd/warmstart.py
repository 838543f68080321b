"""Learned warm start of axisymmetric pulsar magnetospheres.

Pipeline, all on one GPU with the project's kernels:
  1. relax: a batch of log-spherical runs (rotating-dipole start, R_LC drawn
     from a range) advances to t = periods * 2 pi R_LC in one batched,
     graph-replayed solver; their final states are the training targets;
  2. train: the U-Net (coffeegpu_amd.surrogate) learns the map
     (dipole start, R_LC) -> final state, as a correction to the start, in
     orthonormal components times r^RPOW;
  3. evaluate: held-out R_LC values run from the dipole start and from the
     network's prediction; the spin-down luminosity at 1.5 R_LC is sampled
     every `check` steps, and each start's cost is the first step after which
     L stays within `tol` of the dipole run's final luminosity.  The field
     error of the prediction against the relaxed state is reported too, and
     the solver's own residual monitor (max |div B|, max |E.B|/B^2,
     max (E^2 - B^2)/B^2 and non-finite cells, coffee_sph_batch_residuals)
     of both starts and of both runs at the end.

The dipole start remains the fallback: a prediction is used only as initial
data and the solver does all the evolution.

usage: python -m coffeegpu_amd.warmstart [--nr 256 --nth 128 --train 128 --test 16 ...]
prints one JSON line.
"""

from __future__ import annotations

import argparse
import ctypes
import json
import math
import sys
import time
from typing import List, Sequence

import numpy as np

from . import _lib
from . import initial_data as ic
from .params import SimParams
from .solver import SimData, SimEnvironment
from .spherical import SphBatch
from .surrogate import RPOW, UNet, pad16


def _torch():
    import torch

    return torch


def run_params(nr: int, nth: int, rlc: float, r_out: float = 60.0, r_abs: float = 48.0) -> SimParams:
    """Star R = 1, b0 = 1, x = ln r from -0.1 to ln r_out, theta in [0, pi],
    absorbing layer r > r_abs, Omega = 1/R_LC, dt = 0.4 min(dx, dtheta).

    Kreiss-Oliger uses the flat 2D stencil (KO_geometry off): with the
    sqrt(gamma)-weighted form, whose weights vanish on the axis, the open
    field of runs with R_LC >~ 8 dissipates and the spin-down luminosity
    decays for as long as the run lasts, so there is no steady state to
    learn or to reach."""
    x0 = -0.1
    size_x = math.log(r_out) - x0
    dx, dth = size_x / nr, math.pi / nth
    pml = int(math.ceil((math.log(r_out) - math.log(r_abs)) / dx))
    return SimParams(dt=0.4 * min(dx, dth), N=[nr, nth, 1], guard=[3, 4, 0], lower=[x0, 0.0, 0.0],
                     size=[size_x, math.pi, 0.0], nodes=[1, 1, 1],
                     periodic_boundary=[False, False, False], pml=[pml, 0, 0], pmllen=10,
                     sigpml=1.0, divB_clean=True, ch2=1.0, tau=2.0, KOeps=0.1, KO_geometry=False,
                     use_edotb_damping=False, calc_current=True, clean_ep=True, check_egb=True,
                     problem=0, radius=1.0, omega=1.0 / rlc, b0=1.0, dw0=0.0, dw1=0.0)


class Runs:
    """A batch of runs with their environments and caller-owned state."""

    def __init__(self, rlcs: Sequence[float], nr: int, nth: int, nsteps: Sequence[int],
                 dtype=np.float32):
        self.rlcs = list(rlcs)
        self.envs, self.datas = [], []
        for rlc in self.rlcs:
            sp = run_params(nr, nth, rlc)
            env = SimEnvironment(sp, dtype, "cuda:0")
            d = SimData(env)
            f = ic.pulsar_dipole_sph(env.grid(), sp, dtype=dtype)
            for name, arr in zip(ic.FIELDS, d.state_arrays()):
                arr.set(f[name])
            self.envs.append(env)
            self.datas.append(d)
        self.nsteps = list(nsteps)
        self.batch = SphBatch(self.envs, self.datas, nsteps=self.nsteps)
        self.batch.use_graph(True)

    def close(self):
        self.batch.close()


def _field_ptrs(d: SimData):
    return (ctypes.c_void_p * 6)(*[a.dev_ptr() for a in d.state_arrays()[:6]])


def _geom(env):
    g = env.grid()
    return g, g.reduced_dim(0), g.reduced_dim(1)


def pack_runs(runs: Runs, out, cst: int, out_fp32: bool, extra_scale: float = 1.0 / 20.0):
    """States of all runs into an NHWC tensor (image b = run b)."""
    torch = _torch()
    L = _lib.load()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for b, (env, d, rlc) in enumerate(zip(runs.envs, runs.datas, runs.rlcs)):
        g, n0, n1 = _geom(env)
        _lib.check(L.coffee_unet_pack(_field_ptrs(d), env.dtype.itemsize, g.dims[0], g.guard[0],
                                      g.guard[1], n0, n1, g.lower[0], g.delta[0], g.delta[1], RPOW,
                                      rlc * extra_scale, ctypes.c_void_p(out.data_ptr()),
                                      int(out_fp32), cst, b, s))


def unpack_runs(runs: Runs, src, cst: int):
    torch = _torch()
    L = _lib.load()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for b, (env, d) in enumerate(zip(runs.envs, runs.datas)):
        g, n0, n1 = _geom(env)
        _lib.check(L.coffee_unet_unpack(ctypes.c_void_p(src.data_ptr()), cst, b, _field_ptrs(d),
                                        env.dtype.itemsize, g.dims[0], g.guard[0], g.guard[1], n0,
                                        n1, g.lower[0], g.delta[0], g.delta[1], RPOW, s))


def corotate_runs(runs: Runs):
    """Replace E by the corotating field of B, E = -(Omega x r) x B, in every
    run's interior (the steady aligned rotator satisfies it on every field
    line, so a predicted B fixes E)."""
    torch = _torch()
    L = _lib.load()
    s = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for env, d in zip(runs.envs, runs.datas):
        g, n0, n1 = _geom(env)
        _lib.check(L.coffee_unet_corotate(_field_ptrs(d), env.dtype.itemsize, g.dims[0],
                                          g.guard[0], g.guard[1], n0, n1, g.lower[0], g.delta[0],
                                          g.delta[1], env.params().omega, s))


def relax(runs: Runs, chunk: int = 500, log=None):
    """Advance every run to its own step target."""
    torch = _torch()
    target = max(runs.nsteps)
    done = 0
    while done < target:
        k = min(chunk, target - done)
        runs.batch.step(k)
        done += k
        torch.cuda.synchronize()
        if log:
            log(f"relax {done}/{target}")


def luminosity_history(runs: Runs, check: int, total: Sequence[int], log=None):
    """Step every run to total[b] while sampling L(1.5 R_LC) every `check`
    steps.  Returns per-run lists of (step, L)."""
    torch = _torch()
    hist = [[] for _ in runs.rlcs]
    radii = 1.5 * np.asarray(runs.rlcs)
    horizon = max(total)
    step = 0
    n = 0
    while step < horizon:
        runs.batch.step(check)
        step += check
        L = runs.batch.luminosity(radii)
        for b in range(len(runs.rlcs)):
            if step <= total[b]:
                hist[b].append((step, float(L[b])))
        n += 1
        if log and n % 50 == 0:
            log(f"evaluate {step}/{horizon}")
    torch.cuda.synchronize()
    return hist


def _res_summary(r: dict) -> dict:
    """The residual monitor's figures of one run (coffee_sph_batch_residuals):
    max |div B|, max |E.B|/B^2, max (E^2 - B^2)/B^2 (metric norms) and the
    number of non-finite cells."""
    return {q: float(r[q]) for q in ("max_divB", "max_EdotB", "max_E2mB2", "nonfinite")}


def settle_step(hist, L_ref: float, tol: float):
    """First sampled step after which L stays within tol of L_ref."""
    last_bad = 0
    for step, L in hist:
        if not abs(L - L_ref) <= tol * abs(L_ref):
            last_bad = step
    if last_bad == hist[-1][0]:
        return None
    nxt = [s for s, _ in hist if s > last_bad]
    return nxt[0] if nxt else None


def train(net: UNet, X, T, n: int, h: int, w: int, epochs: int, batch: int, lr: float, seed: int = 0,
          log=None) -> List[float]:
    """Adam on mini-batches of consecutive images, order reshuffled per epoch."""
    torch = _torch()
    cin = pad16(net.cin)
    rng = np.random.default_rng(seed)
    nb = n // batch
    losses = []
    for ep in range(epochs):
        tot = 0.0
        it_in_epoch = 0
        for b in rng.permutation(nb):
            x = X[b * batch * h * w * cin:(b + 1) * batch * h * w * cin]
            t = T[b * batch * h * w * net.cout:(b + 1) * batch * h * w * net.cout]
            out = net.forward(x, batch, h, w)
            loss, d = net.loss_and_grad(out, t, batch, h, w)
            net.backward(d)
            # linear warm-up over the first two epochs, then cosine decay
            warm = min(1.0, (ep * nb + it_in_epoch + 1) / (2.0 * nb))
            it_in_epoch += 1
            net.adam_step(lr=lr * warm * (0.5 * (1 + math.cos(math.pi * ep / epochs))))
            tot += loss
        losses.append(tot / nb)
        if log and (ep % max(1, epochs // 10) == 0 or ep == epochs - 1):
            log(f"epoch {ep} loss {losses[-1]:.4e}")
    torch.cuda.synchronize()
    return losses


def experiment(nr=256, nth=128, n_train=128, n_test=16, periods=3.0, rlc_lo=3.0, rlc_hi=20.0,
               levels=4, base=32, epochs=60, batch=8, lr=1e-3, tol=0.01, seed=0, log=None,
               eval_periods=None, corotate=False):
    torch = _torch()
    t_start = time.perf_counter()
    rng = np.random.default_rng(seed)
    rl_train = rng.uniform(rlc_lo, rlc_hi, n_train)
    rl_test = rng.uniform(rlc_lo, rlc_hi, n_test)
    dt = run_params(nr, nth, 5.0).dt
    steps_of = lambda r: int(math.ceil(periods * 2 * math.pi * r / dt))  # noqa: E731
    cin, cout = pad16(7), 6
    H, W = nth, nr

    # 1. training targets from the batched solver
    runs = Runs(rl_train, nr, nth, [steps_of(r) for r in rl_train])
    X = torch.zeros(n_train * H * W * cin, dtype=torch.int16, device="cuda:0")
    X0 = torch.zeros(n_train * H * W * cout, dtype=torch.float32, device="cuda:0")
    pack_runs(runs, X, cin, False)
    pack_runs(runs, X0, cout, True)
    t0 = time.perf_counter()
    relax(runs, log=log)
    t_relax = time.perf_counter() - t0
    T = torch.zeros(n_train * H * W * cout, dtype=torch.float32, device="cuda:0")
    pack_runs(runs, T, cout, True)
    T -= X0  # the network learns the correction to the dipole start
    cells = sum(nr * nth * s for s in runs.nsteps)
    runs.close()
    del X0
    # per-channel scale: every channel of the correction enters the loss at unit RMS
    Tv = T.view(-1, cout)
    chan_rms = torch.sqrt((Tv * Tv).mean(dim=0)).clamp_min(1e-12)
    Tv /= chan_rms

    # 2. training
    net = UNet(levels, base, 7, 6, seed=seed)
    t0 = time.perf_counter()
    losses = train(net, X, T, n_train, H, W, epochs, batch, lr, seed=seed, log=log)
    t_train = time.perf_counter() - t0
    del X, T

    # 3. evaluation on held-out R_LC: dipole start vs predicted start, both run
    # past the training horizon so that the dipole run's late luminosity is a
    # relaxed reference
    ep = eval_periods if eval_periods else periods + 1.0
    steps = [int(math.ceil(ep * 2 * math.pi * r / dt)) for r in rl_test]
    check = max(1, min(steps) // 60)
    dip = Runs(rl_test, nr, nth, steps)
    warm = Runs(rl_test, nr, nth, steps)
    Xt = torch.zeros(n_test * H * W * cin, dtype=torch.int16, device="cuda:0")
    pack_runs(warm, Xt, cin, False)
    pred = torch.zeros(n_test * H * W * cout, dtype=torch.float32, device="cuda:0")
    base0 = torch.zeros_like(pred)
    pack_runs(warm, base0, cout, True)
    for b in range(0, n_test, batch):
        m = min(batch, n_test - b)
        out = net.forward(Xt[b * H * W * cin:(b + m) * H * W * cin], m, H, W)
        o = (out.view(m * H * W, pad16(cout))[:, :cout] * chan_rms).reshape(-1)
        pred[b * H * W * cout:(b + m) * H * W * cout] = o + base0[b * H * W * cout:(b + m) * H * W * cout]
    unpack_runs(warm, pred, cout)
    if corotate:
        corotate_runs(warm)
    # the solver's own residual monitor on both starts and at the end
    r_start = (dip.batch.residuals(), warm.batch.residuals())
    h_dip = luminosity_history(dip, check, steps, log=log)
    h_warm = luminosity_history(warm, check, steps, log=log)
    r_end = (dip.batch.residuals(), warm.batch.residuals())
    # field error of the prediction vs the relaxed (dipole-run) state, in network units
    relaxed = torch.zeros_like(pred)
    pack_runs(dip, relaxed, cout, True)
    err = []
    for b in range(n_test):
        sl = slice(b * H * W * cout, (b + 1) * H * W * cout)
        num = torch.linalg.vector_norm(pred[sl] - relaxed[sl]).item()
        den0 = torch.linalg.vector_norm(base0[sl] - relaxed[sl]).item()
        err.append((num / max(torch.linalg.vector_norm(relaxed[sl]).item(), 1e-30),
                    den0 / max(torch.linalg.vector_norm(relaxed[sl]).item(), 1e-30)))
    res = []
    for b in range(n_test):
        # relaxed reference: the dipole run's mean luminosity over its last quarter
        tail = [L for _, L in h_dip[b][-max(1, len(h_dip[b]) // 4):]]
        L_ref = float(np.mean(tail))
        sd = settle_step(h_dip[b], L_ref, tol)
        sw = settle_step(h_warm[b], L_ref, tol)
        by_tol = {str(tt): [settle_step(h_dip[b], L_ref, tt), settle_step(h_warm[b], L_ref, tt)]
                  for tt in (0.01, 0.02, 0.05, 0.1)}
        stride = max(1, len(h_dip[b]) // 24)
        res.append({"rlc": float(rl_test[b]), "L_ref": L_ref, "steps_dipole": sd, "steps_warm": sw,
                    "settle_by_tol": by_tol,
                    "total_steps": steps[b], "field_err_warm": err[b][0],
                    "field_err_dipole": err[b][1],
                    "residuals": {"start_dipole": _res_summary(r_start[0][b]),
                                  "start_warm": _res_summary(r_start[1][b]),
                                  "end_dipole": _res_summary(r_end[0][b]),
                                  "end_warm": _res_summary(r_end[1][b])},
                    "L_dipole": [round(L / L_ref, 4) for _, L in h_dip[b][::stride]],
                    "L_warm": [round(L / L_ref, 4) for _, L in h_warm[b][::stride]],
                    "L_steps": [st for st, _ in h_dip[b][::stride]]})
    dip.close()
    warm.close()
    gains = [r["steps_dipole"] / r["steps_warm"] for r in res
             if r["steps_dipole"] and r["steps_warm"]]
    gain_by_tol = {}
    for tt in ("0.01", "0.02", "0.05", "0.1"):
        g = [r["settle_by_tol"][tt][0] / r["settle_by_tol"][tt][1] for r in res
             if r["settle_by_tol"][tt][0] and r["settle_by_tol"][tt][1]]
        gain_by_tol[tt] = {"median_gain": float(np.median(g)) if g else None, "n": len(g),
                           "dipole_settled": sum(1 for r in res if r["settle_by_tol"][tt][0]),
                           "warm_settled": sum(1 for r in res if r["settle_by_tol"][tt][1])}
    return {
        "grid": [nr, nth], "n_train": n_train, "n_test": n_test, "periods": periods,
        "eval_periods": ep, "tol": tol, "corotate": corotate,
        "rlc_range": [rlc_lo, rlc_hi], "levels": levels, "base": base,
        "params": net.num_params(), "epochs": epochs, "loss_first": losses[0],
        "loss_last": losses[-1], "relax_s": t_relax, "relax_cell_updates_per_s": cells / t_relax,
        "train_s": t_train, "median_gain": float(np.median(gains)) if gains else None,
        "gain_by_tol": gain_by_tol,
        "n_settled_both": len(gains),
        "median_residuals": {
            k: {q: float(np.median([r["residuals"][k][q] for r in res]))
                for q in ("max_divB", "max_EdotB", "max_E2mB2", "nonfinite")}
            for k in ("start_dipole", "start_warm", "end_dipole", "end_warm")},
        "median_field_err_warm": float(np.median([r["field_err_warm"] for r in res])),
        "median_field_err_dipole": float(np.median([r["field_err_dipole"] for r in res])),
        "runs": res, "wall_s": time.perf_counter() - t_start,
    }


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--nr", type=int, default=256)
    ap.add_argument("--nth", type=int, default=128)
    ap.add_argument("--train", type=int, default=128)
    ap.add_argument("--test", type=int, default=16)
    ap.add_argument("--periods", type=float, default=3.0)
    ap.add_argument("--rlc", type=float, nargs=2, default=[3.0, 20.0])
    ap.add_argument("--levels", type=int, default=4)
    ap.add_argument("--base", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=60)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval-periods", type=float, default=0.0)
    ap.add_argument("--corotate", action="store_true",
                    help="set the warm start's E from its B by corotation")
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)

    def log(msg):
        print(f"[warmstart] {msg}", file=sys.stderr, flush=True)

    out = experiment(a.nr, a.nth, a.train, a.test, a.periods, a.rlc[0], a.rlc[1], a.levels, a.base,
                     a.epochs, a.batch, a.lr, seed=a.seed, log=log,
                     eval_periods=a.eval_periods or None, corotate=a.corotate)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
