"""Learned warm start for the batched log-spherical pulsar runs.

A 2D convolutional U-Net (BF16 weights and activations, FP32 accumulation)
maps the rotating-dipole start of a run (6 field channels, orthonormal
components times r^RPOW, and R_LC as a 7th broadcast channel) to its
near-steady magnetosphere.  Forward, backward and Adam are the project's own
gfx950 kernels (include/coffee_unet.h, coffeegpu_amd/csrc/unet.hip); torch
tensors only hold device memory.  The plain dipole start stays the fallback.

Network (levels L, base channels c0, c_l = c0 2^l): encoder level l = two
3x3 conv + ReLU after a 2x2 average pool (l > 0); decoder level l = nearest
2x upsample, 3x3 conv + ReLU, concatenation with the encoder output of level
l, two 3x3 conv + ReLU; head = 3x3 conv to 6 channels.  Channel counts are
padded to multiples of 16 on the device (padded weights stay zero).
"""

from __future__ import annotations

import ctypes
from typing import Dict, List, Optional, Tuple

import numpy as np

from . import _lib

RPOW = 2.0  # fields enter the network as orthonormal components times r^RPOW


def _torch():
    import torch

    return torch


def pad16(c: int) -> int:
    return (c + 15) // 16 * 16


def layer_specs(levels: int, base: int, cin: int, cout: int) -> List[Tuple[str, int, int, bool]]:
    """(name, cin, cout, relu) of every convolution, in forward order."""
    specs = []
    c_prev = cin
    for l in range(levels):
        c = base << l
        specs.append((f"enc{l}a", c_prev, c, True))
        specs.append((f"enc{l}b", c, c, True))
        c_prev = c
    for l in range(levels - 2, -1, -1):
        c = base << l
        specs.append((f"dec{l}up", c << 1, c, True))
        specs.append((f"dec{l}a", 2 * c, c, True))
        specs.append((f"dec{l}b", c, c, True))
    specs.append(("head", base, cout, False))
    return specs


def _p(t) -> ctypes.c_void_p:
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


class UNet:
    """U-Net on the device.  Parameters: FP32 masters W[tap][co][ci] (padded),
    biases, Adam moments, BF16 copies W and WT."""

    def __init__(self, levels: int = 4, base: int = 32, cin: int = 7, cout: int = 6,
                 seed: int = 0, device: str = "cuda:0", stream=None):
        self.levels, self.base, self.cin, self.cout = levels, base, cin, cout
        self.specs = layer_specs(levels, base, cin, cout)
        self.device = device
        self._lib = _lib.load()
        self._stream = stream
        self.params: Dict[str, dict] = {}
        rng = np.random.default_rng(seed)
        for name, ci, co, relu in self.specs:
            cip, cop = pad16(ci), pad16(co)
            w = np.zeros((9, cop, cip), dtype=np.float32)
            std = np.sqrt(2.0 / (9 * ci))
            w[:, :co, :ci] = (rng.standard_normal((9, co, ci)) * std).astype(np.float32)
            b = np.zeros(cop, dtype=np.float32)
            self.params[name] = self._alloc_param(w, b, ci, co, relu)
        self.t = 0
        self._bufs: Dict[tuple, object] = {}

    # ---- parameters ------------------------------------------------------
    def _alloc_param(self, w: np.ndarray, b: np.ndarray, ci: int, co: int, relu: bool) -> dict:
        torch = _torch()
        dev = self.device
        cop, cip = w.shape[1], w.shape[2]
        P = {"ci": ci, "co": co, "cip": cip, "cop": cop, "relu": relu,
             "w": torch.from_numpy(np.ascontiguousarray(w)).to(dev),
             "b": torch.from_numpy(np.ascontiguousarray(b)).to(dev)}
        for k in ("mw", "vw", "gw"):
            P[k] = torch.zeros_like(P["w"])
        for k in ("mb", "vb", "gb"):
            P[k] = torch.zeros_like(P["b"])
        P["wb"] = torch.empty(w.size, dtype=torch.int16, device=dev)
        P["wbt"] = torch.empty(w.size, dtype=torch.int16, device=dev)
        self._refresh_bf16(P)
        return P

    def _refresh_bf16(self, P):
        _lib.check(self._lib.coffee_unet_weights_bf16(_p(P["w"]), _p(P["wb"]), _p(P["wbt"]),
                                                      P["cop"], P["cip"], self._s()))

    def set_params(self, params_np: Dict[str, Tuple[np.ndarray, np.ndarray]]):
        """Load weights given as W[ky, kx, ci, co], b[co] (the oracle's layout)."""
        torch = _torch()
        for name, ci, co, _ in self.specs:
            W, b = params_np[name]
            P = self.params[name]
            w = np.zeros((9, P["cop"], P["cip"]), dtype=np.float32)
            w[:, :co, :ci] = np.transpose(W.reshape(9, ci, co), (0, 2, 1))
            bb = np.zeros(P["cop"], dtype=np.float32)
            bb[:co] = b
            P["w"].copy_(torch.from_numpy(w))
            P["b"].copy_(torch.from_numpy(bb))
            self._refresh_bf16(P)

    def get_grads(self) -> Dict[str, Tuple[np.ndarray, np.ndarray]]:
        """Gradients in the oracle's layout (W[ky, kx, ci, co], b[co])."""
        out = {}
        for name, ci, co, _ in self.specs:
            P = self.params[name]
            g = P["gw"].cpu().numpy()[:, :co, :ci]
            out[name] = (np.transpose(g, (0, 2, 1)).reshape(3, 3, ci, co).astype(np.float64),
                         P["gb"].cpu().numpy()[:co].astype(np.float64))
        return out

    def num_params(self) -> int:
        return sum(9 * ci * co + co for _, ci, co, _ in self.specs)

    def state_dict(self) -> Dict[str, np.ndarray]:
        out = {}
        for name, P in self.params.items():
            out[name + ".w"] = P["w"].cpu().numpy()
            out[name + ".b"] = P["b"].cpu().numpy()
        return out

    # ---- buffers ---------------------------------------------------------
    def _s(self):
        torch = _torch()
        s = self._stream if self._stream is not None else torch.cuda.current_stream()
        return ctypes.c_void_p(s.cuda_stream)

    def _buf(self, key, numel, dtype):
        torch = _torch()
        t = self._bufs.get(key)
        if t is None or t.numel() < numel or t.dtype != dtype:
            t = torch.empty(numel, dtype=dtype, device=self.device)
            self._bufs[key] = t
        return t

    # ---- kernels -----------------------------------------------------------
    def _conv(self, x, P, n, h, w, out, out32=None, relu=None):
        _lib.check(self._lib.coffee_unet_conv3x3(
            _p(x), _p(P["wb"]), _p(P["b"]), _p(out), _p(out32), n, h, w, P["cip"], P["cop"], 0,
            int(P["relu"] if relu is None else relu), self._s()))

    def forward(self, x, n: int, h: int, w: int):
        """x: BF16 (int16 storage) NHWC tensor with pad16(cin) channels.
        Returns the FP32 NHWC output (pad16(cout) channels, first cout valid)."""
        torch = _torch()
        L = self.levels
        tape = []
        it = iter(self.specs)
        skips = []
        cur, cc, hh, ww = x, pad16(self.cin), h, w
        for l in range(L):
            if l > 0:
                pooled = self._buf(("pool", l), n * (hh // 2) * (ww // 2) * cc, torch.int16)
                _lib.check(self._lib.coffee_unet_pool2(_p(cur), _p(pooled), n, hh, ww, cc, self._s()))
                cur, hh, ww = pooled, hh // 2, ww // 2
            for part in ("a", "b"):
                name = next(it)[0]
                P = self.params[name]
                y = self._buf(("y", name), n * hh * ww * P["cop"], torch.int16)
                self._conv(cur, P, n, hh, ww, y)
                tape.append((name, cur, y, hh, ww))
                cur, cc = y, P["cop"]
            skips.append((cur, cc, hh, ww))
        for l in range(L - 2, -1, -1):
            up = self._buf(("up", l), n * (hh * 2) * (ww * 2) * cc, torch.int16)
            hh, ww = hh * 2, ww * 2
            _lib.check(self._lib.coffee_unet_up2(_p(cur), _p(up), n, hh, ww, cc, self._s()))
            name = next(it)[0]
            P = self.params[name]
            y = self._buf(("y", name), n * hh * ww * P["cop"], torch.int16)
            self._conv(up, P, n, hh, ww, y)
            tape.append((name, up, y, hh, ww))
            sk, skc, _, _ = skips[l]
            c = P["cop"]
            cat = self._buf(("cat", l), n * hh * ww * 2 * c, torch.int16)
            npix = n * hh * ww
            _lib.check(self._lib.coffee_unet_combine(_p(cat), 2 * c, 0, _p(y), c, 0, None, 0, 0,
                                                     None, 0, 0, c, npix, self._s()))
            _lib.check(self._lib.coffee_unet_combine(_p(cat), 2 * c, c, _p(sk), skc, 0, None, 0,
                                                     0, None, 0, 0, c, npix, self._s()))
            cur, cc = cat, 2 * c
            for part in ("a", "b"):
                name = next(it)[0]
                P = self.params[name]
                y = self._buf(("y", name), n * hh * ww * P["cop"], torch.int16)
                self._conv(cur, P, n, hh, ww, y)
                tape.append((name, cur, y, hh, ww))
                cur, cc = y, P["cop"]
        name = next(it)[0]
        P = self.params[name]
        out32 = self._buf(("out32",), n * hh * ww * P["cop"], torch.float32)
        self._conv(cur, P, n, hh, ww, None, out32=out32, relu=False)
        tape.append((name, cur, None, hh, ww))
        self._tape = (tape, n)
        return out32[: n * hh * ww * P["cop"]]

    def backward(self, dout):
        """dout: BF16 gradient of the loss w.r.t. the FP32 output (same layout).
        Fills every parameter's gradient (gw, gb)."""
        torch = _torch()
        tape, n = self._tape
        pos = [len(tape) - 1]

        def conv_back(dz):
            name, xin, y, hh, ww = tape[pos[0]]
            pos[0] -= 1
            P = self.params[name]
            npix = n * hh * ww
            nsplit = max(1, min(256, npix // 512))
            part = self._buf(("part",), nsplit * 9 * P["cop"] * P["cip"], torch.float32)
            _lib.check(self._lib.coffee_unet_conv3x3_wgrad(
                _p(xin), _p(dz), _p(part), nsplit, _p(P["gw"]), _p(P["gb"]), n, hh, ww, P["cip"],
                P["cop"], self._s()))
            dx = self._buf(("dx", name), npix * P["cip"], torch.int16)
            _lib.check(self._lib.coffee_unet_conv3x3(
                _p(dz), _p(P["wbt"]), None, _p(dx), None, n, hh, ww, P["cop"], P["cip"], 1, 0,
                self._s()))
            return dx, P["cip"], hh, ww

        def masked(key, a, ast, aoff, c, y, npix, b=None, bst=0, boff=0):
            out = self._buf(("dz", key), npix * c, torch.int16)
            _lib.check(self._lib.coffee_unet_combine(_p(out), c, 0, _p(a), ast, aoff, _p(b), bst,
                                                     boff, _p(y), c, 0, c, npix, self._s()))
            return out

        L = self.levels
        d, dc, hh, ww = conv_back(dout)  # head (no activation)
        dcat = [None] * L
        for l in range(0, L - 1):
            for part in ("b", "a"):
                name, _, y, hh, ww = tape[pos[0]]
                c = self.params[name]["cop"]
                dz = masked(name, d, dc, 0, c, y, n * hh * ww)
                d, dc, hh, ww = conv_back(dz)
            # d: gradient of concat(y_up, skip_l), 2c channels
            dcat[l] = (d, dc)
            name, _, y, hh, ww = tape[pos[0]]
            c = self.params[name]["cop"]
            dz = masked(name, d, dc, 0, c, y, n * hh * ww)
            d, dc, hh, ww = conv_back(dz)
            up = self._buf(("dup", l), n * (hh // 2) * (ww // 2) * dc, torch.int16)
            _lib.check(self._lib.coffee_unet_up2_bwd(_p(d), _p(up), n, hh, ww, dc, self._s()))
            d, hh, ww = up, hh // 2, ww // 2
        for l in range(L - 1, -1, -1):
            name, _, y, hh, ww = tape[pos[0]]
            c = self.params[name]["cop"]
            if dcat[l] is not None:
                dcl, dcc = dcat[l]
                dz = masked(name, d, dc, 0, c, y, n * hh * ww, b=dcl, bst=dcc, boff=dcc // 2)
            else:
                dz = masked(name, d, dc, 0, c, y, n * hh * ww)
            d, dc, hh, ww = conv_back(dz)
            name, _, y, hh, ww = tape[pos[0]]
            c = self.params[name]["cop"]
            dz = masked(name, d, dc, 0, c, y, n * hh * ww)
            d, dc, hh, ww = conv_back(dz)
            if l > 0:
                pd = self._buf(("dpool", l), n * (hh * 2) * (ww * 2) * dc, torch.int16)
                _lib.check(self._lib.coffee_unet_pool2_bwd(_p(d), _p(pd), n, hh * 2, ww * 2, dc,
                                                           self._s()))
                d, hh, ww = pd, hh * 2, ww * 2
        self.dx = d
        return d

    def loss_and_grad(self, out32, target, n: int, h: int, w: int) -> Tuple[float, object]:
        torch = _torch()
        cop = pad16(self.cout)
        dout = self._buf(("dout",), n * h * w * cop, torch.int16)
        loss = ctypes.c_double()
        _lib.check(self._lib.coffee_unet_mse(_p(out32), cop, _p(target), self.cout, n * h * w,
                                             _p(dout), ctypes.byref(loss), self._s()))
        return loss.value, dout

    def adam_step(self, lr: float = 1e-3, beta1: float = 0.9, beta2: float = 0.999,
                  eps: float = 1e-8):
        self.t += 1
        for P in self.params.values():
            _lib.check(self._lib.coffee_unet_adam(
                _p(P["w"]), _p(P["gw"]), _p(P["mw"]), _p(P["vw"]), P["w"].numel(), lr, beta1,
                beta2, eps, self.t, _p(P["wb"]), _p(P["wbt"]), P["cop"], P["cip"], self._s()))
            _lib.check(self._lib.coffee_unet_adam(
                _p(P["b"]), _p(P["gb"]), _p(P["mb"]), _p(P["vb"]), P["b"].numel(), lr, beta1,
                beta2, eps, self.t, None, None, 0, 0, self._s()))
