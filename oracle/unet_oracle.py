"""FP64 CPU oracle of the surrogate U-Net (TEST INFRASTRUCTURE ONLY).

The learned warm start has no counterpart in the reference; this numpy
restatement of the network the HIP kernels implement (forward and backward)
is the checker for those kernels.  Only tests/ load it.  It is pinned by
central finite differences of its own loss (tests/test_unet_oracle.py).

Layout: activations NHWC (N, H, W, C); a 3x3 convolution (zero padding 1)
has weights W[ky, kx, ci, co] and bias b[co]:
    out[n, y, x, co] = b[co] + sum_{ky,kx,ci} W[ky,kx,ci,co] in[n, y+ky-1, x+kx-1, ci].
Network (levels L, base channels c0, c_l = c0 2^l):
    encoder  level l: conv(c_{l-1} -> c_l) + ReLU, conv(c_l -> c_l) + ReLU -> skip_l,
                      2x2 average pool before levels 1..L-1 (input c_{-1} = cin);
    decoder  level l = L-2 .. 0: nearest 2x upsample, conv(c_{l+1} -> c_l) + ReLU,
                      concat(., skip_l), conv(2 c_l -> c_l) + ReLU, conv(c_l -> c_l) + ReLU;
    head     conv(c_0 -> cout), no activation.
Loss: mean squared error over all outputs.
"""

from __future__ import annotations

from typing import Dict, List, Tuple

import numpy as np


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


def init_params(specs, seed: int = 0) -> Dict[str, Tuple[np.ndarray, np.ndarray]]:
    rng = np.random.default_rng(seed)
    out = {}
    for name, ci, co, _ in specs:
        std = np.sqrt(2.0 / (9 * ci))
        out[name] = (rng.standard_normal((3, 3, ci, co)) * std, rng.standard_normal(co) * 0.01)
    return out


def conv3x3(x, W, b):
    n, h, w, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    out = np.broadcast_to(b, (n, h, w, W.shape[3])).astype(np.float64).copy()
    for ky in range(3):
        for kx in range(3):
            out += xp[:, ky:ky + h, kx:kx + w, :] @ W[ky, kx]
    return out


def conv3x3_bwd(x, W, dz):
    n, h, w, _ = x.shape
    xp = np.pad(x, ((0, 0), (1, 1), (1, 1), (0, 0)))
    dxp = np.zeros_like(xp)
    dW = np.zeros_like(W)
    for ky in range(3):
        for kx in range(3):
            sl = xp[:, ky:ky + h, kx:kx + w, :]
            dW[ky, kx] = np.einsum("nyxi,nyxo->io", sl, dz)
            dxp[:, ky:ky + h, kx:kx + w, :] += dz @ W[ky, kx].T
    return dxp[:, 1:h + 1, 1:w + 1, :], dW, dz.sum(axis=(0, 1, 2))


def pool2(x):
    n, h, w, c = x.shape
    return x.reshape(n, h // 2, 2, w // 2, 2, c).mean(axis=(2, 4))


def pool2_bwd(dy):
    return np.repeat(np.repeat(dy, 2, axis=1), 2, axis=2) / 4.0


def up2(x):
    return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)


def up2_bwd(dy):
    n, h, w, c = dy.shape
    return dy.reshape(n, h // 2, 2, w // 2, 2, c).sum(axis=(2, 4))


def forward(params, specs, x, levels: int):
    """Returns (output, tape) with the tape holding what backward needs."""
    tape = []
    it = iter(specs)
    skips = []
    h = x
    for l in range(levels):
        if l > 0:
            h = pool2(h)
        for _ in range(2):
            name, _, _, relu = next(it)
            z = conv3x3(h, *params[name])
            y = np.maximum(z, 0.0) if relu else z
            tape.append((name, h, y, relu))
            h = y
        skips.append(h)
    for l in range(levels - 2, -1, -1):
        h = up2(h)
        name, _, _, relu = next(it)
        y = np.maximum(conv3x3(h, *params[name]), 0.0)
        tape.append((name, h, y, relu))
        h = np.concatenate([y, skips[l]], axis=3)
        for _ in range(2):
            name, _, _, relu = next(it)
            y = np.maximum(conv3x3(h, *params[name]), 0.0)
            tape.append((name, h, y, relu))
            h = y
    name, _, _, relu = next(it)
    out = conv3x3(h, *params[name])
    tape.append((name, h, out, relu))
    return out, tape


def mse(out, target):
    d = out - target
    return float((d * d).mean()), 2.0 * d / d.size


def backward(params, specs, tape, dout, levels: int):
    """Gradients {name: (dW, db)} of the loss whose output gradient is dout."""
    grads = {}
    pos = len(tape) - 1

    def conv_back(dy):
        nonlocal pos
        name, xin, y, relu = tape[pos]
        pos -= 1
        dz = dy * (y > 0) if relu else dy
        dx, dW, db = conv3x3_bwd(xin, params[name][0], dz)
        grads[name] = (dW, db)
        return dx

    d = conv_back(dout)  # head
    c0 = specs[0][2]
    dskips = [None] * levels
    for l in range(0, levels - 1):
        c = c0 << l
        d = conv_back(d)  # dec{l}b
        d = conv_back(d)  # dec{l}a -> gradient of concat(y_up, skip_l)
        dskips[l] = d[..., c:]
        d = conv_back(d[..., :c])  # dec{l}up
        d = up2_bwd(d)
    for l in range(levels - 1, -1, -1):
        if dskips[l] is not None:
            d = d + dskips[l]
        d = conv_back(d)  # enc{l}b
        d = conv_back(d)  # enc{l}a
        if l > 0:
            d = pool2_bwd(d)
    return grads, d
