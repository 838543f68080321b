"""Pins the FP64 U-Net oracle (oracle/unet_oracle.py) by central finite
differences of its loss, and its building blocks by direct definitions."""

import numpy as np
import pytest

import unet_oracle as U


def test_conv3x3_matches_direct_sum():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((2, 5, 6, 3))
    W = rng.standard_normal((3, 3, 3, 4))
    b = rng.standard_normal(4)
    out = U.conv3x3(x, W, b)
    n, y, xx, co = 1, 0, 5, 2
    want = b[co]
    for ky in range(3):
        for kx in range(3):
            yy, xs = y + ky - 1, xx + kx - 1
            if 0 <= yy < 5 and 0 <= xs < 6:
                want += (W[ky, kx, :, co] * x[n, yy, xs, :]).sum()
    assert abs(out[n, y, xx, co] - want) < 1e-12


def test_pool_and_upsample_adjoint():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 4, 6, 3))
    y = rng.standard_normal((2, 2, 3, 3))
    assert abs((U.pool2(x) * y).sum() - (x * U.pool2_bwd(y)).sum()) < 1e-12
    assert abs((U.up2(y) * x).sum() - (y * U.up2_bwd(x)).sum()) < 1e-12


@pytest.mark.parametrize("levels", [2, 3])
def test_backward_matches_finite_differences(levels):
    specs = U.layer_specs(levels, 4, 3, 2)
    params = U.init_params(specs, seed=levels)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((2, 8, 8, 3))
    t = rng.standard_normal((2, 8, 8, 2))
    out, tape = U.forward(params, specs, x, levels)
    _, dout = U.mse(out, t)
    grads, dx = U.backward(params, specs, tape, dout, levels)

    def loss():
        return U.mse(U.forward(params, specs, x, levels)[0], t)[0]

    h = 1e-6
    checked = 0
    for name, _, _, _ in specs:
        W, b = params[name]
        for idx in [(0, 0, 0, 0), (1, 2, W.shape[2] - 1, W.shape[3] - 1), (2, 1, 0, W.shape[3] // 2)]:
            old = W[idx]
            W[idx] = old + h
            lp = loss()
            W[idx] = old - h
            lm = loss()
            W[idx] = old
            num = (lp - lm) / (2 * h)
            ana = grads[name][0][idx]
            assert abs(num - ana) <= 1e-6 * max(1.0, abs(num)) + 1e-9, (name, idx, num, ana)
            checked += 1
        old = b[0]
        b[0] = old + h
        lp = loss()
        b[0] = old - h
        lm = loss()
        b[0] = old
        assert abs((lp - lm) / (2 * h) - grads[name][1][0]) <= 1e-6 * max(1.0, abs(grads[name][1][0])) + 1e-9
    # input gradient
    for idx in [(0, 3, 4, 1), (1, 7, 0, 2)]:
        old = x[idx]
        x[idx] = old + h
        lp = loss()
        x[idx] = old - h
        lm = loss()
        x[idx] = old
        assert abs((lp - lm) / (2 * h) - dx[idx]) <= 1e-6 * max(1.0, abs(dx[idx])) + 1e-9
    assert checked == 3 * len(specs)


def test_parameter_count_of_the_warm_start_net():
    specs = U.layer_specs(4, 32, 7, 6)
    n = sum(9 * ci * co + co for _, ci, co, _ in specs)
    assert 2.0e6 < n < 2.3e6
