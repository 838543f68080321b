"""GPU U-Net kernels (include/coffee_unet.h) against the FP64 CPU oracle.

Tolerances: BF16 inputs and weights are rounded identically on both sides,
so a single convolution differs only by FP32 accumulation (<= 1e-5 relative);
through a whole network the BF16 rounding of every stored activation and
activation gradient adds about 2^-9 relative per layer, so forward outputs
agree to 3e-2 and parameter gradients to 1e-1 relative L2 (17 layers at 4
levels), with cosine similarity >= 0.995.
"""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")
if not torch.cuda.is_available():  # pragma: no cover
    pytest.skip("no GPU", allow_module_level=True)

import ctypes  # noqa: E402

import unet_oracle as UO  # noqa: E402
from coffeegpu_amd import _lib  # noqa: E402
from coffeegpu_amd.surrogate import UNet, pad16  # noqa: E402


def bf16_bits(x):
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_round(x):
    return (bf16_bits(x).astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def dev_bf16(x):
    return torch.from_numpy(bf16_bits(x).view(np.int16)).to("cuda:0")


def s():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else ctypes.c_void_p(0)


def test_bf16_conversion_round_to_nearest_even():
    L = _lib.load()
    x = np.array([1.0, 1.00390625, 1.01171875, -3.14159, 1e-30, 65504.0, 3.3895e38], np.float32)
    src = torch.from_numpy(x).to("cuda:0")
    out = torch.empty(len(x), dtype=torch.int16, device="cuda:0")
    _lib.check(L.coffee_unet_to_bf16(p(src), p(out), len(x), s()))
    assert np.array_equal(out.cpu().numpy().view(np.uint16), bf16_bits(x))


# cin 16: the 16x16x16 kernel; cin % 32 == 0: the K = 32 kernel with LDS-staged
# weights (one or two 16-pixel blocks per wavefront, 1/2/4 channel blocks,
# partial pixel and channel tiles); cin % 32 == 0, cout % 64 == 0 and
# w % 16 == 0: the LDS-tiled kernel (8-row tiles, and 16-row tiles where they
# make at least 512 workgroups, as in the last case; partial last tile rows,
# several channel blocks and chunks, flipped taps)
@pytest.mark.parametrize("cin,cout,flip,n,h,w", [
    (16, 16, 0, 2, 7, 9), (16, 48, 0, 2, 7, 9), (32, 64, 0, 2, 7, 9), (64, 32, 1, 2, 7, 9),
    (96, 80, 1, 2, 7, 9), (32, 32, 0, 2, 7, 9), (32, 64, 0, 1, 256, 256),
    (64, 16, 1, 1, 257, 259), (64, 128, 1, 2, 20, 48), (32, 64, 0, 1, 10, 32),
    (128, 64, 0, 1, 33, 16), (96, 192, 1, 1, 16, 64), (32, 64, 1, 4, 121, 256)])
def test_conv3x3_matches_oracle(cin, cout, flip, n, h, w):
    rng = np.random.default_rng(cin + cout + flip + h)
    x = bf16_round(rng.standard_normal((n, h, w, cin)))
    W = bf16_round(rng.standard_normal((3, 3, cin, cout)) * 0.2)
    b = rng.standard_normal(cout).astype(np.float32).astype(np.float64)
    Wk = W[::-1, ::-1] if flip else W  # the backward data pass mirrors the taps
    want = UO.conv3x3(x, Wk, b)
    wdev = np.transpose(W.reshape(9, cin, cout), (0, 2, 1))  # [tap][co][ci]
    out32 = torch.empty(n * h * w * cout, dtype=torch.float32, device="cuda:0")
    bd = torch.from_numpy(b.astype(np.float32)).to("cuda:0")
    # device buffers stay referenced until the asynchronous launch has finished
    xd, wd = dev_bf16(x), dev_bf16(wdev)
    _lib.check(_lib.load().coffee_unet_conv3x3(p(xd), p(wd), p(bd), None, p(out32), n, h, w, cin,
                                               cout, flip, 0, s()))
    got = out32.cpu().numpy().reshape(n, h, w, cout)
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()


@pytest.mark.parametrize("cin,cout,h,w", [(64, 64, 20, 32), (32, 32, 9, 7)])
def test_conv3x3_bf16_output_is_rounded_fp32_output(cin, cout, h, w):
    """One launch writing both outputs (bias and ReLU applied): the BF16 output
    is the FP32 output rounded to nearest even, in both epilogues (LDS-tiled
    kernel and K = 32 kernel)."""
    rng = np.random.default_rng(5)
    n = 2
    x = bf16_round(rng.standard_normal((n, h, w, cin)))
    W = bf16_round(rng.standard_normal((9, cout, cin)) * 0.2)
    b = rng.standard_normal(cout).astype(np.float32)
    xd, wd = dev_bf16(x), dev_bf16(W)
    bd = torch.from_numpy(b).to("cuda:0")
    out32 = torch.empty(n * h * w * cout, dtype=torch.float32, device="cuda:0")
    out = torch.empty(n * h * w * cout, dtype=torch.bfloat16, device="cuda:0")
    _lib.check(_lib.load().coffee_unet_conv3x3(p(xd), p(wd), p(bd), p(out), p(out32), n, h, w,
                                               cin, cout, 0, 1, s()))
    torch.cuda.synchronize()
    assert float(out32.min()) >= 0.0  # ReLU
    assert torch.equal(out32.to(torch.bfloat16), out)


# the last five take the spatial-tile kernel (cin, cout multiples of 64,
# w of 16): 8-row tiles with a partial last row, 4-row tiles, two channel
# blocks, one range (bias by the separate passes), more ranges than tiles
@pytest.mark.parametrize("n,h,w,cin,cout,nsplit", [(2, 6, 10, 32, 16, 3), (1, 40, 52, 64, 48, 7),
                                                    (2, 12, 16, 16, 272, 2), (2, 13, 32, 64, 64, 5),
                                                    (1, 6, 16, 128, 64, 3), (2, 20, 48, 64, 128, 4),
                                                    (1, 16, 32, 64, 64, 1), (2, 9, 16, 128, 128, 40)])
def test_conv3x3_wgrad_matches_oracle(n, h, w, cin, cout, nsplit):
    rng = np.random.default_rng(11 + cout)
    x = bf16_round(rng.standard_normal((n, h, w, cin)))
    dz = bf16_round(rng.standard_normal((n, h, w, cout)))
    _, dW, db = UO.conv3x3_bwd(x, np.zeros((3, 3, cin, cout)), dz)
    part = torch.empty(nsplit * 9 * cout * cin, dtype=torch.float32, device="cuda:0")
    gw = torch.empty(9 * cout * cin, dtype=torch.float32, device="cuda:0")
    gb = torch.empty(cout, dtype=torch.float32, device="cuda:0")
    xd, dzd = dev_bf16(x), dev_bf16(dz)
    _lib.check(_lib.load().coffee_unet_conv3x3_wgrad(p(xd), p(dzd), p(part), nsplit, p(gw), p(gb),
                                                     n, h, w, cin, cout, s()))
    got = np.transpose(gw.cpu().numpy().reshape(9, cout, cin), (0, 2, 1)).reshape(3, 3, cin, cout)
    assert np.abs(got - dW).max() <= 1e-5 * np.abs(dW).max()
    assert np.abs(gb.cpu().numpy() - db).max() <= 1e-5 * np.abs(db).max()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


@pytest.mark.parametrize("levels,base", [(2, 16), (3, 16), (4, 32)])
def test_network_forward_and_gradients_match_oracle(levels, base):
    cin, cout = 7, 6
    n, h, w = 2, 16, 32
    specs = UO.layer_specs(levels, base, cin, cout)
    params = UO.init_params(specs, seed=levels)
    params = {k: (bf16_round(W), b.astype(np.float32).astype(np.float64)) for k, (W, b) in params.items()}
    rng = np.random.default_rng(levels)
    x = bf16_round(rng.standard_normal((n, h, w, cin)))
    t = rng.standard_normal((n, h, w, cout)).astype(np.float32).astype(np.float64)
    want, tape = UO.forward(params, specs, x, levels)
    _, dout = UO.mse(want, t)
    grads, _ = UO.backward(params, specs, tape, dout, levels)

    net = UNet(levels, base, cin, cout, seed=0)
    net.set_params(params)
    xp = np.zeros((n, h, w, pad16(cin)))
    xp[..., :cin] = x
    out = net.forward(dev_bf16(xp), n, h, w)
    got = out.cpu().numpy().reshape(n, h, w, pad16(cout))[..., :cout]
    ef = _rel(got, want)
    loss, dgrad = net.loss_and_grad(out, torch.from_numpy(t.astype(np.float32)).to("cuda:0"), n, h, w)
    net.backward(dgrad)
    g = net.get_grads()
    worst = 0.0
    for name, _, _, _ in specs:
        ew = _rel(g[name][0], grads[name][0])
        cos = float((g[name][0] * grads[name][0]).sum() /
                    (np.linalg.norm(g[name][0]) * np.linalg.norm(grads[name][0]) + 1e-300))
        worst = max(worst, ew)
        assert cos >= 0.995, (name, cos)
        assert _rel(g[name][1], grads[name][1]) <= 1e-1, name
    print(f"levels {levels} base {base}: forward rel {ef:.2e}, worst weight-grad rel {worst:.2e}, "
          f"loss {loss:.4e} vs {UO.mse(want, t)[0]:.4e}")
    assert ef <= 3e-2
    assert worst <= 1e-1
    assert abs(loss - UO.mse(want, t)[0]) <= 3e-2 * UO.mse(want, t)[0]


def test_training_reduces_loss():
    n, h, w = 4, 16, 16
    net = UNet(2, 16, 7, 6, seed=1)
    rng = np.random.default_rng(2)
    yy, xx = np.mgrid[0:h, 0:w] / 8.0
    xp = np.zeros((n, h, w, 16))
    feats = [yy, xx, np.sin(yy), np.cos(xx), yy * xx / 4, 1 - yy]
    for k, f in enumerate(feats):
        xp[..., k] = f
    xp[..., 6] = 0.1 * rng.standard_normal((n, h, w))
    X = dev_bf16(xp)
    T = np.stack([np.sin(yy + k) * np.cos(xx * (k + 1)) for k in range(6)], -1)[None].repeat(n, 0)
    Td = torch.from_numpy(T.astype(np.float32)).to("cuda:0")
    losses = []
    for it in range(150):
        out = net.forward(X, n, h, w)
        loss, d = net.loss_and_grad(out, Td, n, h, w)
        net.backward(d)
        net.adam_step(lr=3e-3)
        losses.append(loss)
    print(f"loss {losses[0]:.4f} -> {losses[-1]:.4f}")
    assert losses[-1] < 0.2 * losses[0]


def test_pack_unpack_round_trip():
    from coffeegpu_amd.solver import SimData, SimEnvironment
    from tests.golden.cases_sph import case_sph

    sp, U0 = case_sph(np.float32, n=(64, 32), omega=0.2)
    env = SimEnvironment(sp, np.float32, "cuda:0")
    d = SimData(env)
    d.set_state(U0)
    g = env.grid()
    n0, n1 = g.reduced_dim(0), g.reduced_dim(1)
    fields = (ctypes.c_void_p * 6)(*[a.dev_ptr() for a in d.state_arrays()[:6]])
    buf = torch.zeros(2 * n0 * n1 * 8, dtype=torch.float32, device="cuda:0")
    L = _lib.load()
    _lib.check(L.coffee_unet_pack(fields, 4, g.dims[0], g.guard[0], g.guard[1], n0, n1, g.lower[0],
                                  g.delta[0], g.delta[1], 2.0, 0.25, p(buf), 1, 8, 1, s()))
    arr = buf.cpu().numpy().reshape(2, n1, n0, 8)
    assert np.all(arr[0] == 0) and np.all(arr[1, :, :, 6] == np.float32(0.25))
    # zero the arrays, unpack, compare the interior (the phi components vanish on the axis row)
    for a in d.state_arrays():
        a.assign(0.0)
    _lib.check(L.coffee_unet_unpack(p(buf), 8, 1, fields, 4, g.dims[0], g.guard[0], g.guard[1], n0,
                                    n1, g.lower[0], g.delta[0], g.delta[1], 2.0, s()))
    got = d.get_state()
    it = (slice(0, 6),) + g.interior()
    ref = U0[it].copy()
    ref[2, 0, 0] = ref[5, 0, 0] = 0.0  # theta = 0 row: h_phi = 0
    gi = got[it]
    gi[2, 0, 0] = gi[5, 0, 0] = 0.0
    assert np.abs(gi - ref).max() <= 1e-5 * np.abs(ref).max()
