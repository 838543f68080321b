"""CPU emulation of the index arithmetic of the LDS-tiled kernels
(coffeegpu_amd/csrc/ez_flat_tiled.hpp): every global load/store index stays
inside the guarded array and every LDS slot inside its buffer, for periodic
and non-periodic axes, partial tiles and odd z chunks; and the cells an
active lane reads are exactly the stencil neighbours (with periodic wrap)."""

import itertools

import numpy as np
import pytest


def amap(v, n, g, per):
    v = np.clip(v, -g, n - 1 + g)
    if per:
        v = np.where(v < 0, v + n, np.where(v >= n, v - n, v))
    return v + g


def halo_xy(h, TX, TY, R):
    if h < 2 * R * TX:
        r = h // TX
        return h % TX, (r - R if r < R else TY + r - R)
    u = h - 2 * R * TX
    c = u // TY
    return (c - R if c < R else TX + c - R), u % TY


def emulate(n, g, per, TY, R, zc):
    TX = 64
    dims = [n[a] + 2 * g for a in range(3)]
    LX, LY = TX + 2 * R, TY + 2 * R
    NT = TX * TY
    H = 2 * R * (TX + TY)
    HPT = (H + NT - 1) // NT
    ntx, nty = -(-n[0] // TX), -(-n[1] // TY)
    nzc = -(-n[2] // zc)
    total = dims[0] * dims[1] * dims[2]
    t = np.arange(NT)
    lx, ly = t % TX, t // TX
    for tx, ty, tz in itertools.product(range(ntx), range(nty), range(nzc)):
        i0, j0, z0 = tx * TX, ty * TY, tz * zc
        z1 = min(z0 + zc, n[2])
        gi, gj = i0 + lx, j0 + ly
        active = (gi < n[0]) & (gj < n[1])
        ai = amap(gi, n[0], g, per[0])
        aj = amap(gj, n[1], g, per[1])
        col = ai + aj * dims[0]
        slot = (ly + R) * LX + (lx + R)
        assert slot.min() >= 0 and slot.max() < LX * LY
        for h0 in range(HPT):
            idx = t + h0 * NT
            ok = idx < H
            for tt in np.nonzero(ok)[0]:
                hx, hy = halo_xy(int(idx[tt]), TX, TY, R)
                hs = (hy + R) * LX + (hx + R)
                assert 0 <= hs < LX * LY
                hc = amap(i0 + hx, n[0], g, per[0]) + amap(j0 + hy, n[1], g, per[1]) * dims[0]
                assert 0 <= hc < dims[0] * dims[1]
        zs = list(range(z0 - R, z0 + R + 1)) + [z + R + 1 for z in range(z0, z1)] + \
            list(range(z0, z1))
        for z in zs:
            off = amap(z, n[2], g, per[2]) * dims[0] * dims[1]
            idx = col + off
            assert idx.min() >= 0 and idx.max() < total
        # stores / dU of active lanes: interior cells only
        for z in range(z0, z1):
            c = col[active] + (z + g) * dims[0] * dims[1]
            assert c.min() >= 0 and c.max() < total
        # active lanes read true stencil neighbours
        for d in range(-R, R + 1):
            a = amap(gi[active] + d, n[0], g, per[0])
            want = (gi[active] + d) % n[0] + g if per[0] else gi[active] + d + g
            assert np.array_equal(a, want)


@pytest.mark.parametrize("n", [(16, 16, 16), (16, 8, 8), (32, 16, 16), (70, 13, 9), (128, 24, 7)])
@pytest.mark.parametrize("per", [(1, 1, 1), (0, 0, 0), (1, 0, 1)])
@pytest.mark.parametrize("TY,R", [(8, 2), (8, 3), (4, 2), (4, 3)])
@pytest.mark.parametrize("zc", [3, 5, 64])
def test_tiled_indices_in_bounds(n, per, TY, R, zc):
    if any(per[a] and n[a] < 3 for a in range(3)):
        pytest.skip("periodic axis shorter than the guard")
    emulate(n, 3, per, TY, R, zc)
