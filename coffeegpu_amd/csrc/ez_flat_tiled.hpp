// LDS-tiled, z-marching kernels of the flat EZ path (gfx950).
//
// A workgroup owns a TX x TY tile of the x-y plane and marches it through a
// chunk of z planes.  Per plane, every field of the tile plus an R-cell halo
// in x and y is staged in LDS (double-buffered, one barrier per plane); x and y
// derivatives read LDS, z derivatives read a per-thread register queue of the
// 2R+1 planes around the current one.  Each plane of each input field is
// therefore fetched from HBM once per chunk (plus the x-y halo, which the
// neighbouring tiles of the same XCD have usually brought into L2, and 2R
// extra planes per chunk).  Halo values of the next plane are prefetched into
// registers while the current plane is computed.
//
// Rows of TX = 64, 128 or 256 cells map to one to four wavefronts, so all
// plane loads and stores are fully coalesced rows.  Tiles are assigned XCD-contiguously:
// workgroups are dispatched round-robin over the 8 XCDs, and the remap gives
// each XCD a contiguous range of tiles so that halo rows hit its own L2.
//
// O32 (Kreiss-Oliger kernel): fields addressed by 32-bit byte offsets
// (arrays below 4 GiB, chosen by the host), see ldf / stf.
//
// Partial tiles (n not a multiple of the tile) are handled by clamping loads
// of inactive lanes into the array and masking their stores, so every grid
// the simple kernels accept runs here and gives bitwise the same results
// (same per-cell arithmetic, stage_cell / ko_cell_finish).
#pragma once

#include "ez_flat.hpp"

namespace coffee {

struct TileMap {
  int ntx, nty, nzc, zc;
};

__device__ __forceinline__ int xcd_remap(int b, int nb) {
  if ((nb & 7) != 0) return b;
  return (b & 7) * (nb >> 3) + (b >> 3);
}

// Interior-relative coordinate v to array coordinate: periodic axes wrap
// into the interior, other axes read guard cells.  Loads that feed a result
// have v in [-R, n-1+R] with R <= g; anything else (inactive lanes, the
// look-ahead plane past the last one) is clamped to [-g, n-1+g] first, so
// every access stays inside the guarded array (g >= 3 and n >= g are
// enforced at context creation).
__device__ __forceinline__ int amap(int v, int n, int g, int per) {
  v = v < -g ? -g : (v > n - 1 + g ? n - 1 + g : v);
  if (per) {
    if (v < 0) v += n;
    else if (v >= n) v -= n;
  }
  return v + g;
}

// Whether a tiled kernel keeps the x-y halo of R+LA planes in registers
// (fetched with the centre columns) or of one plane.
template <typename T, bool STAR, int HPT, int NT>
constexpr bool deep_halo() {
  return !STAR && HPT == 1 && sizeof(T) * NT <= 4096;
}

// Planes of look-ahead of the centre columns (LA) of the stage kernels: two
// where the registers allow (FP32 kernels with the deep halo, but not the
// last stage, whose absorbing-layer operands take the room), else one.
template <typename T, bool STAR, bool LAST, int HPT, int NT>
constexpr int centre_lookahead() {
  return deep_halo<T, STAR, HPT, NT>() && sizeof(T) == 4 && !LAST ? 2 : 1;
}

// Halo cell h (0 <= h < 2R(TX+TY)) of a TX x TY tile: R rows below/above
// the tile (x-extent TX), then R columns left/right (y-extent TY).
template <int TX, int TY, int R>
__device__ __forceinline__ void halo_xy(int h, int& hx, int& hy) {
  if (h < 2 * R * TX) {
    const int r = h / TX;
    hx = h % TX;
    hy = r < R ? r - R : TY + r - R;
  } else {
    const int u = h - 2 * R * TX;
    const int cidx = u / TY;
    hy = u % TY;
    hx = cidx < R ? cidx - R : TX + cidx - R;
  }
}

template <typename T, int TX, int TY, bool FIRST, bool LAST, bool DIAG, bool STAR>
__global__ void __launch_bounds__(TX * TY)
    ez_stage_tiled(const FlatK<T> k, const FlatIO<T> io, const T As, const T Bs, const T ts,
                   const TileMap tm) {
  constexpr int R = 2, LX = TX + 2 * R, LY = TY + 2 * R, NF = 7, NT = TX * TY;
  constexpr int H = 2 * R * (TX + TY);
  constexpr int HPT = (H + NT - 1) / NT;
  constexpr int Q = 2 * R + 1;
  __shared__ T lds[2][NF][LY * LX];

  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = L % tm.ntx;
  const int rest = L / tm.ntx;
  const int ty = rest % tm.nty;
  const int tz = rest / tm.nty;
  const int i0 = tx * TX, j0 = ty * TY, z0 = tz * tm.zc;
  const int z1 = min(z0 + tm.zc, k.n[2]);
  const int t = threadIdx.x;
  const int lx = t % TX, ly = t / TX;
  const int gi = i0 + lx, gj = j0 + ly;
  const bool active = gi < k.n[0] && gj < k.n[1];
  const int ai = amap(gi, k.n[0], k.guard[0], k.per[0]);
  const int aj = amap(gj, k.n[1], k.guard[1], k.per[1]);
  const int col = ai + aj * k.dims[0];
  const int slot = (ly + R) * LX + (lx + R);

  int hcol[HPT], hslot[HPT];
  bool hok[HPT];
#pragma unroll
  for (int h = 0; h < HPT; h++) {
    const int idx = t + h * NT;
    hok[h] = idx < H;
    int hx = 0, hy = 0;
    if (hok[h]) halo_xy<TX, TY, R>(idx, hx, hy);
    hcol[h] = amap(i0 + hx, k.n[0], k.guard[0], k.per[0]) +
              amap(j0 + hy, k.n[1], k.guard[1], k.per[1]) * k.dims[0];
    hslot[h] = (hy + R) * LX + (hx + R);
  }

  // q[0..Q-1] = planes z-R..z+R; q[Q..Q+LA-1] hold the next LA planes, the
  // last of which is requested during the iteration of plane z, so every
  // global load has LA full iterations to land.
  constexpr int LA = centre_lookahead<T, STAR, LAST, HPT, NT>();
  T q[Q + LA][NF];
#pragma unroll
  for (int d = 0; d < Q + LA - 1; d++) {
    const int off = amap(z0 - R + d, k.n[2], k.guard[2], k.per[2]) * k.sz;
#pragma unroll
    for (int f = 0; f < NF; f++) q[d][f] = io.in[f][col + off];
  }
  // hv[d] = halo of plane z+d.  The halo of a plane is fetched HD = R+LA
  // planes ahead, in the same iteration as that plane's centre columns, i.e.
  // when the neighbouring tiles fetch the same cells as their own centre, so
  // the shared lines are still in this XCD's L2 (fetched a plane later they
  // have been evicted by the 28 plane-streams every tile moves per plane).
  // Only where the registers are free: the STAR kernels, tiles with more
  // than one halo cell per thread and 1024-thread FP64 tiles keep one plane.
  constexpr int HD = deep_halo<T, STAR, HPT, NT>() ? R + LA : 1;
  T hv[HD][HPT][NF];
#pragma unroll
  for (int d = 0; d < HD; d++) {
    const int off = amap(z0 + d, k.n[2], k.guard[2], k.per[2]) * k.sz;
#pragma unroll
    for (int h = 0; h < HPT; h++)
      if (hok[h])
#pragma unroll
        for (int f = 0; f < NF; f++) hv[d][h][f] = io.in[f][hcol[h] + off];
  }
  // dU of the current plane, prefetched one plane ahead
  T dp[7];
#pragma unroll
  for (int f = 0; f < 7; f++) dp[f] = T(0);
  if (!FIRST && active) {
    const int c0 = col + (z0 + k.guard[2]) * k.sz;
#pragma unroll
    for (int f = 0; f < 7; f++) dp[f] = io.dU[f][c0];
  }

  for (int z = z0; z < z1; z++) {
    const int buf = (z - z0) & 1;
    {
      const int off = amap(z + R + LA, k.n[2], k.guard[2], k.per[2]) * k.sz;
#pragma unroll
      for (int f = 0; f < NF; f++) q[Q + LA - 1][f] = io.in[f][col + off];
    }
#pragma unroll
    for (int f = 0; f < NF; f++) lds[buf][f][slot] = q[R][f];
#pragma unroll
    for (int h = 0; h < HPT; h++)
      if (hok[h])
#pragma unroll
        for (int f = 0; f < NF; f++) lds[buf][f][hslot[h]] = hv[0][h][f];
#pragma unroll
    for (int d = 0; d + 1 < HD; d++)
#pragma unroll
      for (int h = 0; h < HPT; h++)
#pragma unroll
        for (int f = 0; f < NF; f++) hv[d][h][f] = hv[d + 1][h][f];
    if (z + HD < z1) {
      const int off = amap(z + HD, k.n[2], k.guard[2], k.per[2]) * k.sz;
#pragma unroll
      for (int h = 0; h < HPT; h++)
        if (hok[h])
#pragma unroll
          for (int f = 0; f < NF; f++) hv[HD - 1][h][f] = io.in[f][hcol[h] + off];
    }
    T dn[7];
#pragma unroll
    for (int f = 0; f < 7; f++) dn[f] = T(0);
    if (!FIRST && active && z + 1 < z1) {
      const int c1 = col + (z + 1 + k.guard[2]) * k.sz;
#pragma unroll
      for (int f = 0; f < 7; f++) dn[f] = io.dU[f][c1];
    }
    __syncthreads();
    if (active) {
      T d[7][3];
      T u[7];
#pragma unroll
      for (int f = 0; f < NF; f++) {
        const T* p = &lds[buf][f][slot];
        d[f][0] = fma(T(8), p[1] - p[-1], p[-2] - p[2]) * k.c12[0];
        d[f][1] = fma(T(8), p[LX] - p[-LX], p[-2 * LX] - p[2 * LX]) * k.c12[1];
        d[f][2] = fma(T(8), q[3][f] - q[1][f], q[0][f] - q[4][f]) * k.c12[2];
        u[f] = q[R][f];
      }
      const int az = z + k.guard[2];
      const int c = col + az * k.sz;
      stage_cell<T, FIRST, LAST, DIAG, STAR>(k, io, ai, aj, az, c, d, u, dp, As, Bs, ts);
    }
#pragma unroll
    for (int d = 0; d < Q + LA - 1; d++)
#pragma unroll
      for (int f = 0; f < NF; f++) q[d][f] = q[d + 1][f];
#pragma unroll
    for (int f = 0; f < 7; f++) dp[f] = dn[f];
  }
}

template <typename T, int TX, int TY, bool STAR, bool O32>
__global__ void __launch_bounds__(TX * TY)
    ez_ko_tiled(const FlatK<T> k, const FlatIO<T> io, const T tko, const TileMap tm) {
  constexpr int R = 3, LX = TX + 2 * R, LY = TY + 2 * R, NF = 6, NT = TX * TY;
  constexpr int H = 2 * R * (TX + TY);
  constexpr int HPT = (H + NT - 1) / NT;
  constexpr int Q = 2 * R + 1;
  __shared__ T lds[2][NF][LY * LX];

  const int L = xcd_remap(blockIdx.x, gridDim.x);
  const int tx = L % tm.ntx;
  const int rest = L / tm.ntx;
  const int ty = rest % tm.nty;
  const int tz = rest / tm.nty;
  const int i0 = tx * TX, j0 = ty * TY, z0 = tz * tm.zc;
  const int z1 = min(z0 + tm.zc, k.n[2]);
  const int t = threadIdx.x;
  const int lx = t % TX, ly = t / TX;
  const int gi = i0 + lx, gj = j0 + ly;
  const bool active = gi < k.n[0] && gj < k.n[1];
  const int ai = amap(gi, k.n[0], k.guard[0], k.per[0]);
  const int aj = amap(gj, k.n[1], k.guard[1], k.per[1]);
  const int col = ai + aj * k.dims[0];
  const int slot = (ly + R) * LX + (lx + R);

  int hcol[HPT], hslot[HPT];
  bool hok[HPT];
#pragma unroll
  for (int h = 0; h < HPT; h++) {
    const int idx = t + h * NT;
    hok[h] = idx < H;
    int hx = 0, hy = 0;
    if (hok[h]) halo_xy<TX, TY, R>(idx, hx, hy);
    hcol[h] = amap(i0 + hx, k.n[0], k.guard[0], k.per[0]) +
              amap(j0 + hy, k.n[1], k.guard[1], k.per[1]) * k.dims[0];
    hslot[h] = (hy + R) * LX + (hx + R);
  }

  // centre columns of planes z-R .. z+R+LA (see ez_stage_tiled); a second
  // plane of look-ahead measured slower here (1.97 against 1.94 ms at 512^3)
  constexpr int LA = 1;
  T q[Q + LA][NF];
#pragma unroll
  for (int d = 0; d < Q + LA - 1; d++) {
    const int off = amap(z0 - R + d, k.n[2], k.guard[2], k.per[2]) * k.sz;
#pragma unroll
    for (int f = 0; f < NF; f++) q[d][f] = ldf<T, O32>(io.in[f], col + off);
  }
  // halo of planes z .. z+HD-1, fetched with the centre columns
  constexpr int HD = deep_halo<T, STAR, HPT, NT>() ? R + LA : 1;
  T hv[HD][HPT][NF];
#pragma unroll
  for (int d = 0; d < HD; d++) {
    const int off = amap(z0 + d, k.n[2], k.guard[2], k.per[2]) * k.sz;
#pragma unroll
    for (int h = 0; h < HPT; h++)
      if (hok[h])
#pragma unroll
        for (int f = 0; f < NF; f++) hv[d][h][f] = ldf<T, O32>(io.in[f], hcol[h] + off);
  }
  // P passes through unchanged when the caller routes it here (io.in[6]
  // set); the step schedule leaves it out (in[6] = out[6] = null)
  const bool has_p = io.in[6] != nullptr;
  T pc = T(0);
  if (active && has_p) pc = ldf<T, O32>(io.in[6], col + (z0 + k.guard[2]) * k.sz);

  for (int z = z0; z < z1; z++) {
    const int buf = (z - z0) & 1;
    {
      const int off = amap(z + R + LA, k.n[2], k.guard[2], k.per[2]) * k.sz;
#pragma unroll
      for (int f = 0; f < NF; f++) q[Q + LA - 1][f] = ldf<T, O32>(io.in[f], col + off);
    }
    const int az = z + k.guard[2];
    const int c = col + az * k.sz;
#pragma unroll
    for (int f = 0; f < NF; f++) lds[buf][f][slot] = q[R][f];
#pragma unroll
    for (int h = 0; h < HPT; h++)
      if (hok[h])
#pragma unroll
        for (int f = 0; f < NF; f++) lds[buf][f][hslot[h]] = hv[0][h][f];
#pragma unroll
    for (int d = 0; d + 1 < HD; d++)
#pragma unroll
      for (int h = 0; h < HPT; h++)
#pragma unroll
        for (int f = 0; f < NF; f++) hv[d][h][f] = hv[d + 1][h][f];
    if (z + HD < z1) {
      const int off = amap(z + HD, k.n[2], k.guard[2], k.per[2]) * k.sz;
#pragma unroll
      for (int h = 0; h < HPT; h++)
        if (hok[h])
#pragma unroll
          for (int f = 0; f < NF; f++) hv[HD - 1][h][f] = ldf<T, O32>(io.in[f], hcol[h] + off);
    }
    T pn = T(0);
    if (active && has_p && z + 1 < z1) pn = ldf<T, O32>(io.in[6], c + k.sz);
    __syncthreads();
    if (active) {
      T u[7];
#pragma unroll
      for (int f = 0; f < NF; f++) {
        const T* p = &lds[buf][f][slot];
        const T kx = d6v(p[-3], p[-2], p[-1], p[0], p[1], p[2], p[3]);
        const T ky = d6v(p[-3 * LX], p[-2 * LX], p[-LX], p[0], p[LX], p[2 * LX], p[3 * LX]);
        const T kz = d6v(q[0][f], q[1][f], q[2][f], q[3][f], q[4][f], q[5][f], q[6][f]);
        u[f] = fma(kx + ky + kz, k.KOfac, q[R][f]);
      }
      u[6] = pc;
      ko_cell_finish<T, STAR, false, O32>(k, io, ai, aj, az, c, u, tko);
    }
#pragma unroll
    for (int d = 0; d < Q + LA - 1; d++)
#pragma unroll
      for (int f = 0; f < NF; f++) q[d][f] = q[d + 1][f];
    pc = pn;
  }
}

}  // namespace coffee
