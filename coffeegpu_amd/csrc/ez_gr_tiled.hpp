// LDS-tiled, z-marching stage kernel of the Kerr-Schild GR path (gfx950).
//
// One launch per RK stage replaces the point sweep + stage sweep of the
// simple schedule: a 64 x TY workgroup marches its tile column through a
// chunk of z planes and computes E_i, H_i, sqrt(g) D, sqrt(g) B, P at every
// point it needs on the fly (FP64 metric per point), so the only HBM traffic
// is the state U, the LSRK register dU and their updates.
//   x-y derivatives read 11 per-point values staged in double-buffered LDS
//   (tile + 2-cell halo, one barrier per plane);
//   z derivatives read a register queue of the 7 z-differenced values of the
//   planes z-2 .. z+2; the raw state and metric of planes z .. z+2 wait in
//   registers, plane z+3 is loaded one iteration ahead.
// Each point's values come from the same functions as the simple schedule
// (gr_point_values, cks_point) and the cell update is the same gr_stage_core,
// so both schedules give bitwise the same results.
#pragma once

#include "ez_flat_tiled.hpp"
#include "ez_gr.hpp"

// Same contraction as the solver arithmetic of ez_gr.hpp (multiply-adds
// fused within a source expression), so the tiled and the point schedules
// compute bitwise the same values.
#pragma clang fp contract(on)

namespace coffee {

template <typename T, bool FAST = sizeof(T) == 4>
__device__ __forceinline__ CksT<T> gr_metric_at(const GrK<T>& k, int ai, int aj, int ak) {
  return gr_metric<T, FAST>(k.a, gr_pos(k, 0, ai), gr_pos(k, 1, aj), gr_pos(k, 2, ak));
}

// Metric of a point from the table io.W[0..6] (s, l_x, l_y, l_z, alpha,
// sqrt(gamma), bu as gr_metric_table_kernel stores them; r is not stored and
// not needed by gr_point_full).
template <typename T>
__device__ __forceinline__ CksT<T> gr_metric_load(const GrIO<T>& io, int c) {
  CksT<T> m;
  m.s = io.W[0][c];
  m.l[0] = io.W[1][c];
  m.l[1] = io.W[2][c];
  m.l[2] = io.W[3][c];
  m.alpha = io.W[4][c];
  m.sg = io.W[5][c];
  m.bfac = io.W[6][c];
  m.r = 0.0;
  return m;
}

// Fills the metric table io.W[0..6] at every point of the guarded grid with
// the values gr_metric_at computes, so a kernel reading it sees bitwise the
// same metric.
template <typename T>
__global__ void __launch_bounds__(256) gr_metric_table_kernel(const GrK<T> k, const GrIO<T> io) {
  const int total = k.dims[0] * k.dims[1] * k.dims[2];
  for (int c = blockIdx.x * 256 + (int)threadIdx.x; c < total; c += gridDim.x * 256) {
    const int ai = c % k.dims[0], aj = (c / k.dims[0]) % k.dims[1], ak = c / k.sz;
    const CksT<T> m = gr_metric_at(k, ai, aj, ak);
    io.W[0][c] = m.s;
    io.W[1][c] = m.l[0];
    io.W[2][c] = m.l[1];
    io.W[3][c] = m.l[2];
    io.W[4][c] = m.alpha;
    io.W[5][c] = m.sg;
    io.W[6][c] = m.bfac;
  }
}

// LDS slots: 0-2 E_i, 3-5 H_i, 6 sqrt(g) D^x, 7 sqrt(g) D^y, 8 sqrt(g) B^x,
// 9 sqrt(g) B^y, 10 P.
template <typename T>
__device__ __forceinline__ void gr_point_full(const CksT<T>& m, const T u[7], T v[11]) {
  const T D[3] = {u[0], u[1], u[2]};
  const T B[3] = {u[3], u[4], u[5]};
  T E[3], H[3];
  gr_point_values(m, D, B, E, H);
  v[0] = E[0];
  v[1] = E[1];
  v[2] = E[2];
  v[3] = H[0];
  v[4] = H[1];
  v[5] = H[2];
  v[6] = m.sg * D[0];
  v[7] = m.sg * D[1];
  v[8] = m.sg * B[0];
  v[9] = m.sg * B[1];
  v[10] = u[6];
}

// The values differenced along z, in the order of gr_stage_core's ddz.
template <typename T>
__device__ __forceinline__ void gr_point_z(const CksT<T>& m, const T u[7], T v[7]) {
  const T D[3] = {u[0], u[1], u[2]};
  const T B[3] = {u[3], u[4], u[5]};
  T E[3], H[3];
  gr_point_values(m, D, B, E, H);
  v[0] = E[0];
  v[1] = E[1];
  v[2] = H[0];
  v[3] = H[1];
  v[4] = m.sg * D[2];
  v[5] = m.sg * B[2];
  v[6] = u[6];
}

// MT: the halo points' metric is read from the table io.W[0..6] (one plane
// ahead, with their state) instead of being computed in FP64 every plane.
// HD: planes of look-ahead of the x-y halo loads (1 with the table).  With
// HD = 3 the halo of plane z + 3 is fetched in the iteration that fetches the
// centre columns of plane z + 4, close to when the neighbouring tiles fetch
// the same cells as their own centre, so more of them come from L2.
template <typename T, int TY, bool FIRST, bool LAST, bool DIAG, bool MT, int HD = 1>
__global__ void __launch_bounds__(64 * TY)
    gr_stage_tiled(const GrK<T> k, const GrIO<T> io, const T As, const T Bs, const TileMap tm) {
  static_assert(HD >= 1 && (!MT || HD == 1), "halo look-ahead");
  constexpr int TX = 64, R = 2, LX = TX + 2 * R, LY = TY + 2 * R, NL = 11, NT = TX * TY;
  constexpr int H = 2 * R * (TX + TY);
  constexpr int HPT = (H + NT - 1) / NT;
  constexpr bool O32 = !FIRST;  // the first stage measured slower with it (2.33 -> 2.42 ms)
  __shared__ T lds[2][NL][LY * LX];
  // LDS slots differenced along x and along y (the order of ddx / ddy)
  constexpr int SX[7] = {1, 2, 4, 5, 6, 8, 10};
  constexpr int SY[7] = {0, 2, 3, 5, 7, 9, 10};

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
  // guards are filled, so no axis wraps; clamping keeps every access inside
  // the guarded array (guard >= 3 is enforced at context creation)
  const int ai = amap(gi, k.n[0], k.guard[0], 0);
  const int aj = amap(gj, k.n[1], k.guard[1], 0);
  const int col = ai + aj * k.sy;
  const int slot = (ly + R) * LX + (lx + R);

  int hcol[HPT], hslot[HPT], hai[HPT], haj[HPT];
  bool hok[HPT];
#pragma unroll
  for (int h = 0; h < HPT; h++) {
    const int idx = t + h * NT;
    hok[h] = idx < H;
    int hx = 0, hy = 0;
    if (hok[h]) halo_xy<TX, TY, R>(idx, hx, hy);
    hai[h] = amap(i0 + hx, k.n[0], k.guard[0], 0);
    haj[h] = amap(j0 + hy, k.n[1], k.guard[1], 0);
    hcol[h] = hai[h] + haj[h] * k.sy;
    hslot[h] = (hy + R) * LX + (hx + R);
  }

  T qz[5][7];     // z-differenced values of planes z-2 .. z+2
  T rawq[3][7];   // state of planes z .. z+2 at this column
  CksT<T> mq[3];  // metric of planes z .. z+2 at this column
  T inc[7];       // state of plane z+3 (in flight)
#pragma unroll
  for (int d = 0; d < 5; d++) {
    const int ak = amap(z0 - 2 + d, k.n[2], k.guard[2], 0);
    T u[7];
#pragma unroll
    for (int f = 0; f < 7; f++) u[f] = ldf<T, O32>(io.in[f], col + ak * k.sz);
    const CksT<T> m = gr_metric_at(k, ai, aj, ak);
    gr_point_z(m, u, qz[d]);
    if (d >= 2) {
#pragma unroll
      for (int f = 0; f < 7; f++) rawq[d - 2][f] = u[f];
      mq[d - 2] = m;
    }
  }
  {
    const int ak = amap(z0 + 3, k.n[2], k.guard[2], 0);
#pragma unroll
    for (int f = 0; f < 7; f++) inc[f] = ldf<T, O32>(io.in[f], col + ak * k.sz);
  }
  T hq[HD][HPT][7];  // halo state of planes z .. z + HD - 1
  CksT<T> hm[HPT];   // halo metric (MT)
#pragma unroll
  for (int d = 0; d < HD; d++) {
    const int ak = amap(z0 + d, k.n[2], k.guard[2], 0);
#pragma unroll
    for (int h = 0; h < HPT; h++)
      if (hok[h]) {
#pragma unroll
        for (int f = 0; f < 7; f++) hq[d][h][f] = ldf<T, O32>(io.in[f], hcol[h] + ak * k.sz);
        if constexpr (MT) hm[h] = gr_metric_load(io, hcol[h] + ak * k.sz);
      }
  }
  // dU (all but the first stage) and, in the last stage, Un of the current
  // plane, prefetched one plane ahead
  const double px = gr_pos(k, 0, ai), py = gr_pos(k, 1, aj);
  T dp[7], up[7];
#pragma unroll
  for (int f = 0; f < 7; f++) {
    dp[f] = T(0);
    up[f] = T(0);
  }
  if (!FIRST && active) {
    const int c0 = col + (z0 + k.guard[2]) * k.sz;
#pragma unroll
    for (int f = 0; f < 7; f++) dp[f] = ldf<T, O32>(io.dU[f], c0);
    if (LAST && gr_needs_un(k, px, py, gr_pos(k, 2, z0 + k.guard[2]), mq[0].r))
#pragma unroll
      for (int f = 0; f < 7; f++) up[f] = ldf<T, O32>(io.Un[f], c0);
  }

  for (int z = z0; z < z1; z++) {
    const int buf = (z - z0) & 1;
    const int akz = z + k.guard[2];
    T nxt[7];
    {
      const int ak = amap(z + 4, k.n[2], k.guard[2], 0);
#pragma unroll
      for (int f = 0; f < 7; f++) nxt[f] = ldf<T, O32>(io.in[f], col + ak * k.sz);
    }
    T Ec[3];
    {
      T v[11];
      gr_point_full(mq[0], rawq[0], v);
#pragma unroll
      for (int f = 0; f < NL; f++) lds[buf][f][slot] = v[f];
      Ec[0] = v[0];
      Ec[1] = v[1];
      Ec[2] = v[2];
    }
#pragma unroll
    for (int h = 0; h < HPT; h++)
      if (hok[h]) {
        CksT<T> m;
        if constexpr (MT) m = hm[h];
        else m = gr_metric_at(k, hai[h], haj[h], akz);
        T v[11];
        gr_point_full(m, hq[0][h], v);
#pragma unroll
        for (int f = 0; f < NL; f++) lds[buf][f][hslot[h]] = v[f];
      }
#pragma unroll
    for (int d = 0; d + 1 < HD; d++)
#pragma unroll
      for (int h = 0; h < HPT; h++)
#pragma unroll
        for (int f = 0; f < 7; f++) hq[d][h][f] = hq[d + 1][h][f];
    if (z + HD < z1) {
      const int off = (akz + HD) * k.sz;
#pragma unroll
      for (int h = 0; h < HPT; h++)
        if (hok[h]) {
#pragma unroll
          for (int f = 0; f < 7; f++) hq[HD - 1][h][f] = ldf<T, O32>(io.in[f], hcol[h] + off);
          if constexpr (MT) hm[h] = gr_metric_load(io, hcol[h] + off);
        }
    }
    T dn[7], un[7];
#pragma unroll
    for (int f = 0; f < 7; f++) {
      dn[f] = T(0);
      un[f] = T(0);
    }
    if (!FIRST && active && z + 1 < z1) {
      const int c1 = col + (akz + 1) * k.sz;
#pragma unroll
      for (int f = 0; f < 7; f++) dn[f] = ldf<T, O32>(io.dU[f], c1);
      if (LAST && gr_needs_un(k, px, py, gr_pos(k, 2, akz + 1), mq[1].r))
#pragma unroll
        for (int f = 0; f < 7; f++) un[f] = ldf<T, O32>(io.Un[f], c1);
    }
    // the metric of plane z + 3 at this column: in the first and last stage
    // while the plane's loads and the barrier are pending (last stage 2.20 ->
    // 2.08 ms at 384^3), in the middle stages after the stores (2.02 against
    // 2.04 ms the other way)
    CksT<T> m3;
    if constexpr (FIRST || LAST) m3 = gr_metric_at(k, ai, aj, amap(z + 3, k.n[2], k.guard[2], 0));
    __syncthreads();
    if (active) {
      T ddx[7], ddy[7], ddz[7];
#pragma unroll
      for (int q = 0; q < 7; q++) {
        const T* px = &lds[buf][SX[q]][slot];
        ddx[q] = fma(T(8), px[1] - px[-1], px[-2] - px[2]) * k.c12[0];
        const T* py = &lds[buf][SY[q]][slot];
        ddy[q] = fma(T(8), py[LX] - py[-LX], py[-2 * LX] - py[2 * LX]) * k.c12[1];
        ddz[q] = fma(T(8), qz[3][q] - qz[1][q], qz[0][q] - qz[4][q]) * k.c12[2];
      }
      T u[7];
#pragma unroll
      for (int f = 0; f < 7; f++) u[f] = rawq[0][f];
      const int c = col + akz * k.sz;
      gr_stage_core<T, FIRST, LAST, DIAG, O32>(k, io, c, px, py, gr_pos(k, 2, akz), mq[0], ddx, ddy,
                                          ddz, Ec, u, dp, up, As, Bs);
    }
    // advance the queues: plane z+3 enters
    {
      if constexpr (!(FIRST || LAST)) m3 = gr_metric_at(k, ai, aj, amap(z + 3, k.n[2], k.guard[2], 0));
#pragma unroll
      for (int d = 0; d < 4; d++)
#pragma unroll
        for (int f = 0; f < 7; f++) qz[d][f] = qz[d + 1][f];
      gr_point_z(m3, inc, qz[4]);
#pragma unroll
      for (int f = 0; f < 7; f++) {
        rawq[0][f] = rawq[1][f];
        rawq[1][f] = rawq[2][f];
        rawq[2][f] = inc[f];
        inc[f] = nxt[f];
        dp[f] = dn[f];
        up[f] = un[f];
      }
      mq[0] = mq[1];
      mq[1] = mq[2];
      mq[2] = m3;
    }
  }
}

// KO on D, B, P (.cu:537-608) + the metric projections, tiled like the flat
// KO kernel: the 7 fields of the tile + 3-cell halo in double-buffered LDS,
// z neighbours in a register queue.  Same arithmetic as gr_ko_kernel.
template <typename T, int TY>
__global__ void __launch_bounds__(64 * TY)
    gr_ko_tiled(const GrK<T> k, const GrIO<T> io, const TileMap tm) {
  constexpr int TX = 64, R = 3, LX = TX + 2 * R, LY = TY + 2 * R, NF = 7, NT = TX * TY;
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
  const int ai = amap(gi, k.n[0], k.guard[0], 0);
  const int aj = amap(gj, k.n[1], k.guard[1], 0);
  const int col = ai + aj * k.sy;
  const int slot = (ly + R) * LX + (lx + R);

  int hcol[HPT], hslot[HPT];
  bool hok[HPT];
#pragma unroll
  for (int h = 0; h < HPT; h++) {
    const int idx = t + h * NT;
    hok[h] = idx < H;
    int hx = 0, hy = 0;
    if (hok[h]) halo_xy<TX, TY, R>(idx, hx, hy);
    hcol[h] = amap(i0 + hx, k.n[0], k.guard[0], 0) + amap(j0 + hy, k.n[1], k.guard[1], 0) * k.sy;
    hslot[h] = (hy + R) * LX + (hx + R);
  }
  T q[Q + 1][NF];
#pragma unroll
  for (int d = 0; d < Q; d++) {
    const int off = amap(z0 - R + d, k.n[2], k.guard[2], 0) * k.sz;
#pragma unroll
    for (int f = 0; f < NF; f++) q[d][f] = io.in[f][col + off];
  }
  T hv[HPT][NF];
  {
    const int off = amap(z0, k.n[2], k.guard[2], 0) * k.sz;
#pragma unroll
    for (int h = 0; h < HPT; h++)
      if (hok[h])
#pragma unroll
        for (int f = 0; f < NF; f++) hv[h][f] = io.in[f][hcol[h] + off];
  }
  for (int z = z0; z < z1; z++) {
    const int buf = (z - z0) & 1;
    {
      const int off = amap(z + R + 1, k.n[2], k.guard[2], 0) * k.sz;
#pragma unroll
      for (int f = 0; f < NF; f++) q[Q][f] = io.in[f][col + off];
    }
    const int az = z + k.guard[2];
#pragma unroll
    for (int f = 0; f < NF; f++) lds[buf][f][slot] = q[R][f];
#pragma unroll
    for (int h = 0; h < HPT; h++)
      if (hok[h])
#pragma unroll
        for (int f = 0; f < NF; f++) lds[buf][f][hslot[h]] = hv[h][f];
    if (z + 1 < z1) {
      const int off = (az + 1) * k.sz;
#pragma unroll
      for (int h = 0; h < HPT; h++)
        if (hok[h])
#pragma unroll
          for (int f = 0; f < NF; f++) hv[h][f] = io.in[f][hcol[h] + off];
    }
    // the cell's metric while the plane's loads and the barrier are pending
    const CksT<T> m = gr_metric_at<T>(k, ai, aj, az);
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
      gr_project(k, m, u);
      const int c = col + az * k.sz;
#pragma unroll
      for (int f = 0; f < 7; f++) io.out[f][c] = u[f];
    }
#pragma unroll
    for (int d = 0; d < Q; d++)
#pragma unroll
      for (int f = 0; f < NF; f++) q[d][f] = q[d + 1][f];
  }
}

}  // namespace coffee
