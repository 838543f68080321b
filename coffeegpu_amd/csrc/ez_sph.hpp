// Log-spherical 2D (axisymmetric) EZ kernels for gfx950, batched.
//
// Semantics: src/algorithms/field_solver_EZ_spherical.cu (the reference's
// only implementation of this path).  Coordinates x = ln r, y = theta; arrays
// hold upper components, lowered ones (gamma_ii F^i, kernel_compute_ElBl
// .cu:175-198) are formed on the fly from per-run tables r(i), sin(theta)(j).
//
// Every kernel takes an array of run descriptors and a list of work items
// (run, tile): one launch advances all runs of a batch by one substep, so a
// single run is simply a batch of one.  Runs carry their own time and step
// counters in device memory; a run whose step counter reached its target is
// inactive and its items are skipped, so a fixed launch sequence (captured
// once into a HIP graph) can be replayed until every run has finished.
//
// One stage launch does, per interior cell, the reference's rk_step phases
// 1+2 (.cu:200-359), clean_epar / check_eGTb (.cu:361-442, flag-gated),
// boundary_pulsar (.cu:617-688) and, in the last stage, the radial absorbing
// layer (.cu:750-802); U ping-pongs between two buffers.  The theta-axis
// boundary (.cu:690-748) is a separate small launch after each stage, because
// it reads interior rows that other workgroups update.
#pragma once

#include <hip/hip_runtime.h>

#include "ez_flat.hpp"

namespace coffee {

template <typename T>
struct SphRun {
  int dims0, dims1, g0, g1, n0, n1;
  T dx, dy, x0, y0;     // delta and lower (x = ln r, y = theta)
  T idx12, idy12;       // inv_delta / 12
  T inv_dx, inv_dy;     // inv_delta
  T dt, ch2, tau, KOfac, damp_gamma, radius, b0, omega;
  T sigpml, pmllen, xh, rh;
  T tp_start, tp_end, dw0, nT, thp1, thp2;
  T tp_start1, tp_end1, dw1, nT1, thp11, thp21;
  int calc_current, use_edotb, divB_clean, clean_ep, check_egb, KO_geometry, pml_on;
  const T* r;     // r(i) = exp(x_i), dims0 entries
  const T* sth;   // sin(theta_j), dims1 entries
  const T* r3;    // r(i)^3: sqrt(gamma) = r^3 sin(theta)
  const T* ir2;   // 1 / r(i)^2 = gamma^11 = gamma^22
  T inv_tau;
  T* U[2][7];     // [0] = caller's state, [1] = ping-pong partner
  T* dU[7];
  T* Un[7];       // step-start values in the absorbing layer
  T* J[3];
  T* divE;
  T* divB;
  double t;       // time of the step being taken
  int step, nsteps, active;
};

struct SphItem {
  int run, i0, j0;
};

template <typename T>
__device__ __forceinline__ T sph_sqrtg(T r, T s) {
  // metric_sph.h:124-133: sqrt(g11 g22 g33), g11 = g22 = r^2, g33 = (r sin)^2
  const T a = r * s;
  return sqrt((r * r) * (r * r) * (a * a));
}

// dipole_sph_2d (pulsar_impl.hpp:131-139) and wpert_sph_angle (.cu:576-595)
template <typename T>
__device__ __forceinline__ T wpert_angle(T t, T th, T tps, T tpe, T dw0, T nT, T a, T b) {
  if (a > b) {
    T tmp = a;
    a = b;
    b = tmp;
  }
  const T mu = (a + b) / T(2);
  if (t >= tps && t <= tpe && th >= a && th <= b) {
    const T c = cos((th - mu) * T(M_PI) / (b - a));
    const T s = sin((t - tps) * T(M_PI) / (tpe - tps));
    return dw0 * (c * c) * (s * s) * sin((t - tps) * T(2) * T(M_PI) * nT / (tpe - tps));
  }
  return T(0);
}

// metric-normed projections (kernel_Epar_sph / kernel_EgtB_sph)
template <typename T>
__device__ __forceinline__ void sph_project(const SphRun<T>& R, T g1, T g2, T g3, T (&u)[7]) {
  if (R.clean_ep) {
    T B2 = g1 * u[3] * u[3] + g2 * u[4] * u[4] + g3 * u[5] * u[5];
    if (B2 < TinyOf<T>::v) B2 = TinyOf<T>::v;
    const T EB = g1 * u[0] * u[3] + g2 * u[1] * u[4] + g3 * u[2] * u[5];
    const T f = EB / B2;
    u[0] -= f * u[3];
    u[1] -= f * u[4];
    u[2] -= f * u[5];
  }
  if (R.check_egb) {
    T B2 = g1 * u[3] * u[3] + g2 * u[4] * u[4] + g3 * u[5] * u[5];
    if (B2 < TinyOf<T>::v) B2 = TinyOf<T>::v;
    const T E2 = g1 * u[0] * u[0] + g2 * u[1] * u[1] + g3 * u[2] * u[2];
    if (E2 > B2) {
      const T s = sqrt(B2 / E2);
      u[0] *= s;
      u[1] *= s;
      u[2] *= s;
    }
  }
}

// boundary_pulsar_sph on one cell (rotating conductor r < R + dx)
template <typename T>
__device__ __forceinline__ void sph_pulsar(const SphRun<T>& R, T r, T th, T sth, T g1, T g2,
                                           T g3, T t, T (&u)[7]) {
  if (!(r < R.radius + R.dx)) return;
  const T w = R.omega + wpert_angle(t, th, R.tp_start, R.tp_end, R.dw0, R.nT, R.thp1, R.thp2) +
              wpert_angle(t, th, R.tp_start1, R.tp_end1, R.dw1, R.nT1, R.thp11, R.thp21);
  const T r3 = r * r * r;
  const T bxn = R.b0 * (T(2) * cos(th) / r3);
  const T byn = R.b0 * (sth / r3);
  const T bzn = T(0);
  const T v3n = w * r * sth;
  const T exn = v3n * byn;
  const T eyn = -v3n * bxn;
  const T s1 = sqrt(g1), s2 = sqrt(g2);
  T s3 = sqrt(g3);
  if (s3 < TinyOf<T>::v) s3 = TinyOf<T>::v;
  if (fabs(r - R.radius) < R.dx / T(2)) {
    u[3] = bxn / s1;
    u[1] = eyn / s2;
    u[2] = T(0);
  } else if (r < R.radius - TinyOf<T>::v) {
    u[3] = bxn / s1;
    u[4] = byn / s2;
    u[5] = bzn / s3;
    u[0] = exn / s1;
    u[1] = eyn / s2;
    u[2] = T(0);
  }
}

template <int TX, int TY>
__device__ __forceinline__ bool sph_cell(const SphItem& it, int dims0g, int& i, int& j) {
  i = it.i0 + (int)(threadIdx.x % TX);
  j = it.j0 + (int)(threadIdx.x / TX);
  return true;
}

// One LSRK stage at interior cell (i, j) of run R.  F[f] points at the cell's
// value of field f in the buffer read (global memory or an LDS tile) and sy
// is that storage's row stride, so both kernels below run the same
// arithmetic and give bitwise the same results.
template <typename T, bool FIRST, bool LAST, bool DIAG>
__device__ __forceinline__ void sph_stage_point(const SphRun<T>& R, int i, int j, int src, T As,
                                                T Bs, T cs, const T* const* F, int sy) {
  const int c = i + j * R.dims0;
  const T* Ex = F[0];
  const T* Ey = F[1];
  const T* Ez = F[2];
  const T* Bx = F[3];
  const T* By = F[4];
  const T* Bz = F[5];
  const T* P = F[6];
  // metric at the cell and its neighbours
  T rr[5], ss[5];
#pragma unroll
  for (int q = 0; q < 5; q++) {
    rr[q] = R.r[i - 2 + q];
    ss[q] = R.sth[j - 2 + q];
  }
  const T r = rr[2], sth = ss[2];
  const T g1 = r * r, g2 = r * r, rs = r * sth, g3 = rs * rs;
  // sqrt(gamma) = r^3 |sin(theta)| from the per-run tables (metric_sph.h:124-133
  // takes the square root of g11 g22 g33, which is positive also in the guard
  // rows beyond the axis, where sin(theta) < 0)
  T r3[5];
#pragma unroll
  for (int q = 0; q < 5; q++) r3[q] = R.r3[i - 2 + q];
  T gm = r3[2] * fabs(sth);
  if (gm < TinyOf<T>::v) gm = TinyOf<T>::v;
  const T igm = T(1) / gm;
  // lowered components at x-neighbours (row j) and y-neighbours (column i)
  auto dxl = [&](const T* f, int comp) {
    // d/dx of gamma_cc F (gamma depends on r only for c = 1, 2; on r and theta for c = 3)
    T v[5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
      const T rq = rr[q];
      const T gq = comp < 2 ? rq * rq : (rq * sth) * (rq * sth);
      v[q] = q == 2 ? T(0) : gq * f[q - 2];
    }
    return fma(T(8), v[3] - v[1], v[0] - v[4]) * R.idx12;
  };
  auto dyl = [&](const T* f, int comp) {
    T v[5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
      const T gq = comp < 2 ? r * r : (r * ss[q]) * (r * ss[q]);
      v[q] = q == 2 ? T(0) : gq * f[(q - 2) * sy];
    }
    return fma(T(8), v[3] - v[1], v[0] - v[4]) * R.idy12;
  };
  const T rotBx = dyl(Bz, 2) * igm;
  const T rotEx = dyl(Ez, 2) * igm;
  const T rotBy = -dxl(Bz, 2) * igm;
  const T rotBz = (dxl(By, 1) - dyl(Bx, 0)) * igm;
  const T rotEy = -dxl(Ez, 2) * igm;
  const T rotEz = (dxl(Ey, 1) - dyl(Ex, 0)) * igm;
  // div4_sph (.cu:99-124): d/dx(sqrt(g) F^x) + d/dy(sqrt(g) F^y), over sqrt(g)
  auto divf = [&](const T* fx, const T* fy) {
    T vx[5], vy[5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
      vx[q] = q == 2 ? T(0) : (r3[q] * fabs(sth)) * fx[q - 2];
      vy[q] = q == 2 ? T(0) : (r3[2] * fabs(ss[q])) * fy[(q - 2) * sy];
    }
    const T tx = fma(T(8), vx[3] - vx[1], vx[0] - vx[4]) * R.idx12;
    const T ty = fma(T(8), vy[3] - vy[1], vy[0] - vy[4]) * R.idy12;
    return (tx + ty) * igm;
  };
  const T divE = divf(Ex, Ey);
  const T divB = divf(Bx, By);
  T u[7] = {Ex[0], Ey[0], Ez[0], Bx[0], By[0], Bz[0], P[0]};
  const T elx = g1 * u[0], ely = g2 * u[1], elz = g3 * u[2];
  const T blx = g1 * u[3], bly = g2 * u[4], blz = g3 * u[5];
  T B2 = u[3] * blx + u[4] * bly + u[5] * blz;
  if (B2 < TinyOf<T>::v) B2 = TinyOf<T>::v;
  const T iB2 = T(1) / B2;
  T Jx = 0, Jy = 0, Jz = 0;
  if (R.calc_current) {
    T Jp = (blx * rotBx + bly * rotBy + blz * rotBz) - (elx * rotEx + ely * rotEy + elz * rotEz);
    if (R.use_edotb) {
      const T E2 = u[0] * elx + u[1] * ely + u[2] * elz;
      const T chi2 = B2 - E2;
      const T EdotB = u[0] * blx + u[1] * bly + u[2] * blz;
      const T E02 = T(0.5) * (sqrt(chi2 * chi2 + T(4) * EdotB * EdotB) - chi2);
      Jp += R.damp_gamma * EdotB / R.dt;
      const T a = divE * igm / (E02 + B2);
      const T JpB = Jp * iB2;
      Jx = a * (ely * blz - elz * bly) + JpB * u[3];
      Jy = a * (elz * blx - elx * blz) + JpB * u[4];
      Jz = a * (elx * bly - ely * blx) + JpB * u[5];
    } else {
      const T a = divE * igm;
      Jx = (a * (ely * blz - elz * bly) + Jp * u[3]) * iB2;
      Jy = (a * (elz * blx - elx * blz) + Jp * u[4]) * iB2;
      Jz = (a * (elx * bly - ely * blx) + Jp * u[5]) * iB2;
    }
  }
  T Px = 0, Py = 0;
  if (R.divB_clean) {
    const T ig = R.ir2[i];
    Px = fma(T(8), P[1] - P[-1], P[-2] - P[2]) * R.idx12 * ig;
    Py = fma(T(8), P[sy] - P[-sy], P[-2 * sy] - P[2 * sy]) * R.idy12 * ig;
  }
  if (DIAG) {
    R.J[0][c] = Jx;
    R.J[1][c] = Jy;
    R.J[2][c] = Jz;
    R.divE[c] = divE;
    R.divB[c] = divB;
  }
  const T dt = R.dt;
  T du[7];
  du[0] = dt * (rotBx - Jx);
  du[1] = dt * (rotBy - Jy);
  du[2] = dt * (rotBz - Jz);
  du[3] = -dt * (rotEx + Px);
  du[4] = -dt * (rotEy + Py);
  du[5] = -dt * rotEz;
  du[6] = -dt * R.ch2 * divB - u[6] * R.inv_tau;
  if (!FIRST) {
#pragma unroll
    for (int q = 0; q < 7; q++) du[q] = fma(As, R.dU[q][c], du[q]);
  }
  if (!LAST) {
#pragma unroll
    for (int q = 0; q < 7; q++) R.dU[q][c] = du[q];
  }
  const T x = R.x0 + R.dx * T(i - R.g0);
  if (FIRST && R.pml_on && x > R.xh) {
#pragma unroll
    for (int q = 0; q < 7; q++) R.Un[q][c] = u[q];
  }
#pragma unroll
  for (int q = 0; q < 7; q++) u[q] = fma(Bs, du[q], u[q]);
  sph_project(R, g1, g2, g3, u);
  const T th = R.y0 + R.dy * T(j - R.g1);
  const T tstage = (T)R.t + cs * dt;
  sph_pulsar(R, r, th, sth, g1, g2, g3, tstage, u);
  if (LAST && R.pml_on && x > R.xh) {
    T sig = T(0);
    if (r > R.rh) {
      const T q = (r - R.rh) / R.pmllen;
      sig = R.sigpml * (q * q * q);
    }
    if (sig > TinyOf<T>::v) {
      const T e = exp(-sig), f = (T(1) - e) / sig;
#pragma unroll
      for (int q = 0; q < 6; q++) {
        const T un = R.Un[q][c];
        u[q] = e * un + f * (u[q] - un);
      }
    }
  }
  T* const* out = R.U[src ^ 1];
#pragma unroll
  for (int q = 0; q < 7; q++) out[q][c] = u[q];
}

// One LSRK stage for every active run, stencil operands read from global
// memory (through the caches).  src = buffer index read (0/1).
template <typename T, int TX, int TY, bool FIRST, bool LAST, bool DIAG>
__global__ void __launch_bounds__(TX * TY)
    sph_stage_kernel(SphRun<T>* __restrict__ runs, const SphItem* __restrict__ items, int nitems,
                     int src, T As, T Bs, T cs) {
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const SphItem it = items[w];
    const SphRun<T>& R = runs[it.run];
    if (!R.active) continue;
    const int li = (int)(threadIdx.x % TX), lj = (int)(threadIdx.x / TX);
    const int ii = it.i0 + li, jj = it.j0 + lj;  // interior-relative
    if (ii >= R.n0 || jj >= R.n1) continue;
    const int i = ii + R.g0, j = jj + R.g1;
    const int c = i + j * R.dims0;
    const T* F[7];
#pragma unroll
    for (int f = 0; f < 7; f++) F[f] = R.U[src][f] + c;
    sph_stage_point<T, FIRST, LAST, DIAG>(R, i, j, src, As, Bs, cs, F, R.dims0);
  }
}

// Stage a TX x TY tile of all 7 fields plus a RH-cell halo of run R into LDS
// (rows of TX + 2 RH).  Cells beyond the array (partial tiles at the far
// edges) are clamped to the last column/row; they feed no result.
template <typename T, int TX, int TY, int RH>
__device__ __forceinline__ void sph_fill_tile(const SphRun<T>& R, const SphItem& it, int src,
                                              T (*lds)[(TX + 2 * RH) * (TY + 2 * RH)]) {
  constexpr int LXW = TX + 2 * RH, NCELL = LXW * (TY + 2 * RH);
  for (int idx = threadIdx.x; idx < NCELL; idx += TX * TY) {
    const int lx = idx % LXW, ly = idx / LXW;
    int gi = it.i0 + lx - RH + R.g0, gj = it.j0 + ly - RH + R.g1;
    gi = gi > R.dims0 - 1 ? R.dims0 - 1 : gi;
    gj = gj > R.dims1 - 1 ? R.dims1 - 1 : gj;
    const int g = gi + gj * R.dims0;
#pragma unroll
    for (int f = 0; f < 7; f++) lds[f][idx] = R.U[src][f][g];
  }
}

// The same stage with the tile and its 2-cell halo staged in LDS: one load
// per field and cell instead of one per stencil tap.
template <typename T, int TX, int TY, bool FIRST, bool LAST, bool DIAG>
__global__ void __launch_bounds__(TX * TY)
    sph_stage_lds_kernel(SphRun<T>* __restrict__ runs, const SphItem* __restrict__ items,
                         int nitems, int src, T As, T Bs, T cs) {
  constexpr int RH = 2, LXW = TX + 2 * RH;
  __shared__ T lds[7][LXW * (TY + 2 * RH)];
  const int li = (int)(threadIdx.x % TX), lj = (int)(threadIdx.x / TX);
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const SphItem it = items[w];
    const SphRun<T>& R = runs[it.run];
    if (!R.active) continue;  // uniform: the whole workgroup works on one item
    sph_fill_tile<T, TX, TY, RH>(R, it, src, lds);
    __syncthreads();
    const int ii = it.i0 + li, jj = it.j0 + lj;
    if (ii < R.n0 && jj < R.n1) {
      const int slot = (lj + RH) * LXW + li + RH;
      const T* F[7];
#pragma unroll
      for (int f = 0; f < 7; f++) F[f] = &lds[f][slot];
      sph_stage_point<T, FIRST, LAST, DIAG>(R, ii + R.g0, jj + R.g1, src, As, Bs, cs, F, LXW);
    }
    __syncthreads();
  }
}

// Kreiss-Oliger on all 7 fields (kernel_KO_step1/2_sph, .cu:444-555):
// KO_2d6 (sqrt(g)-weighted, KO_geometry) or KO_2d0, then projections and the
// pulsar boundary at t + dt.  F, sy as in sph_stage_point.
template <typename T>
__device__ __forceinline__ void sph_ko_point(const SphRun<T>& R, int i, int j, int src,
                                             const T* const* F, int sy) {
  const int c = i + j * R.dims0;
  T rr[7], ss[7];
#pragma unroll
  for (int q = 0; q < 7; q++) {
    rr[q] = R.r[i - 3 + q];
    ss[q] = R.sth[j - 3 + q];
  }
  const T r = rr[3], sth = ss[3];
  const T r3c = R.r3[i];
  T gm = r3c * fabs(sth);
  if (gm < TinyOf<T>::v) gm = TinyOf<T>::v;
  T wx[7], wy[7];
#pragma unroll
  for (int q = 0; q < 7; q++) {
    wx[q] = R.KO_geometry ? R.r3[i - 3 + q] * fabs(sth) : T(1);
    wy[q] = R.KO_geometry ? r3c * fabs(ss[q]) : T(1);
  }
  const T iden = R.KO_geometry ? T(1) / gm : T(1);
  T u[7];
#pragma unroll
  for (int f = 0; f < 7; f++) {
    const T* p = F[f];
    const T kx = d6v(wx[0] * p[-3], wx[1] * p[-2], wx[2] * p[-1], wx[3] * p[0], wx[4] * p[1],
                     wx[5] * p[2], wx[6] * p[3]);
    const T ky = d6v(wy[0] * p[-3 * sy], wy[1] * p[-2 * sy], wy[2] * p[-sy], wy[3] * p[0],
                     wy[4] * p[sy], wy[5] * p[2 * sy], wy[6] * p[3 * sy]);
    u[f] = fma((kx + ky) * iden, R.KOfac, p[0]);
  }
  const T g1 = r * r, g2 = r * r, rs = r * sth, g3 = rs * rs;
  sph_project(R, g1, g2, g3, u);
  const T th = R.y0 + R.dy * T(j - R.g1);
  sph_pulsar(R, r, th, sth, g1, g2, g3, (T)R.t + R.dt, u);
  T* const* out = R.U[src ^ 1];
#pragma unroll
  for (int q = 0; q < 7; q++) out[q][c] = u[q];
}

template <typename T, int TX, int TY>
__global__ void __launch_bounds__(TX * TY)
    sph_ko_kernel(SphRun<T>* __restrict__ runs, const SphItem* __restrict__ items, int nitems,
                  int src) {
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const SphItem it = items[w];
    const SphRun<T>& R = runs[it.run];
    if (!R.active) continue;
    const int ii = it.i0 + (int)(threadIdx.x % TX), jj = it.j0 + (int)(threadIdx.x / TX);
    if (ii >= R.n0 || jj >= R.n1) continue;
    const int i = ii + R.g0, j = jj + R.g1;
    const int c = i + j * R.dims0;
    const T* F[7];
#pragma unroll
    for (int f = 0; f < 7; f++) F[f] = R.U[src][f] + c;
    sph_ko_point<T>(R, i, j, src, F, R.dims0);
  }
}

template <typename T, int TX, int TY>
__global__ void __launch_bounds__(TX * TY)
    sph_ko_lds_kernel(SphRun<T>* __restrict__ runs, const SphItem* __restrict__ items, int nitems,
                      int src) {
  constexpr int RH = 3, LXW = TX + 2 * RH;
  __shared__ T lds[7][LXW * (TY + 2 * RH)];
  const int li = (int)(threadIdx.x % TX), lj = (int)(threadIdx.x / TX);
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const SphItem it = items[w];
    const SphRun<T>& R = runs[it.run];
    if (!R.active) continue;  // uniform: the whole workgroup works on one item
    sph_fill_tile<T, TX, TY, RH>(R, it, src, lds);
    __syncthreads();
    const int ii = it.i0 + li, jj = it.j0 + lj;
    if (ii < R.n0 && jj < R.n1) {
      const int slot = (lj + RH) * LXW + li + RH;
      const T* F[7];
#pragma unroll
      for (int f = 0; f < 7; f++) F[f] = &lds[f][slot];
      sph_ko_point<T>(R, ii + R.g0, jj + R.g1, src, F, LXW);
    }
    __syncthreads();
  }
}

// theta-axis boundary (kernel_boundary_axis_sph, .cu:690-748) on buffer `buf`
// of every active run: one thread per interior radial index.
template <typename T>
__global__ void __launch_bounds__(256)
    sph_axis_kernel(SphRun<T>* __restrict__ runs, int nruns, int buf) {
  const int run = blockIdx.y;
  if (run >= nruns) return;
  const SphRun<T>& R = runs[run];
  if (!R.active) return;
  for (int ii = blockIdx.x * blockDim.x + threadIdx.x; ii < R.n0; ii += gridDim.x * blockDim.x) {
    const int i = ii + R.g0;
    const int s = R.dims0;
    T* const* U = R.U[buf];
    int j = R.g1;
    int c = i + j * s;
    T y = R.y0 + R.dy * T(j - R.g1);
    if (fabs(y) < R.dy / T(2)) {
      U[0][c] = U[0][c + s];
      U[1][c] = T(0);
      U[2][c] = T(0);
      U[3][c] = U[3][c + s];
      U[4][c] = T(0);
      U[5][c] = T(0);
      U[6][c] = U[6][c + s];
      for (int l = 1; l <= 3; ++l) {
        U[0][c - l * s] = U[0][c + l * s];
        U[1][c - l * s] = -U[1][c + l * s];
        U[2][c - l * s] = -U[2][c + l * s];
        U[3][c - l * s] = U[3][c + l * s];
        U[4][c - l * s] = -U[4][c + l * s];
        U[5][c - l * s] = -U[5][c + l * s];
      }
    }
    j = R.dims1 - R.g1;
    c = i + j * s;
    y = R.y0 + R.dy * T(j - R.g1);
    if (fabs(y - T(M_PI)) < R.dy / T(2)) {
      U[0][c] = U[0][c - s];
      U[1][c] = T(0);
      U[2][c] = T(0);
      U[3][c] = U[3][c - s];
      U[4][c] = T(0);
      U[5][c] = T(0);
      U[6][c] = U[6][c - s];
      for (int l = 1; l <= 3; ++l) {
        U[0][c + l * s] = U[0][c - l * s];
        U[1][c + l * s] = -U[1][c - l * s];
        U[2][c + l * s] = -U[2][c - l * s];
        U[3][c + l * s] = U[3][c - l * s];
        U[4][c + l * s] = -U[4][c - l * s];
        U[5][c + l * s] = -U[5][c - l * s];
      }
    }
  }
}

// Copy every guard cell (radial guard columns and theta guard rows) of the
// caller's buffer into the partner buffer, so frozen guards are seen by the
// stages that read the partner.
template <typename T>
__global__ void __launch_bounds__(256)
    sph_guard_copy_kernel(SphRun<T>* __restrict__ runs, int nruns) {
  const int run = blockIdx.y;
  if (run >= nruns) return;
  const SphRun<T>& R = runs[run];
  if (!R.active) return;
  const long total = (long)R.dims0 * R.dims1;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t % R.dims0), j = (int)(t / R.dims0);
    const bool guard = i < R.g0 || i >= R.dims0 - R.g0 || j < R.g1 || j >= R.dims1 - R.g1;
    if (!guard) continue;
#pragma unroll
    for (int q = 0; q < 7; q++) R.U[1][q][t] = R.U[0][q][t];
  }
}

// End of a step: t += dt, step += 1; a run that reached its target step
// count becomes inactive.
template <typename T>
__global__ void sph_advance_kernel(SphRun<T>* __restrict__ runs, int nruns) {
  const int run = blockIdx.x * blockDim.x + threadIdx.x;
  if (run >= nruns) return;
  SphRun<T>& R = runs[run];
  if (!R.active) return;
  R.t = (double)((T)R.t + R.dt);
  R.step += 1;
  if (R.nsteps > 0 && R.step >= R.nsteps) R.active = 0;
}

// Spin-down luminosity through the sphere r = rad[run] (a diagnostic of this
// project; the reference has none): L = 2 pi r^2 sum_j S_r sin(theta_j) dtheta,
// S_r = E_th B_ph - E_ph B_th the radial Poynting flux in orthonormal
// components (E_th = r E^theta, E_ph = r sin(theta) E^phi of the upper
// components stored for x = ln r), r^2 S_r linearly interpolated in x between
// the two nodes around ln(rad).  One workgroup per run; out[run] in double.
template <typename T>
__global__ void __launch_bounds__(256)
    sph_luminosity_kernel(const SphRun<T>* __restrict__ runs, const double* __restrict__ rad,
                          double* __restrict__ out) {
  __shared__ double red[256];
  const SphRun<T>& R = runs[blockIdx.x];
  const double fi = (log(rad[blockIdx.x]) - (double)R.x0) / (double)R.dx + R.g0;
  int i0 = (int)floor(fi);
  i0 = i0 < R.g0 ? R.g0 : (i0 > R.g0 + R.n0 - 2 ? R.g0 + R.n0 - 2 : i0);
  double w = fi - i0;
  w = w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w);
  double acc = 0.0;
  for (int j = R.g1 + threadIdx.x; j < R.g1 + R.n1; j += blockDim.x) {
    const double s = (double)R.sth[j];
    double f2[2];
    for (int q = 0; q < 2; q++) {
      const int i = i0 + q;
      const int c = i + j * R.dims0;
      const double r = (double)R.r[i];
      const double Eth = r * (double)R.U[0][1][c], Eph = r * s * (double)R.U[0][2][c];
      const double Bth = r * (double)R.U[0][4][c], Bph = r * s * (double)R.U[0][5][c];
      f2[q] = r * r * (Eth * Bph - Eph * Bth);
    }
    acc += ((1.0 - w) * f2[0] + w * f2[1]) * s;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int m = 128; m > 0; m >>= 1) {
    if ((int)threadIdx.x < m) red[threadIdx.x] += red[threadIdx.x + m];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[blockIdx.x] = 2.0 * M_PI * red[0] * (double)R.dy;
}

}  // namespace coffee
