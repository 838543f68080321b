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

// Global-memory access through run descriptors.  The field and table pointers
// are read from the descriptor array, so the compiler cannot tell that they
// address global memory and would emit flat instructions, which also count
// against the LDS/scalar wait counter: every wait for an LDS read would then
// wait for all outstanding HBM loads too, and a prefetch issued before a
// barrier would be waited for at the barrier.  These cast to the global
// address space.  gld_at / gst_at address cell `idx` of an array by a 32-bit
// byte offset (every spherical array is < 4 GiB, checked when a run is set
// up), so an access is a scalar base plus one vector offset shared by all
// fields, without 64-bit address arithmetic per field.
template <typename T>
using GlobalCPtr = const __attribute__((address_space(1))) T*;
template <typename T>
using GlobalPtr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ T gld(const T* p) {
  return *(GlobalCPtr<T>)(p);
}
template <typename T>
__device__ __forceinline__ void gst(T* p, T v) {
  *(GlobalPtr<T>)(p) = v;
}
template <typename T>
__device__ __forceinline__ T gld_at(const T* base, unsigned idx) {
  return *(GlobalCPtr<T>)((const char*)base + idx * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ void gst_at(T* base, unsigned idx, T v) {
  *(GlobalPtr<T>)((char*)base + idx * (unsigned)sizeof(T)) = v;
}

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

// Stencil operands of one cell.  x(f, d) is field f at radial offset d and
// y(f, d) at theta offset d (|d| <= the kernel's halo).  StrideTaps reads
// them from global memory or an LDS tile; QueueTaps takes the theta column
// from a register queue and the radial neighbours from an LDS row.  Every
// kernel below feeds the same per-cell arithmetic through one of these, so
// all of them give bitwise the same results.
template <typename T, bool GLOBAL>
struct StrideTaps {
  const T* const* F;  // F[f] = the cell's value of field f
  int sy;             // row stride of that storage
  __device__ __forceinline__ T x(int f, int d) const { return GLOBAL ? gld(F[f] + d) : F[f][d]; }
  __device__ __forceinline__ T y(int f, int d) const {
    return GLOBAL ? gld(F[f] + d * sy) : F[f][d * sy];
  }
};

template <typename T, int RH, int LW>
struct QueueTaps {
  const T (&q)[7][2 * RH + 1];  // q[f][RH + d] = field f at theta offset d
  const T (&row)[7][LW];        // the cell's row, radial halo included
  int lx;                       // the cell's slot in `row`
  __device__ __forceinline__ T x(int f, int d) const { return d == 0 ? q[f][RH] : row[f][lx + d]; }
  __device__ __forceinline__ T y(int f, int d) const { return q[f][RH + d]; }
};

// Metric tables around cell (i, j): r, r^3 at radial offsets -H..H, 1/r^2 at
// the cell (stage only, H = 2), sin(theta) at theta offsets -H..H.
template <typename T, int H>
struct SphGeo {
  T r[2 * H + 1], r3[2 * H + 1], s[2 * H + 1], ir2;
  __device__ __forceinline__ void radial(const SphRun<T>& R, int i) {
#pragma unroll
    for (int q = 0; q <= 2 * H; q++) {
      r[q] = gld(R.r + i - H + q);
      r3[q] = gld(R.r3 + i - H + q);
    }
    ir2 = H == 2 ? gld(R.ir2 + i) : T(0);
  }
  __device__ __forceinline__ void polar(const SphRun<T>& R, int j) {
#pragma unroll
    for (int q = 0; q <= 2 * H; q++) s[q] = gld(R.sth + j - H + q);
  }
  // from a copy of the sin(theta) table, sl[d] = sin(theta_{j + d})
  __device__ __forceinline__ void polar(const T* sl) {
#pragma unroll
    for (int q = 0; q <= 2 * H; q++) s[q] = sl[q - H];
  }
};

// One LSRK stage at interior cell (i, j) of run R: S supplies the stencil
// operands, G the metric tables, dprev the cell's LSRK register (ignored in
// the first stage).
template <typename T, bool FIRST, bool LAST, bool DIAG, typename Taps>
__device__ __forceinline__ void sph_stage_point(const SphRun<T>& R, int i, int j, int src, T As,
                                                T Bs, T cs, const Taps& S, const SphGeo<T, 2>& G,
                                                const T (&dprev)[7]) {
  const unsigned c = (unsigned)(i + j * R.dims0);
  enum { EX, EY, EZ, BX, BY, BZ, PP };
  const T r = G.r[2], sth = G.s[2];
  const T g1 = r * r, g2 = r * r, rs = r * sth, g3 = rs * rs;
  // sqrt(gamma) = r^3 |sin(theta)| from the per-run tables (metric_sph.h:124-133
  // takes the square root of g11 g22 g33, which is positive also in the guard
  // rows beyond the axis, where sin(theta) < 0)
  T gm = G.r3[2] * fabs(sth);
  if (gm < TinyOf<T>::v) gm = TinyOf<T>::v;
  const T igm = T(1) / gm;
  // lowered components at x-neighbours (row j) and y-neighbours (column i)
  auto dxl = [&](int f, int comp) {
    // d/dx of gamma_cc F (gamma depends on r only for c = 1, 2; on r and theta for c = 3)
    T v[5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
      const T rq = G.r[q];
      const T gq = comp < 2 ? rq * rq : (rq * sth) * (rq * sth);
      v[q] = q == 2 ? T(0) : gq * S.x(f, q - 2);
    }
    return fma(T(8), v[3] - v[1], v[0] - v[4]) * R.idx12;
  };
  auto dyl = [&](int f, int comp) {
    T v[5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
      const T gq = comp < 2 ? r * r : (r * G.s[q]) * (r * G.s[q]);
      v[q] = q == 2 ? T(0) : gq * S.y(f, q - 2);
    }
    return fma(T(8), v[3] - v[1], v[0] - v[4]) * R.idy12;
  };
  const T rotBx = dyl(BZ, 2) * igm;
  const T rotEx = dyl(EZ, 2) * igm;
  const T rotBy = -dxl(BZ, 2) * igm;
  const T rotBz = (dxl(BY, 1) - dyl(BX, 0)) * igm;
  const T rotEy = -dxl(EZ, 2) * igm;
  const T rotEz = (dxl(EY, 1) - dyl(EX, 0)) * igm;
  // div4_sph (.cu:99-124): d/dx(sqrt(g) F^x) + d/dy(sqrt(g) F^y), over sqrt(g)
  auto divf = [&](int fx, int fy) {
    T vx[5], vy[5];
#pragma unroll
    for (int q = 0; q < 5; q++) {
      vx[q] = q == 2 ? T(0) : (G.r3[q] * fabs(sth)) * S.x(fx, q - 2);
      vy[q] = q == 2 ? T(0) : (G.r3[2] * fabs(G.s[q])) * S.y(fy, q - 2);
    }
    const T tx = fma(T(8), vx[3] - vx[1], vx[0] - vx[4]) * R.idx12;
    const T ty = fma(T(8), vy[3] - vy[1], vy[0] - vy[4]) * R.idy12;
    return (tx + ty) * igm;
  };
  const T divE = divf(EX, EY);
  const T divB = divf(BX, BY);
  T u[7];
#pragma unroll
  for (int f = 0; f < 7; f++) u[f] = S.x(f, 0);
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
    const T ig = G.ir2;
    Px = fma(T(8), S.x(PP, 1) - S.x(PP, -1), S.x(PP, -2) - S.x(PP, 2)) * R.idx12 * ig;
    Py = fma(T(8), S.y(PP, 1) - S.y(PP, -1), S.y(PP, -2) - S.y(PP, 2)) * R.idy12 * ig;
  }
  if (DIAG) {
    gst_at(R.J[0], c, Jx);
    gst_at(R.J[1], c, Jy);
    gst_at(R.J[2], c, Jz);
    gst_at(R.divE, c, divE);
    gst_at(R.divB, c, divB);
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
    for (int q = 0; q < 7; q++) du[q] = fma(As, dprev[q], du[q]);
  }
  if (!LAST) {
#pragma unroll
    for (int q = 0; q < 7; q++) gst_at(R.dU[q], c, du[q]);
  }
  const T x = R.x0 + R.dx * T(i - R.g0);
  if (FIRST && R.pml_on && x > R.xh) {
#pragma unroll
    for (int q = 0; q < 7; q++) gst_at(R.Un[q], c, u[q]);
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
        const T un = gld_at(R.Un[q], c);
        u[q] = e * un + f * (u[q] - un);
      }
    }
  }
  T* const* out = R.U[src ^ 1];
#pragma unroll
  for (int q = 0; q < 7; q++) gst_at(out[q], c, u[q]);
}

// The LSRK register of cell c (not read in the first stage).
template <typename T, bool FIRST>
__device__ __forceinline__ void sph_load_dprev(const SphRun<T>& R, unsigned c, T (&dp)[7]) {
#pragma unroll
  for (int q = 0; q < 7; q++) dp[q] = FIRST ? T(0) : gld_at(R.dU[q], c);
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
    const unsigned c = (unsigned)(i + j * R.dims0);
    const T* F[7];
#pragma unroll
    for (int f = 0; f < 7; f++) F[f] = R.U[src][f] + c;
    SphGeo<T, 2> G;
    G.radial(R, i);
    G.polar(R, j);
    T dp[7];
    sph_load_dprev<T, FIRST>(R, c, dp);
    sph_stage_point<T, FIRST, LAST, DIAG>(R, i, j, src, As, Bs, cs, StrideTaps<T, true>{F, R.dims0}, G,
                                          dp);
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
    for (int f = 0; f < 7; f++) lds[f][idx] = gld(R.U[src][f] + g);
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
      const int i = ii + R.g0, j = jj + R.g1;
      const int slot = (lj + RH) * LXW + li + RH;
      const T* F[7];
#pragma unroll
      for (int f = 0; f < 7; f++) F[f] = &lds[f][slot];
      SphGeo<T, 2> G;
      G.radial(R, i);
      G.polar(R, j);
      T dp[7];
      sph_load_dprev<T, FIRST>(R, (unsigned)(i + j * R.dims0), dp);
      sph_stage_point<T, FIRST, LAST, DIAG>(R, i, j, src, As, Bs, cs, StrideTaps<T, false>{F, LXW}, G,
                                            dp);
    }
    __syncthreads();
  }
}

// Theta-marching traversal shared by the march kernels.  A workgroup of TX
// lanes owns a band of TX radial cells and walks up to NJ rows of it in
// theta.  Each lane keeps the 2 RH + 1 theta rows of its column in a register
// queue; the centre row, with the band's RH-cell radial halo fetched by the
// edge lanes, goes through a double-buffered LDS row for the radial taps.
// Loads are ordered so that no wait drains a prefetch (the vector memory
// counter completes in order): per row, the halo of the next row, the cell's
// own operands, then the row entering the queue two rows ahead (into one of
// two alternating register sets); after the cell, the next row's LDS buffer
// (centre and halo) is filled.  Every value is read from memory once per band.
// `cell(q, row, lx, i, j, live, sl)` runs the per-cell work; lanes beyond the
// run's last column (live = false) only help fill the rows.  Column and row
// indices are clamped to the array; clamped values feed no result.
template <typename T, int TX, int NJ, int RH, typename Cell>
__device__ __forceinline__ void sph_march(const SphRun<T>& R, const SphItem& it, int src,
                                          T (&lds)[2][7][TX + 2 * RH], T (&sl)[NJ + 2 * RH],
                                          Cell&& cell) {
  static_assert(TX > 2 * RH, "band narrower than its halo");
  constexpr int NQ = 2 * RH + 1;
  const int li = (int)threadIdx.x;
  const int s = R.dims0, jmax = R.dims1 - 1;
  const int ii = it.i0 + li;
  const bool live = ii < R.n0;
  const int gi = min(ii + R.g0, s - 1);
  // edge lanes also fetch the radial halo: lanes 0..RH-1 the columns left of
  // the band (slots 0..RH-1), lanes TX-RH..TX-1 those right of it
  int hc = -1, hslot = 0;
  if (li < RH) {
    hc = it.i0 + R.g0 - RH + li;  // >= 0: the radial guard is >= RH
    hslot = li;
  } else if (li >= TX - RH) {
    hc = min(it.i0 + R.g0 + li + RH, s - 1);
    hslot = li + 2 * RH;
  }
  const T* const* U = R.U[src];
  const int jc = it.j0 + R.g1;
  const int nj = min(NJ, R.n1 - it.j0);
  // sin(theta) of rows jc - RH .. jc + nj - 1 + RH (within the theta guards)
  for (int t = li; t < nj + 2 * RH; t += TX) sl[t] = gld(R.sth + jc - RH + t);
  auto load_row = [&](T(&v)[7], int j) {
    const unsigned o = (unsigned)(gi + min(j, jmax) * s);
#pragma unroll
    for (int f = 0; f < 7; f++) v[f] = gld_at(U[f], o);
  };
  auto load_halo = [&](T(&v)[7], int j) {
    if (hc >= 0) {
      const unsigned o = (unsigned)(hc + min(j, jmax) * s);
#pragma unroll
      for (int f = 0; f < 7; f++) v[f] = gld_at(U[f], o);
    }
  };
  auto put_row = [&](int buf, const T(&qq)[7][NQ], const T(&h)[7]) {
#pragma unroll
    for (int f = 0; f < 7; f++) {
      lds[buf][f][li + RH] = qq[f][RH];
      if (hc >= 0) lds[buf][f][hslot] = h[f];
    }
  };
  T q[7][NQ], h[7] = {};
  // queue rows jc - RH .. jc + RH; row jc (centre and halo) into buffer 0
#pragma unroll
  for (int d = 0; d < NQ; d++) {
    const unsigned o = (unsigned)(gi + (jc - RH + d) * s);
#pragma unroll
    for (int f = 0; f < 7; f++) q[f][d] = gld_at(U[f], o);
  }
  load_halo(h, jc);
  put_row(0, q, h);
  T na[7], nb[7];
  load_row(na, jc + RH + 1);
  // row jc + k from buffer k & 1; `use` (row jc + k + RH + 1, loaded an
  // iteration earlier) enters the queue afterwards, `pre` is loaded for the
  // next iteration
  auto body = [&](int k, T(&use)[7], T(&pre)[7]) {
    const int j = jc + k;
    load_halo(h, j + 1);
    cell.prefetch(gi, j, live);
    load_row(pre, j + RH + 2);
    __syncthreads();
    cell(q, lds[k & 1], li + RH, gi, j, live, &sl[k + RH]);
#pragma unroll
    for (int f = 0; f < 7; f++) {
#pragma unroll
      for (int d = 0; d < NQ - 1; d++) q[f][d] = q[f][d + 1];
      q[f][NQ - 1] = use[f];
    }
    // buffer (k + 1) & 1 was last read in the previous iteration, which every
    // lane finished before this iteration's barrier
    put_row((k + 1) & 1, q, h);
  };
  for (int k = 0; k < nj; k += 2) {
    body(k, na, nb);
    if (k + 1 < nj) body(k + 1, nb, na);
  }
  // the next item's first row goes to buffer 0 (and its sin(theta) window to
  // sl), which lanes may still read
  __syncthreads();
}

// Per-cell work of the march kernels: prefetch() requests the cell's own
// operands that are not stencil values (the LSRK register), operator() runs
// the cell.
template <typename T, bool FIRST, bool LAST, bool DIAG>
struct SphStageCell {
  static constexpr int RH = 2;
  const SphRun<T>& R;
  SphGeo<T, 2>& G;
  int src;
  T As, Bs, cs;
  T dp[7];
  __device__ __forceinline__ void prefetch(int i, int j, bool live) {
    if (live) sph_load_dprev<T, FIRST>(R, (unsigned)(i + j * R.dims0), dp);
  }
  template <int LW>
  __device__ __forceinline__ void operator()(const T (&q)[7][2 * RH + 1], const T (&row)[7][LW],
                                             int lx, int i, int j, bool live, const T* sl) {
    if (!live) return;
    G.polar(sl);
    sph_stage_point<T, FIRST, LAST, DIAG>(R, i, j, src, As, Bs, cs,
                                          QueueTaps<T, RH, LW>{q, row, lx}, G, dp);
  }
};

// The stage as a theta march (bands of TX radial cells, up to NJ rows).
template <typename T, int TX, int NJ, bool FIRST, bool LAST, bool DIAG>
__global__ void __launch_bounds__(TX)
    sph_stage_march_kernel(SphRun<T>* __restrict__ runs, const SphItem* __restrict__ items,
                           int nitems, int src, T As, T Bs, T cs) {
  __shared__ T lds[2][7][TX + 4];
  __shared__ T sl[NJ + 4];
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const SphItem it = items[w];
    const SphRun<T>& R = runs[it.run];
    if (!R.active) continue;  // uniform: the whole workgroup works on one item
    SphGeo<T, 2> G;
    const int ii = it.i0 + (int)threadIdx.x;
    if (ii < R.n0) G.radial(R, ii + R.g0);
    SphStageCell<T, FIRST, LAST, DIAG> cell{R, G, src, As, Bs, cs, {}};
    sph_march<T, TX, NJ, 2>(R, it, src, lds, sl, cell);
  }
}

// Kreiss-Oliger on all 7 fields (kernel_KO_step1/2_sph, .cu:444-555):
// KO_2d6 (sqrt(g)-weighted, KO_geometry) or KO_2d0, then projections and the
// pulsar boundary at t + dt.  S supplies the operands, G the metric tables.
template <typename T, typename Taps>
__device__ __forceinline__ void sph_ko_point(const SphRun<T>& R, int i, int j, int src,
                                             const Taps& S, const SphGeo<T, 3>& G) {
  const unsigned c = (unsigned)(i + j * R.dims0);
  const T r = G.r[3], sth = G.s[3];
  const T r3c = G.r3[3];
  T gm = r3c * fabs(sth);
  if (gm < TinyOf<T>::v) gm = TinyOf<T>::v;
  T wx[7], wy[7];
#pragma unroll
  for (int q = 0; q < 7; q++) {
    wx[q] = R.KO_geometry ? G.r3[q] * fabs(sth) : T(1);
    wy[q] = R.KO_geometry ? r3c * fabs(G.s[q]) : T(1);
  }
  const T iden = R.KO_geometry ? T(1) / gm : T(1);
  // the delta^6 coefficients (1, -6, 15, -20, 15, -6, 1) times the tap
  // weights, once per cell for the seven fields: one FMA per tap
  constexpr T c6[7] = {T(1), T(-6), T(15), T(-20), T(15), T(-6), T(1)};
  T cx[7], cy[7];
#pragma unroll
  for (int q = 0; q < 7; q++) {
    cx[q] = c6[q] * wx[q];
    cy[q] = c6[q] * wy[q];
  }
  T u[7];
#pragma unroll
  for (int f = 0; f < 7; f++) {
    T kx = cx[0] * S.x(f, -3), ky = cy[0] * S.y(f, -3);
#pragma unroll
    for (int q = 1; q < 7; q++) {
      kx = fma(cx[q], S.x(f, q - 3), kx);
      ky = fma(cy[q], q == 3 ? S.x(f, 0) : S.y(f, q - 3), ky);
    }
    u[f] = fma((kx + ky) * iden, R.KOfac, S.x(f, 0));
  }
  const T g1 = r * r, g2 = r * r, rs = r * sth, g3 = rs * rs;
  sph_project(R, g1, g2, g3, u);
  const T th = R.y0 + R.dy * T(j - R.g1);
  sph_pulsar(R, r, th, sth, g1, g2, g3, (T)R.t + R.dt, u);
  T* const* out = R.U[src ^ 1];
#pragma unroll
  for (int q = 0; q < 7; q++) gst_at(out[q], c, u[q]);
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
    const unsigned c = (unsigned)(i + j * R.dims0);
    const T* F[7];
#pragma unroll
    for (int f = 0; f < 7; f++) F[f] = R.U[src][f] + c;
    SphGeo<T, 3> G;
    G.radial(R, i);
    G.polar(R, j);
    sph_ko_point<T>(R, i, j, src, StrideTaps<T, true>{F, R.dims0}, G);
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
      const int i = ii + R.g0, j = jj + R.g1;
      const int slot = (lj + RH) * LXW + li + RH;
      const T* F[7];
#pragma unroll
      for (int f = 0; f < 7; f++) F[f] = &lds[f][slot];
      SphGeo<T, 3> G;
      G.radial(R, i);
      G.polar(R, j);
      sph_ko_point<T>(R, i, j, src, StrideTaps<T, false>{F, LXW}, G);
    }
    __syncthreads();
  }
}

template <typename T>
struct SphKoCell {
  static constexpr int RH = 3;
  const SphRun<T>& R;
  SphGeo<T, 3>& G;
  int src;
  __device__ __forceinline__ void prefetch(int, int, bool) {}
  template <int LW>
  __device__ __forceinline__ void operator()(const T (&q)[7][2 * RH + 1], const T (&row)[7][LW],
                                             int lx, int i, int j, bool live, const T* sl) {
    if (!live) return;
    G.polar(sl);
    sph_ko_point<T>(R, i, j, src, QueueTaps<T, RH, LW>{q, row, lx}, G);
  }
};

// Kreiss-Oliger as a theta march (as the stage).
template <typename T, int TX, int NJ>
__global__ void __launch_bounds__(TX) __attribute__((amdgpu_waves_per_eu(TX >= 512 ? 4 : 1, 8)))
    sph_ko_march_kernel(SphRun<T>* __restrict__ runs, const SphItem* __restrict__ items,
                        int nitems, int src) {
  __shared__ T lds[2][7][TX + 6];
  __shared__ T sl[NJ + 6];
  for (int w = blockIdx.x; w < nitems; w += gridDim.x) {
    const SphItem it = items[w];
    const SphRun<T>& R = runs[it.run];
    if (!R.active) continue;  // uniform: the whole workgroup works on one item
    SphGeo<T, 3> G;
    const int ii = it.i0 + (int)threadIdx.x;
    if (ii < R.n0) G.radial(R, ii + R.g0);
    SphKoCell<T> cell{R, G, src};
    sph_march<T, TX, NJ, 3>(R, it, src, lds, sl, cell);
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
    // One axis row c: E_r, B_r, P from the neighbouring row c + dir s, the
    // other components zero; guard rows c - dir l s (l = 1..3) mirrored from
    // c + dir l s with parity + for r components, - for theta and phi.  The
    // source rows are all read before any row is written (they are disjoint
    // from the written ones), so the copies cost one memory round trip.
    auto axis = [&](int c, int dir) {
      T v[3][6];
#pragma unroll
      for (int l = 0; l < 3; l++)
#pragma unroll
        for (int f = 0; f < 6; f++) v[l][f] = gld(U[f] + c + dir * (l + 1) * s);
      const T p1 = gld(U[6] + c + dir * s);
      gst(U[0] + c, v[0][0]);
      gst(U[1] + c, T(0));
      gst(U[2] + c, T(0));
      gst(U[3] + c, v[0][3]);
      gst(U[4] + c, T(0));
      gst(U[5] + c, T(0));
      gst(U[6] + c, p1);
#pragma unroll
      for (int l = 0; l < 3; l++) {
        const int d = c - dir * (l + 1) * s;
        gst(U[0] + d, v[l][0]);
        gst(U[1] + d, -v[l][1]);
        gst(U[2] + d, -v[l][2]);
        gst(U[3] + d, v[l][3]);
        gst(U[4] + d, -v[l][4]);
        gst(U[5] + d, -v[l][5]);
      }
    };
    int j = R.g1;
    T y = R.y0 + R.dy * T(j - R.g1);
    if (fabs(y) < R.dy / T(2)) axis(i + j * s, 1);
    j = R.dims1 - R.g1;
    y = R.y0 + R.dy * T(j - R.g1);
    if (fabs(y - T(M_PI)) < R.dy / T(2)) axis(i + j * s, -1);
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
  // guard cells only: the g1 full rows below and above the interior, then
  // the g0 cells left and right of every interior row
  const int rows = 2 * R.g1 * R.dims0;
  const int total = rows + 2 * R.g0 * R.n1;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    int i, j;
    if (t < rows) {
      const int r = t / R.dims0;
      i = t % R.dims0;
      j = r < R.g1 ? r : R.dims1 - 2 * R.g1 + r;
    } else {
      const int u = t - rows, c = u % (2 * R.g0);
      j = R.g1 + u / (2 * R.g0);
      i = c < R.g0 ? c : R.dims0 - 2 * R.g0 + c;
    }
    const unsigned o = (unsigned)(i + j * R.dims0);
#pragma unroll
    for (int q = 0; q < 7; q++) gst_at(R.U[1][q], o, gld_at(R.U[0][q], o));
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

// Residual monitor of every run of a batch over its interior cells (no
// reference counterpart; the reference stores per-cell divB/divE each stage,
// field_solver_EZ_spherical.cu:325-329).  Workgroup b of run `run` =
// blockIdx.x / bpr writes NRED doubles to part[blockIdx.x * NRED]:
//   sum (E_i E^i + B_i B^i) sqrt(gamma) dx dtheta,
//   max |div B|, div B = d_i(sqrt(gamma) B^i) / sqrt(gamma) (div4_sph, .cu:99-124;
//     cells on the axis, where sqrt(gamma) = 0, are skipped),
//   max |E_i B^i| / max(B_i B^i, TINY), max (E_i E^i - B_i B^i) / max(B_i B^i, TINY),
//   number of cells with a non-finite E, B or P.
// Evaluated in FP64 from the Scalar arrays.
template <typename T>
__global__ void __launch_bounds__(256)
    sph_residual_kernel(const SphRun<T>* __restrict__ runs, int bpr, double* __restrict__ part) {
  __shared__ double sh[NRED][256];
  const int run = blockIdx.x / bpr, b = blockIdx.x % bpr;
  const SphRun<T>& R = runs[run];
  double acc[NRED] = {0, 0, 0, 0, 0};
  const int total = R.n0 * R.n1;
  const double tiny = (double)TinyOf<T>::v;
  const T* const* U = R.U[0];
  const int sy = R.dims0;
  for (int t = b * (int)blockDim.x + (int)threadIdx.x; t < total; t += bpr * (int)blockDim.x) {
    const int i = t % R.n0 + R.g0, j = t / R.n0 + R.g1;
    const int c = i + j * sy;
    const double r = (double)R.r[i], s = (double)R.sth[j];
    const double g1 = r * r, g3 = (r * s) * (r * s);
    double u[7];
    for (int q = 0; q < 7; q++) u[q] = (double)gld(U[q] + c);
    const double E2 = g1 * (u[0] * u[0] + u[1] * u[1]) + g3 * u[2] * u[2];
    const double B2 = g1 * (u[3] * u[3] + u[4] * u[4]) + g3 * u[5] * u[5];
    const double EB = g1 * (u[0] * u[3] + u[1] * u[4]) + g3 * u[2] * u[5];
    const double B2c = B2 > tiny ? B2 : tiny;
    const double sg = r * r * r * s;
    bool finite = isfinite(E2) && isfinite(B2) && isfinite(u[6]);
    acc[0] += (E2 + B2) * fabs(sg) * (double)R.dx * (double)R.dy;
    if (fabs(sg) > tiny) {
      double tx = 0, ty = 0;
      const int o[4] = {-2, -1, 1, 2};
      const double w[4] = {1.0, -8.0, 8.0, -1.0};
      for (int q = 0; q < 4; q++) {
        const double ri = (double)R.r[i + o[q]];
        tx += w[q] * (double)gld(U[3] + c + o[q]) * ri * ri * ri * s;
        ty += w[q] * (double)gld(U[4] + c + o[q] * sy) * r * r * r * (double)R.sth[j + o[q]];
      }
      const double div = (tx * (double)R.inv_dx + ty * (double)R.inv_dy) / 12.0 / sg;
      acc[1] = fmax(acc[1], fabs(div));
    }
    acc[2] = fmax(acc[2], fabs(EB) / B2c);
    acc[3] = fmax(acc[3], (E2 - B2) / B2c);
    acc[4] += finite ? 0.0 : 1.0;
  }
  for (int q = 0; q < NRED; q++) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int m = blockDim.x / 2; m > 0; m >>= 1) {
    if ((int)threadIdx.x < m) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + m];
      sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][threadIdx.x + m]);
      sh[2][threadIdx.x] = fmax(sh[2][threadIdx.x], sh[2][threadIdx.x + m]);
      sh[3][threadIdx.x] = fmax(sh[3][threadIdx.x], sh[3][threadIdx.x + m]);
      sh[4][threadIdx.x] += sh[4][threadIdx.x + m];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < NRED; q++) part[blockIdx.x * NRED + q] = sh[q][0];
}

}  // namespace coffee
