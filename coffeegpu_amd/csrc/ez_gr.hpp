// Kerr-Schild GR EZ kernels for gfx950 (Cartesian Kerr-Schild, M = 1).
//
// Semantics: src/algorithms/field_solver_gr_EZ.cu, boundary.cu:7-79 and
// metric_cks_impl.hpp (the reference's only implementation of this path).
// Evolved: D (the sim_data "E" arrays), B (upper components) and P.
//
// The 3-metric of Kerr-Schild is gamma_ij = delta_ij + 2H l_i l_j with
// H = r^3 / (r^4 + a^2 z^2), l = ((r x + a y)/(r^2 + a^2), (r y - a x)/(r^2 + a^2), z/r)
// and |l| = 1, so everything the solver needs follows from (r, s = 2H, l):
//   sqrt(gamma) = sqrt(1 + s),  alpha = 1/sqrt(1 + s),
//   gamma_ij v^j = v_i + s l_i (l.v),   gamma^ij w_j = w^i - s/(1+s) l^i (l.w),
//   beta^i = s/(1+s) l^i.
// The reference evaluates some 20 metric functions at each use, each
// recomputing r (metric_cks_impl.hpp:17-181); here one FP64 evaluation of
// (r, s, l) per grid point feeds all of them, in both Scalar precisions.
//
// One step is, per stage: gr_point (E_i, H_i and sqrt(gamma) D, sqrt(gamma) B
// at every point of the guarded grid, kernel_compute_E_gr/_H_gr .cu:90-186),
// then gr_stage (rk_step1 + rk_step2 + clean_epar + check_eGTb, and the
// absorbing layers in the last stage, .cu:188-535, 610-701, boundary.cu:17-79),
// with U ping-ponging between two buffers; then gr_ko (KO on D, B, P +
// projections, .cu:537-608).
#pragma once

#include <hip/hip_runtime.h>

#include "ez_flat.hpp"

namespace coffee {

template <typename T>
struct GrK {
  int dims[3], guard[3], n[3];
  int sy, sz;
  double lower[3], delta[3];  // Grid::pos in FP64
  double a;                   // spin
  T c12[3];                   // inv_delta / 12
  T dt, tau, itau, KOfac;  // itau = 1 / tau
  int calc_current, divB_clean, clean_ep, check_egb;
  int pml_any;
  double xl[3], xh[3];  // outer absorbing-layer bounds (boundary.cu:41-52)
  double pmlscale;      // pmllen * delta[0] on every axis (boundary.cu:55-60)
  double sigpml;
  double r1, r2, dd;    // 0.7 rH, 0.01 rH, 0.2 rH (.cu:625-628)
};

template <typename T>
struct GrIO {
  const T* __restrict__ in[7];
  T* __restrict__ out[7];
  T* __restrict__ dU[7];
  T* __restrict__ Un[7];
  T* __restrict__ J[3];
  T* __restrict__ divE;
  T* __restrict__ divB;
  T* __restrict__ W[12];  // E_i (3), H_i (3), sqrt(g) D^i (3), sqrt(g) B^i (3)
};

// s = 2H feeds gamma_ij, alpha and sqrt(gamma); bu = 2r^3 / (2r^3 + r^4 + a^2 z^2)
// feeds gamma^ij and beta^i.  They are s/(1+s) of each other except where the
// reference's TINY clamp of r^4 + a^2 z^2 acts (grid points on the ring
// singularity), where the reference keeps them apart and so does this code.
struct Cks {
  double r, s, bu, lx, ly, lz;
};

// r of metric_cks_impl.hpp:17-24 and the Kerr-Schild H, l at one point.
// 1/b, 1/sqrt(g) and sqrt(x) in FP64 from the hardware estimates (v_rcp_f64,
// v_rsq_f64) refined by two Newton steps each: within a few ulp of the
// correctly rounded FP64 values, far below the FP32 rounding that the single
// precision solver applies to every metric quantity, in about half the
// instructions of the IEEE division and square root sequences.
__device__ __forceinline__ double rcp_nr(double b) {
  double y = __builtin_amdgcn_rcp(b);
  y = fma(y, fma(-b, y, 1.0), y);
  y = fma(y, fma(-b, y, 1.0), y);
  return y;
}
__device__ __forceinline__ double rsq_nr(double g) {  // y (3 - g y^2) / 2, twice
  const double h = 0.5 * g;
  double y = __builtin_amdgcn_rsq(g);
  y = fma(y, fma(-h * y, y, 0.5), y);
  y = fma(y, fma(-h * y, y, 0.5), y);
  return y;
}
__device__ __forceinline__ double sqrt_nr(double x) { return x > 0.0 ? x * rsq_nr(x) : 0.0; }

// FAST: the Newton-refined reciprocals and roots above (the FP32 solver's
// stages and KO); otherwise IEEE FP64 (the FP64 solver, energy, residuals).
template <bool FAST = false>
__device__ __forceinline__ Cks cks_point(double a, double x, double y, double z) {
  const double tiny = 1e-11;
  const double R2 = x * x + y * y + z * z;
  const double w = R2 - a * a;
  const double az = a * z;
  Cks m;
  if constexpr (FAST)
    m.r = sqrt_nr(0.5 * (w + sqrt_nr(fmax(w * w + 4.0 * (az * az), tiny))));
  else
    m.r = sqrt((w + sqrt(fmax(w * w + 4.0 * (az * az), tiny))) / 2.0);
  const double r = m.r, r2 = r * r;
  const double Q = fmax(r2 * r2 + az * az, tiny);
  const double A = fmax(a * a + r2, tiny);
  const double r3 = 2.0 * r * r2;
  const double Qb = fmax(r3 + r2 * r2 + az * az, tiny);
  if constexpr (FAST) {
    m.s = r3 * rcp_nr(Q);
    m.bu = r3 * rcp_nr(Qb);
  } else {
    m.s = r3 / Q;
    m.bu = r3 / Qb;
  }
  const double iA = FAST ? rcp_nr(A) : 1.0 / A;
  m.lx = (r * x + a * y) * iA;
  m.ly = (r * y - a * x) * iA;
  m.lz = r > 0.0 ? (FAST ? z * rcp_nr(r) : z / r) : 0.0;
  return m;
}

template <typename T>
struct CksT {
  T s, l[3], alpha, sg, bfac;  // bfac = bu (= s / (1 + s) away from the clamp)
  double r;
};

template <typename T, bool FAST>
__device__ __forceinline__ CksT<T> cks_scalar(const Cks& m) {
  const double g = 1.0 + m.s;
  const double ia = FAST ? rsq_nr(g) : 0.0;
  const double sg = FAST ? g * ia : sqrt(g);
  CksT<T> t;
  t.r = m.r;
  t.s = (T)m.s;
  t.l[0] = (T)m.lx;
  t.l[1] = (T)m.ly;
  t.l[2] = (T)m.lz;
  t.sg = (T)sg;
  t.alpha = FAST ? (T)ia : (T)(1.0 / sg);
  t.bfac = (T)m.bu;
  return t;
}

// The metric the solver kernels use at a point: Newton-refined FP64 in the
// FP32 solver, IEEE FP64 in the FP64 solver.  Every schedule of a kernel
// uses the same choice, so the schedules stay bitwise identical.
template <typename T, bool FAST = sizeof(T) == 4>
__device__ __forceinline__ CksT<T> gr_metric(double a, double x, double y, double z) {
  return cks_scalar<T, FAST>(cks_point<FAST>(a, x, y, z));
}

// The library is built with -ffp-contract=off.  The solver arithmetic from
// here on (not the positions and the metric above, not the energy and
// residual reductions at the end of this file, which keep the plain
// operation order of their host counterparts) lets each source expression's
// multiply-add be fused into one FMA: contraction within an expression only,
// decided per expression as written, so every kernel that inlines a routine
// gets the same instructions and the schedules stay bitwise identical.
#pragma clang fp contract(on)

template <typename T>
__device__ __forceinline__ void gr_lower(const CksT<T>& m, const T v[3], T out[3]) {
  const T lv = m.l[0] * v[0] + m.l[1] * v[1] + m.l[2] * v[2];
  const T sl = m.s * lv;
#pragma unroll
  for (int i = 0; i < 3; i++) out[i] = v[i] + sl * m.l[i];
}

template <typename T>
__device__ __forceinline__ T dot3(const T u[3], const T v[3]) {
  return u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
}

template <typename T>
__device__ __forceinline__ double gr_pos(const GrK<T>& k, int ax, int i) {
#pragma clang fp contract(off)
  return k.lower[ax] + k.delta[ax] * (double)(i - k.guard[ax]);
}

// E_i = alpha gamma_ij D^j + sqrt(g) (beta x B)_i, H_i = alpha gamma_ij B^j
// - sqrt(g) (beta x D)_i (.cu:111-184), with beta = bu l:
// E = alpha (D + s l (l.D)) + sqrt(g) bu (l x B), H likewise with B, -D.
template <typename T>
__device__ __forceinline__ void gr_point_values(const CksT<T>& m, const T D[3], const T B[3],
                                                T E[3], T H[3]) {
  const T* l = m.l;
  const T slD = m.s * dot3(l, D), slB = m.s * dot3(l, B);
  const T w = m.sg * m.bfac;
  const T cB[3] = {l[1] * B[2] - l[2] * B[1], l[2] * B[0] - l[0] * B[2], l[0] * B[1] - l[1] * B[0]};
  const T cD[3] = {l[1] * D[2] - l[2] * D[1], l[2] * D[0] - l[0] * D[2], l[0] * D[1] - l[1] * D[0]};
#pragma unroll
  for (int i = 0; i < 3; i++) {
    E[i] = m.alpha * (D[i] + slD * l[i]) + w * cB[i];
    H[i] = m.alpha * (B[i] + slB * l[i]) - w * cD[i];
  }
}

// clean_epar then check_eGTb with the metric (.cu:440-535) on D (u[0..2]).
template <typename T>
__device__ __forceinline__ void gr_project(const GrK<T>& k, const CksT<T>& m, T u[7]) {
  const T tiny = TinyOf<T>::v;
  T D[3] = {u[0], u[1], u[2]};
  const T B[3] = {u[3], u[4], u[5]};
  T Bd[3];
  gr_lower(m, B, Bd);
  const T B2raw = dot3(B, Bd);
  if (k.clean_ep) {
    const T B2 = B2raw < tiny ? tiny : B2raw;
    const T f = dot3(D, Bd) / B2;  // one division for the three components
#pragma unroll
    for (int i = 0; i < 3; i++) D[i] = D[i] - f * B[i];
  }
  if (k.check_egb) {
    T Dd[3];
    gr_lower(m, D, Dd);
    T D2 = dot3(D, Dd);
    if (D2 < tiny) D2 = tiny;
    const T temp = D2 > B2raw ? sqrt(B2raw / D2) : T(1);
#pragma unroll
    for (int i = 0; i < 3; i++) D[i] = temp * D[i];
  }
  u[0] = D[0];
  u[1] = D[1];
  u[2] = D[2];
}

template <typename T>
__device__ __forceinline__ bool gr_in_outer(const GrK<T>& k, double x, double y, double z) {
  return k.pml_any && (x > k.xh[0] || x < k.xl[0] || y > k.xh[1] || y < k.xl[1] ||
                       z > k.xh[2] || z < k.xl[2]);
}

// pmlsigma (boundary.cu:7-15), in the field precision T as the reference's
// Scalar builds evaluate it (positions and bounds are held in double)
template <typename T>
__device__ __forceinline__ T gr_pmlsigma(double x, double xl, double xh, T scale, T sig0) {
  if (x > xh) {
    const T v = (T)(x - xh) / scale;
    return sig0 * v * v * v;
  }
  if (x < xl) {
    const T v = (T)(xl - x) / scale;
    return sig0 * v * v * v;
  }
  return T(0);
}

// Damping factors of the layers: e^-sig and (1 - e^-sig) / sig.  In FP32 the
// second comes from expm1, since 1 - e^-sig cancels for the small sig at the
// inner edge of a layer; FP64 keeps the reference's expression.
template <typename T>
__device__ __forceinline__ void gr_damp(T sig, T& ex, T& w) {
  if constexpr (sizeof(T) == 4) {
    ex = expf(-sig);
    w = -expm1f(-sig) / sig;
  } else {
    ex = exp(-sig);
    w = (T(1) - ex) / sig;
  }
}

// Cells whose step-start copy Un is saved in the first stage and read by
// the absorbing layer in the last: the outer layer and the excision region.
template <typename T>
__device__ __forceinline__ bool gr_needs_un(const GrK<T>& k, double x, double y, double z,
                                            double r) {
  return gr_in_outer(k, x, y, z) || r < k.r1;
}

// kernel_boundary_absorbing_thread (boundary.cu:17-79) then
// kernel_absorbing_inner (.cu:617-701) on D, B; P untouched.  un: the cell's
// Un (read only where gr_needs_un holds).
template <typename T>
__device__ __forceinline__ void gr_absorb(const GrK<T>& k, const T un_[7], double x, double y,
                                          double z, double r, T u[7]) {
  const T tiny = TinyOf<T>::v;
  if (gr_in_outer(k, x, y, z)) {
    const T scale = (T)k.pmlscale, sig0 = (T)k.sigpml;
    const T sig = gr_pmlsigma<T>(x, k.xl[0], k.xh[0], scale, sig0) +
                  gr_pmlsigma<T>(y, k.xl[1], k.xh[1], scale, sig0) +
                  gr_pmlsigma<T>(z, k.xl[2], k.xh[2], scale, sig0);
    if (sig > tiny) {
      T ex, w;
      gr_damp(sig, ex, w);
#pragma unroll
      for (int f = 0; f < 6; f++) {
        const T un = un_[f];
        u[f] = ex * un + w * (u[f] - un);
      }
    }
  }
  if (r < k.r1) {
    const T v = (T)(k.r1 - r) / (T)k.dd;
    const T sig = (T)k.sigpml * v * v * v * k.dt;
    if (sig > tiny) {
      T ex, w;
      gr_damp(sig, ex, w);
#pragma unroll
      for (int f = 0; f < 6; f++) {
        const T un = un_[f];
        u[f] = ex * un + w * (u[f] - un);
      }
    }
  }
  if (r < k.r2) {
#pragma unroll
    for (int f = 0; f < 6; f++) u[f] = T(0);
  }
}

// Every point of the guarded grid: W = (E_i, H_i, sqrt(g) D^i, sqrt(g) B^i).
template <typename T>
__global__ void __launch_bounds__(256) gr_point_kernel(const GrK<T> k, const GrIO<T> io) {
  const int total = k.dims[0] * k.dims[1] * k.dims[2];
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < total; c += gridDim.x * blockDim.x) {
    const int i = c % k.dims[0];
    const int r = c / k.dims[0];
    const int j = r % k.dims[1];
    const int kk = r / k.dims[1];
    const CksT<T> m = gr_metric<T>(k.a, gr_pos(k, 0, i), gr_pos(k, 1, j), gr_pos(k, 2, kk));
    const T D[3] = {io.in[0][c], io.in[1][c], io.in[2][c]};
    const T B[3] = {io.in[3][c], io.in[4][c], io.in[5][c]};
    T E[3], H[3];
    gr_point_values(m, D, B, E, H);
#pragma unroll
    for (int q = 0; q < 3; q++) {
      io.W[q][c] = E[q];
      io.W[3 + q][c] = H[q];
      io.W[6 + q][c] = m.sg * D[q];
      io.W[9 + q][c] = m.sg * B[q];
    }
  }
}

template <typename T>
__device__ __forceinline__ T gr_d1(const T* __restrict__ f, int c, int s, T c12) {
  return fma(T(8), f[c + s] - f[c - s], f[c - 2 * s] - f[c + 2 * s]) * c12;
}

// Derivatives a stage needs at one cell, each a 4th-order centred difference
// times inv_delta: ddx = d/dx of (E_y, E_z, H_y, H_z, sqrt(g) D^x, sqrt(g) B^x, P),
// ddy = d/dy of (E_x, E_z, H_x, H_z, sqrt(g) D^y, sqrt(g) B^y, P),
// ddz = d/dz of (E_x, E_y, H_x, H_y, sqrt(g) D^z, sqrt(g) B^z, P).
// Which W array feeds each slot, per axis:
constexpr int kGrDx[6] = {1, 2, 4, 5, 6, 9};
constexpr int kGrDy[6] = {0, 2, 3, 5, 7, 10};
constexpr int kGrDz[6] = {0, 1, 3, 4, 8, 11};

// One LSRK stage at an interior cell from its derivatives (rk_step1
// .cu:188-373, rk_step2 .cu:375-438, projections .cu:440-535, absorbing in
// the last stage).  u: D, B, P at the cell (in), dprev: dU before the stage,
// Ec: E_i at the cell, un: the cell's Un (last stage, where gr_needs_un).
template <typename T, bool FIRST, bool LAST, bool DIAG, bool O32 = false>
__device__ __forceinline__ void gr_stage_core(const GrK<T>& k, const GrIO<T>& io, int c, double x,
                                              double y, double z, const CksT<T>& m,
                                              const T ddx[7], const T ddy[7], const T ddz[7],
                                              const T Ec[3], T u[7], const T dprev[7],
                                              const T un[7], T As, T Bs) {
  const T tiny = TinyOf<T>::v;
  // rot = curl / sqrt(g) = alpha * curl
  const T rotE[3] = {m.alpha * (ddy[1] - ddz[1]), m.alpha * (ddz[0] - ddx[1]),
                     m.alpha * (ddx[0] - ddy[0])};
  const T rotH[3] = {m.alpha * (ddy[3] - ddz[3]), m.alpha * (ddz[2] - ddx[3]),
                     m.alpha * (ddx[2] - ddy[2])};
  // div4 (.cu:44-88): d_i(sqrt(g) f^i) / sqrt(g)
  const T divD = m.alpha * (ddx[4] + ddy[4] + ddz[4]);
  const T divB = m.alpha * (ddx[5] + ddy[5] + ddz[5]);
  const T D[3] = {u[0], u[1], u[2]};
  const T B[3] = {u[3], u[4], u[5]};
  T Bd[3], Dd[3];
  gr_lower(m, B, Bd);
  gr_lower(m, D, Dd);
  T J[3] = {T(0), T(0), T(0)};
  if (k.calc_current) {
    T B2 = dot3(B, Bd);
    if (B2 < tiny) B2 = tiny;
    const T Jp = dot3(Bd, rotH) - dot3(Dd, rotE);
    const T rho = divD * m.alpha;
    const T iB2 = T(1) / B2;  // one division for the three components
    J[0] = (rho * (Ec[1] * Bd[2] - Ec[2] * Bd[1]) + Jp * B[0]) * iB2;
    J[1] = (rho * (Ec[2] * Bd[0] - Ec[0] * Bd[2]) + Jp * B[1]) * iB2;
    J[2] = (rho * (Ec[0] * Bd[1] - Ec[1] * Bd[0]) + Jp * B[2]) * iB2;
  }
  const T gP[3] = {ddx[6], ddy[6], ddz[6]};
  const T lgP = dot3(m.l, gP);
  const T beta[3] = {m.bfac * m.l[0], m.bfac * m.l[1], m.bfac * m.l[2]};
  const T shape = T(0.5) * (T(1) - tanh((T)(m.r - 5.0)));
  T du[7];
#pragma unroll
  for (int f = 0; f < 7; f++) du[f] = FIRST ? T(0) : dprev[f];
  const T dt = k.dt;
#pragma unroll
  for (int q = 0; q < 3; q++) {
    if (k.divB_clean) {
      const T Pu = gP[q] - m.bfac * m.l[q] * lgP;
      du[3 + q] = As * du[3 + q] + dt * (-rotE[q] + (-m.alpha * Pu + beta[q] * divB) * shape);
    } else {
      du[3 + q] = As * du[3 + q] - dt * rotE[q];
    }
    du[q] = As * du[q] + dt * (rotH[q] - J[q]);
  }
  du[6] = As * du[6] + dt * (-m.alpha * divB + dot3(beta, gP)) - m.alpha * u[6] * k.itau;
  if (FIRST && gr_needs_un(k, x, y, z, m.r)) {
#pragma unroll
    for (int f = 0; f < 7; f++) stf<T, O32>(io.Un[f], c, u[f]);
  }
#pragma unroll
  for (int f = 0; f < 7; f++) u[f] = u[f] + Bs * du[f];
  gr_project(k, m, u);
  if (LAST) gr_absorb(k, un, x, y, z, m.r, u);
#pragma unroll
  for (int f = 0; f < 7; f++) stf<T, O32>(io.out[f], c, u[f]);
  if (!LAST) {
#pragma unroll
    for (int f = 0; f < 7; f++) stf<T, O32>(io.dU[f], c, du[f]);
  }
  if (DIAG) {
    io.J[0][c] = J[0];
    io.J[1][c] = J[1];
    io.J[2][c] = J[2];
    io.divE[c] = divD;
    io.divB[c] = divB;
  }
}

// The stage at interior cell (i, j, kk) with E_i, H_i, sqrt(g) D, sqrt(g) B
// read from the point sweep's arrays W.
template <typename T, bool FIRST, bool LAST, bool DIAG>
__device__ __forceinline__ void gr_stage_cell(const GrK<T>& k, const GrIO<T>& io, int i, int j,
                                              int kk, int c, T As, T Bs) {
  const int sy = k.sy, sz = k.sz;
  const double x = gr_pos(k, 0, i), y = gr_pos(k, 1, j), z = gr_pos(k, 2, kk);
  const CksT<T> m = gr_metric<T>(k.a, x, y, z);
  const T* const* W = (const T* const*)io.W;
  const T* P = io.in[6];
  T ddx[7], ddy[7], ddz[7];
#pragma unroll
  for (int q = 0; q < 6; q++) {
    ddx[q] = gr_d1(W[kGrDx[q]], c, 1, k.c12[0]);
    ddy[q] = gr_d1(W[kGrDy[q]], c, sy, k.c12[1]);
    ddz[q] = gr_d1(W[kGrDz[q]], c, sz, k.c12[2]);
  }
  ddx[6] = gr_d1(P, c, 1, k.c12[0]);
  ddy[6] = gr_d1(P, c, sy, k.c12[1]);
  ddz[6] = gr_d1(P, c, sz, k.c12[2]);
  const T Ec[3] = {W[0][c], W[1][c], W[2][c]};
  T u[7], dp[7], un[7];
  const bool need_un = LAST && gr_needs_un(k, x, y, z, m.r);
#pragma unroll
  for (int f = 0; f < 7; f++) {
    u[f] = io.in[f][c];
    dp[f] = FIRST ? T(0) : io.dU[f][c];
    un[f] = need_un ? io.Un[f][c] : T(0);
  }
  gr_stage_core<T, FIRST, LAST, DIAG>(k, io, c, x, y, z, m, ddx, ddy, ddz, Ec, u, dp, un, As,
                                      Bs);
}

template <typename T, bool FIRST, bool LAST, bool DIAG>
__global__ void __launch_bounds__(256) gr_stage_kernel(const GrK<T> k, const GrIO<T> io, const T As,
                                                       const T Bs) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int kk = blockIdx.z;
  if (i >= k.n[0] || j >= k.n[1]) return;
  const int ai = i + k.guard[0], aj = j + k.guard[1], ak = kk + k.guard[2];
  const int c = ai + aj * k.sy + ak * k.sz;
  gr_stage_cell<T, FIRST, LAST, DIAG>(k, io, ai, aj, ak, c, As, Bs);
}

// KO on D, B, P (.cu:537-608) then the projections (.cu:902-904).
template <typename T>
__global__ void __launch_bounds__(256) gr_ko_kernel(const GrK<T> k, const GrIO<T> io) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int j = blockIdx.y * blockDim.y + threadIdx.y;
  const int kk = blockIdx.z;
  if (i >= k.n[0] || j >= k.n[1]) return;
  const int ai = i + k.guard[0], aj = j + k.guard[1], ak = kk + k.guard[2];
  const int c = ai + aj * k.sy + ak * k.sz;
  const int ox[6] = {-3, -2, -1, 1, 2, 3};
  const int oy[6] = {-3 * k.sy, -2 * k.sy, -k.sy, k.sy, 2 * k.sy, 3 * k.sy};
  const int oz[6] = {-3 * k.sz, -2 * k.sz, -k.sz, k.sz, 2 * k.sz, 3 * k.sz};
  T u[7];
#pragma unroll
  for (int f = 0; f < 7; f++) {
    const T* p = io.in[f];
    const T t = d6(p, c, ox) + d6(p, c, oy) + d6(p, c, oz);
    u[f] = fma(t, k.KOfac, p[c]);
  }
  const CksT<T> m = gr_metric<T>(k.a, gr_pos(k, 0, ai), gr_pos(k, 1, aj), gr_pos(k, 2, ak));
  gr_project(k, m, u);
#pragma unroll
  for (int f = 0; f < 7; f++) io.out[f][c] = u[f];
}

#pragma clang fp contract(off)
// Field energy sum over the interior outside the absorbing layers:
// sum gamma_ij f^i f^j sqrt(gamma) (no reference counterpart: the reference
// declares field_solver_gr_EZ::total_energy but never defines it).
template <typename T>
__global__ void __launch_bounds__(256)
    gr_energy_kernel(const GrK<T> k, const T* __restrict__ f0, const T* __restrict__ f1,
                     const T* __restrict__ f2, double* __restrict__ part) {
  __shared__ double red[256];
  double acc = 0.0;
  const int total = k.n[0] * k.n[1] * k.n[2];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const int i = t % k.n[0] + k.guard[0];
    const int r = t / k.n[0];
    const int j = r % k.n[1] + k.guard[1];
    const int kk = r / k.n[1] + k.guard[2];
    const double x = gr_pos(k, 0, i), y = gr_pos(k, 1, j), z = gr_pos(k, 2, kk);
    const Cks m = cks_point(k.a, x, y, z);
    if (gr_in_outer(k, x, y, z) || m.r < k.r1) continue;
    const int c = i + j * k.sy + kk * k.sz;
    const double v[3] = {(double)f0[c], (double)f1[c], (double)f2[c]};
    const double lv = m.lx * v[0] + m.ly * v[1] + m.lz * v[2];
    acc += ((v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + m.s * lv * lv) * sqrt(1.0 + m.s);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// Residual monitor over interior cells (no reference counterpart; the
// reference stores per-cell divB/divD each stage, field_solver_gr_EZ.cu:336-338),
// NRED doubles per workgroup, in FP64 from the Scalar arrays:
//   sum (D_i D^i + B_i B^i) sqrt(gamma),
//   max |div B|, div B = d_i(sqrt(gamma) B^i) / sqrt(gamma) (div4, .cu:44-88),
//   max |D_i B^i| / max(B_i B^i, TINY), max (D_i D^i - B_i B^i) / max(B_i B^i, TINY),
//   number of cells with a non-finite D, B or P.
template <typename T>
__global__ void __launch_bounds__(256)
    gr_residual_kernel(const GrK<T> k, const FlatIO<T> io, double* __restrict__ part) {
  __shared__ double sh[NRED][256];
  double acc[NRED] = {0, 0, 0, 0, 0};
  const int total = k.n[0] * k.n[1] * k.n[2];
  const double tiny = (double)TinyOf<T>::v;
  const int o[4] = {-2, -1, 1, 2};
  const double wt[4] = {1.0, -8.0, 8.0, -1.0};
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const int i = t % k.n[0] + k.guard[0];
    const int rr = t / k.n[0];
    const int j = rr % k.n[1] + k.guard[1];
    const int kk = rr / k.n[1] + k.guard[2];
    const int c = i + j * k.sy + kk * k.sz;
    const double x = gr_pos(k, 0, i), y = gr_pos(k, 1, j), z = gr_pos(k, 2, kk);
    const Cks m = cks_point(k.a, x, y, z);
    double u[7];
    for (int q = 0; q < 7; q++) u[q] = (double)io.in[q][c];
    const double lD = m.lx * u[0] + m.ly * u[1] + m.lz * u[2];
    const double lB = m.lx * u[3] + m.ly * u[4] + m.lz * u[5];
    const double D2 = u[0] * u[0] + u[1] * u[1] + u[2] * u[2] + m.s * lD * lD;
    const double B2 = u[3] * u[3] + u[4] * u[4] + u[5] * u[5] + m.s * lB * lB;
    const double DB = u[0] * u[3] + u[1] * u[4] + u[2] * u[5] + m.s * lD * lB;
    const double B2c = B2 > tiny ? B2 : tiny;
    const double sg = sqrt(1.0 + m.s);
    double div = 0;
    const int st[3] = {1, k.sy, k.sz};
    for (int ax = 0; ax < 3; ax++) {
      double acc_ax = 0;
      for (int q = 0; q < 4; q++) {
        const double px = x + (ax == 0 ? o[q] * k.delta[0] : 0.0);
        const double py = y + (ax == 1 ? o[q] * k.delta[1] : 0.0);
        const double pz = z + (ax == 2 ? o[q] * k.delta[2] : 0.0);
        const double sgt = sqrt(1.0 + cks_point(k.a, px, py, pz).s);
        acc_ax += wt[q] * (double)io.in[3 + ax][c + o[q] * st[ax]] * sgt;
      }
      div += acc_ax / (12.0 * k.delta[ax]);
    }
    div /= sg;
    const bool finite = isfinite(D2) && isfinite(B2) && isfinite(u[6]);
    acc[0] += (D2 + B2) * sg;
    acc[1] = fmax(acc[1], fabs(div));
    acc[2] = fmax(acc[2], fabs(DB) / B2c);
    acc[3] = fmax(acc[3], (D2 - B2) / B2c);
    acc[4] += finite ? 0.0 : 1.0;
  }
  for (int q = 0; q < NRED; q++) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
      sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][threadIdx.x + w]);
      sh[2][threadIdx.x] = fmax(sh[2][threadIdx.x], sh[2][threadIdx.x + w]);
      sh[3][threadIdx.x] = fmax(sh[3][threadIdx.x], sh[3][threadIdx.x + w]);
      sh[4][threadIdx.x] += sh[4][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < NRED; q++) part[blockIdx.x * NRED + q] = sh[q][0];
}

}  // namespace coffee
