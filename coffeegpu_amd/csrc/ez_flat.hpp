// Flat Cartesian EZ kernels for gfx950 (one LSRK stage per launch, fused).
//
// Semantics: the reference CPU build (src/algorithms/field_solver_EZ.cpp),
// which is normative for the flat path.  One launch of ez_stage does, per
// interior cell, everything the reference does in one pass of the stage loop
// (field_solver_EZ.cpp:1024-1055):
//   rk_step phase 1+2 (stencils, current, cleaning, dU, U += Bs dU),
//   clean_epar_check_eGTb, boundary_pulsar / boundary_alfven,
//   damping_boundary (last stage only),
// and the guard exchange becomes a periodic index wrap in the stencil loads.
// U ping-pongs between two buffers (read `in`, write `out`) so no cell is
// updated while a neighbour may still read it; dU is updated in place
// (pointwise).  ez_ko does Kreiss_Oliger + projection + boundary in one pass.
#pragma once

#include <hip/hip_runtime.h>

namespace coffee {

template <typename T>
struct TinyOf;
template <>
struct TinyOf<float> {
  static constexpr float v = 1e-7f;
};
template <>
struct TinyOf<double> {
  static constexpr double v = 1e-11;
};

// Geometry + parameters in Scalar precision, passed by value to every kernel.
template <typename T>
struct FlatK {
  int dims[3];
  int guard[3];
  int n[3];
  int per[3];
  int sy, sz;  // strides of y and z
  T delta[3];
  T lower[3];
  T c12[3];    // inv_delta / 12
  T dt, ch2, tau, KOfac, damp_over_dt;
  int divB_clean, use_edotb, problem, pert_type, mask_lanes;
  int pml_any;
  T xl[3], xh[3];      // absorbing-layer bounds
  T pmlscale[3];       // pmllen * delta
  T sigpml;
  T radius, omega, b0;
  T p1, p2, p3, q11, q12, q13, q22, q23, qox, qoy, qoz;
  T tp_start, tp_end, rpert1, rpert2, dw0, nT, ph1, ph2, dph, theta0, drpert;
  T tp_start1, tp_end1, rpert11, rpert21, dw1, nT1, ph11, ph21, dph1;
};

template <typename T>
struct FlatIO {
  const T* __restrict__ in[7];
  T* __restrict__ out[7];
  T* __restrict__ dU[7];
  T* __restrict__ Un[7];
  const T* __restrict__ Bbg[3];
  T* __restrict__ J[3];
  T* __restrict__ divE;
  T* __restrict__ divB;
  // energy accounting of diagnostics mode (any may be null): dU_KO,
  // dU_Epar, dU_EgtB, then their _cum arrays
  T* __restrict__ acc[6];
};

// Element `idx` of a field array.  O32: addressed by a 32-bit byte offset
// from the array's base (arrays below 4 GiB; the caller checks), so a load or
// store is one scalar base plus one vector offset shared by every field, with
// no 64-bit address arithmetic per access; the access is cast to the global
// address space.  Otherwise plain 64-bit indexing.  Used by the tiled
// Kreiss-Oliger kernel of the FP32 default tile (1.74 against 1.80 ms at
// 512^3 on one box); the stage kernels ran slower with it (3.46 against 3.37
// ms) and FP64 the same, so they index plainly.
template <typename T, bool O32>
__device__ __forceinline__ T ldf(const T* base, int idx) {
  if (O32)
    return *(const __attribute__((address_space(1))) T*)((const char*)base +
                                                          (unsigned)idx * (unsigned)sizeof(T));
  return base[idx];
}
template <typename T, bool O32>
__device__ __forceinline__ void stf(T* base, int idx, T v) {
  if (O32)
    *(__attribute__((address_space(1))) T*)((char*)base + (unsigned)idx * (unsigned)sizeof(T)) = v;
  else
    base[idx] = v;
}

__device__ __forceinline__ int wrap_idx(int c, int g, int n, int per) {
  if (per) {
    if (c < g) c += n;
    else if (c >= g + n) c -= n;
  }
  return c;
}

template <typename T>
__device__ __forceinline__ T d1(const T* __restrict__ f, int c, int m2, int m1, int p1, int p2,
                                T c12) {
  // 4th-order centred first derivative (finite_diff_simd.h:50-60,95-97).
  return fma(T(8), f[c + p1] - f[c + m1], f[c + m2] - f[c + p2]) * c12;
}

// delta^6 of seven consecutive values m3..p3 with explicit fused multiply-adds.
template <typename T>
__device__ __forceinline__ T d6v(T m3, T m2, T m1, T c0, T p1, T p2, T p3) {
  return fma(T(-20), c0, fma(T(15), m1 + p1, fma(T(-6), m2 + p2, m3 + p3)));
}

template <typename T>
__device__ __forceinline__ T d6(const T* __restrict__ f, int c, const int* o) {
  // delta^6 (finite_diff_simd.h:76-93); o[0..5] = offsets of -3,-2,-1,+1,+2,+3.
  return d6v(f[c + o[0]], f[c + o[1]], f[c + o[2]], f[c], f[c + o[3]], f[c + o[4]], f[c + o[5]]);
}

template <typename T>
__device__ __forceinline__ T node_pos(const FlatK<T>& k, int ax, int n) {
  return k.lower[ax] + k.delta[ax] * T(n - k.guard[ax]);
}

template <typename T>
__device__ __forceinline__ T cube(T x) {
  return x * x * x;
}

template <typename T>
__device__ __forceinline__ T shape_fn(T r, T r0, T del) {
  // pulsar_impl.hpp:52-55
  return T(0.5) * (T(1) - tanh((r - r0) / del));
}

template <typename T>
__device__ T dipole2_d(T x, T y, T z, T p1, T p2, T p3, T phase, int n) {
  // pulsar_impl.hpp:57-80
  const T tiny = TinyOf<T>::v;
  T c = cos(phase), s = sin(phase);
  T p1t = p1 * c - p2 * s, p2t = p1 * s + p2 * c, p3t = p3;
  T r2 = x * x + y * y + z * z;
  T den = sqrt(cube(r2)) + tiny;
  T pr = p1t * x + p2t * y + p3t * z;
  T comp = n == 0 ? x : (n == 1 ? y : z);
  T pt = n == 0 ? p1t : (n == 1 ? p2t : p3t);
  return (T(3) * comp * pr / (r2 + tiny) - pt) / den;
}

template <typename T>
__device__ T quadrupole_d(const FlatK<T>& k, T x, T y, T z, T phase, int n) {
  // pulsar_impl.hpp:82-119
  const T tiny = TinyOf<T>::v;
  T c = cos(phase), s = sin(phase);
  T q11t = k.q11 * c * c - T(2) * k.q12 * c * s + k.q22 * s * s;
  T q12t = k.q12 * cos(T(2) * phase) + (k.q11 - k.q22) * c * s;
  T q22t = k.q22 * c * c + k.q11 * s * s + T(2) * k.q12 * c * s;
  T q13t = k.q13 * c - k.q23 * s;
  T q23t = k.q23 * c + k.q13 * s;
  T q33t = -k.q11 - k.q22;
  T ox = k.qox * c - k.qoy * s, oy = k.qox * s + k.qoy * c;
  T x1 = x - ox, y1 = y - oy, z1 = z - k.qoz;
  T r2 = x1 * x1 + y1 * y1 + z1 * z1;
  T r7 = cube(r2) * sqrt(r2);
  T xqx = q11t * x1 * x1 + T(2) * q12t * x1 * y1 + q22t * y1 * y1 +
          T(2) * (q13t * x1 + q23t * y1) * z1 + q33t * z1 * z1;
  T a0 = n == 0 ? q11t : (n == 1 ? q12t : q13t);
  T a1 = n == 0 ? q12t : (n == 1 ? q22t : q23t);
  T a2 = n == 0 ? q13t : (n == 1 ? q23t : q33t);
  T comp = n == 0 ? x1 : (n == 1 ? y1 : z1);
  return (T(-2) * (a0 * x1 + a1 * y1 + a2 * z1) * r2 + T(5) * comp * xqx) / (r7 + tiny);
}

template <typename T>
__device__ T wpert_d(T t, T th, T tps, T tpe, T dw0, T nT, T rp1, T rp2) {
  // field_solver_EZ.cpp:644-661
  T th1 = acos(sqrt(T(1) - T(1) / rp1));
  T th2 = acos(sqrt(T(1) - T(1) / rp2));
  if (th1 > th2) {
    T tmp = th1;
    th1 = th2;
    th2 = tmp;
  }
  T mu = (th1 + th2) / T(2);
  T s = (mu - th1) / T(3);
  if (t >= tps && t <= tpe && th >= th1 && th <= th2) {
    T q = (th - mu) / s;
    return dw0 * exp(T(-0.5) * q * q) * sin((t - tps) * T(2) * T(M_PI) * nT / (tpe - tps));
  }
  return T(0);
}

template <typename T>
__device__ T wpert3d_d(T t, T th, T ph, T tps, T tpe, T dw0, T nT, T rp1, T rp2, T ph1, T ph2,
                       T dph) {
  // field_solver_EZ.cpp:663-684
  T w = wpert_d(t, th, tps, tpe, dw0, nT, rp1, rp2);
  if (w == T(0)) return w;
  return w * shape_fn(ph, ph2 * T(M_PI), dph * T(M_PI)) *
         (T(1) - shape_fn(ph, ph1 * T(M_PI), dph * T(M_PI)));
}

template <typename T>
__device__ void twistv_d(const FlatK<T>& k, T t, T x, T y, T z, T ri, T v[3]) {
  // field_solver_EZ.cpp:686-709
  const T tiny = TinyOf<T>::v;
  T c0 = cos(k.theta0), s0 = sin(k.theta0);
  T x1 = -x * c0 + z * s0, y1 = y, z1 = x * s0 + z * c0;
  T r = sqrt(x * x + y * y + z * z);
  T R1 = sqrt(x1 * x1 + y1 * y1);
  T costh = x1 / (R1 + tiny), sinth = y1 / (R1 + tiny);
  T dw = T(0);
  if (r > ri && z1 >= T(0) && R1 <= k.drpert && t >= k.tp_start && t <= k.tp_end) {
    T cc = cos(R1 / k.drpert * T(M_PI) / T(2));
    T ss = sin((t - k.tp_start) * T(M_PI) / (k.tp_end - k.tp_start));
    dw = k.dw0 * cc * cc * sin((t - k.tp_start) * T(2) * T(M_PI) * k.nT / (k.tp_end - k.tp_start)) *
         ss * ss;
  }
  v[0] = dw * R1 * sinth * c0;
  v[1] = dw * R1 * costh;
  v[2] = -dw * R1 * sinth * s0;
}

// boundary_pulsar (problem 1, field_solver_EZ.cpp:711-839) and
// boundary_alfven (problem 2, .cpp:841-1005) on one cell's values.
template <typename T>
__device__ void star_boundary(const FlatK<T>& k, T x, T y, T z, T t, T& ex, T& ey, T& ez, T& bx,
                              T& by, T& bz) {
  const T tiny = TinyOf<T>::v;
  T rl = T(1.5) * k.radius, ri = T(0.5) * k.radius;
  T r2 = x * x + y * y + z * z;
  if (r2 < tiny) r2 = tiny;
  T r = sqrt(r2);
  if (!(r < rl)) return;
  T scaleEpar = T(0.5) * k.delta[0], scaleEperp = T(0.25) * k.delta[0];
  T scaleB = scaleEpar;
  T d1 = T(4) * k.delta[0];
  T phase = k.omega * t;
  T vx = 0, vy = 0, vz = 0, bxn, byn, bzn;
  if (k.problem == 1) {
    vx = -k.omega * y;
    vy = k.omega * x;
    bxn = k.b0 * (dipole2_d(x, y, z, k.p1, k.p2, k.p3, phase, 0) + quadrupole_d(k, x, y, z, phase, 0));
    byn = k.b0 * (dipole2_d(x, y, z, k.p1, k.p2, k.p3, phase, 1) + quadrupole_d(k, x, y, z, phase, 1));
    bzn = k.b0 * (dipole2_d(x, y, z, k.p1, k.p2, k.p3, phase, 2) + quadrupole_d(k, x, y, z, phase, 2));
  } else {
    T th = acos(z / r);
    if (k.pert_type == 0 || k.pert_type == 1) {
      T w0, w1;
      if (k.pert_type == 0) {
        w0 = wpert_d(t, th, k.tp_start, k.tp_end, k.dw0, k.nT, k.rpert1, k.rpert2);
        w1 = wpert_d(t, th, k.tp_start1, k.tp_end1, k.dw1, k.nT1, k.rpert11, k.rpert21);
      } else {
        T ph = atan2(y, x);
        w0 = wpert3d_d(t, th, ph, k.tp_start, k.tp_end, k.dw0, k.nT, k.rpert1, k.rpert2, k.ph1,
                       k.ph2, k.dph);
        w1 = wpert3d_d(t, th, ph, k.tp_start1, k.tp_end1, k.dw1, k.nT1, k.rpert11, k.rpert21,
                       k.ph11, k.ph21, k.dph1);
      }
      T w = k.omega + w0 + w1;
      vx = -w * y;
      vy = w * x;
    } else if (k.pert_type == 2) {
      T v[3];
      twistv_d(k, t, x, y, z, ri, v);
      vx = v[0];
      vy = v[1];
      vz = v[2];
    }
    bxn = k.b0 * dipole2_d(x, y, z, k.p1, k.p2, k.p3, phase, 0);
    byn = k.b0 * dipole2_d(x, y, z, k.p1, k.p2, k.p3, phase, 1);
    bzn = k.b0 * dipole2_d(x, y, z, k.p1, k.p2, k.p3, phase, 2);
  }
  T s = shape_fn(r, k.radius - d1, scaleB);
  T bn_r = bxn * x + byn * y + bzn * z;
  T B_r = bx * x + by * y + bz * z;
  // radial part and tangential part use the same blend (scaleBperp == scaleBpar)
  T nbx = bn_r * x / r2 * s + B_r * x / r2 * (T(1) - s);
  T nby = bn_r * y / r2 * s + B_r * y / r2 * (T(1) - s);
  T nbz = bn_r * z / r2 * s + B_r * z / r2 * (T(1) - s);
  nbx += (bxn - bn_r * x / r2) * s + (bx - B_r * x / r2) * (T(1) - s);
  nby += (byn - bn_r * y / r2) * s + (by - B_r * y / r2) * (T(1) - s);
  nbz += (bzn - bn_r * z / r2) * s + (bz - B_r * z / r2) * (T(1) - s);
  bx = nbx;
  by = nby;
  bz = nbz;
  T exn = -vy * bz + vz * by;
  T eyn = vx * bz - vz * bx;
  T ezn = -vx * by + vy * bx;
  T en_r = exn * x + eyn * y + ezn * z;
  T E_r = ex * x + ey * y + ez * z;
  s = shape_fn(r, k.radius, scaleEperp);
  T nex = en_r * x / r2 * s + E_r * x / r2 * (T(1) - s);
  T ney = en_r * y / r2 * s + E_r * y / r2 * (T(1) - s);
  T nez = en_r * z / r2 * s + E_r * z / r2 * (T(1) - s);
  s = shape_fn(r, k.radius, scaleEpar);
  nex += (exn - en_r * x / r2) * s + (ex - E_r * x / r2) * (T(1) - s);
  ney += (eyn - en_r * y / r2) * s + (ey - E_r * y / r2) * (T(1) - s);
  nez += (ezn - en_r * z / r2) * s + (ez - E_r * z / r2) * (T(1) - s);
  ex = nex;
  ey = ney;
  ez = nez;
  if (r < ri) {
    bx = bxn;
    by = byn;
    bz = bzn;
    ex = exn;
    ey = eyn;
    ez = ezn;
  }
}

// clean_epar_check_eGTb (field_solver_EZ.cpp:552-634) on one cell.
template <typename T>
__device__ __forceinline__ void project_cell(T& ex, T& ey, T& ez, T bx, T by, T bz) {
  T B2 = bx * bx + by * by + bz * bz;
  B2 = B2 >= TinyOf<T>::v ? B2 : TinyOf<T>::v;
  T EB = bx * ex + by * ey + bz * ez;
  T f = EB / B2;
  ex -= f * bx;
  ey -= f * by;
  ez -= f * bz;
  T E2 = ex * ex + ey * ey + ez * ez;
  if (E2 > B2) {
    T s = sqrt(B2 / E2);
    ex *= s;
    ey *= s;
    ez *= s;
  }
}

// project_cell with the energy accounting of clean_epar_check_eGTb
// (field_solver_EZ.cpp:588-623): dU_Epar += |E'|^2 - |E|^2 after the E.B
// removal, dU_EgtB += |E''|^2 - |E|^2 after the E>B cap (both from the
// pre-projection |E|^2), and the same into the _cum arrays.
template <typename T>
__device__ __forceinline__ void project_cell_account(const FlatIO<T>& io, int c, T& ex, T& ey,
                                                     T& ez, T bx, T by, T bz) {
  const T u0 = ex * ex + ey * ey + ez * ez;
  T B2 = bx * bx + by * by + bz * bz;
  B2 = B2 >= TinyOf<T>::v ? B2 : TinyOf<T>::v;
  const T EB = bx * ex + by * ey + bz * ez;
  const T f = EB / B2;
  ex -= f * bx;
  ey -= f * by;
  ez -= f * bz;
  const T E2 = ex * ex + ey * ey + ez * ez;
  const T u1 = E2;
  if (E2 > B2) {
    const T s = sqrt(B2 / E2);
    ex *= s;
    ey *= s;
    ez *= s;
  }
  const T u2 = ex * ex + ey * ey + ez * ez;
  if (io.acc[1]) io.acc[1][c] += u1 - u0;
  if (io.acc[4]) io.acc[4][c] += u1 - u0;
  if (io.acc[2]) io.acc[2][c] += u2 - u0;
  if (io.acc[5]) io.acc[5][c] += u2 - u0;
}

template <typename T>
__device__ __forceinline__ T pml_sigma(T x, T xl, T xh, T scale, T sig0) {
  // damping_boundary.h:14-22
  if (x > xh) return sig0 * cube((x - xh) / scale);
  if (x < xl) return sig0 * cube((xl - x) / scale);
  return T(0);
}

template <typename T>
__device__ __forceinline__ bool in_layer(const FlatK<T>& k, T x, T y, T z) {
  return x > k.xh[0] || x < k.xl[0] || y > k.xh[1] || y < k.xl[1] || z > k.xh[2] || z < k.xl[2];
}

// Per-cell body of one LSRK stage, shared by the simple and the tiled kernel.
// d[f][a] = d(field f)/d(axis a) for f in Ex,Ey,Ez,Bx,By,Bz,P (4th order,
// already scaled by 1/delta); u[] = the cell's E,B,P.  FIRST: As = 0 (no dU
// read) and the absorbing-layer cells of U are saved to Un; LAST: dU is not
// written back (the next step starts with As = 0) and the absorbing layer is
// applied.
template <typename T, bool FIRST, bool LAST, bool DIAG, bool STAR>
__device__ __forceinline__ void stage_cell(const FlatK<T>& k, const FlatIO<T>& io, const int i,
                                           const int j, const int kk, const int c,
                                           const T (&d)[7][3], T (&u)[7], const T (&dprev)[7],
                                           const T As, const T Bs, const T tstage) {
  // field_solver_EZ.cpp:139-147
  const T rotBx = d[5][1] - d[4][2];
  const T rotBy = d[3][2] - d[5][0];
  const T rotBz = d[4][0] - d[3][1];
  const T rotEx = d[2][1] - d[1][2];
  const T rotEy = d[0][2] - d[2][0];
  const T rotEz = d[1][0] - d[0][1];
  const T divE = d[0][0] + d[1][1] + d[2][2];
  const T divB = d[3][0] + d[4][1] + d[5][2];
  const T ex = u[0], ey = u[1], ez = u[2], bx = u[3], by = u[4], bz = u[5], pc = u[6];

  T B2 = bx * bx + by * by + bz * bz;
  B2 = B2 >= TinyOf<T>::v ? B2 : TinyOf<T>::v;
  T Jx, Jy, Jz;
  T Jp = (bx * rotBx + by * rotBy + bz * rotBz) - (ex * rotEx + ey * rotEy + ez * rotEz);
  if (k.use_edotb) {
    // field_solver_EZ.cpp:170-186
    T E2 = ex * ex + ey * ey + ez * ez;
    T chi2 = B2 - E2;
    T EdotB = ex * bx + ey * by + ez * bz;
    T E02 = T(0.5) * (sqrt(chi2 * chi2 + T(4) * EdotB * EdotB) - chi2) + B2;
    Jp += EdotB * k.damp_over_dt;
    T a = divE / E02, b = Jp / B2;
    Jx = a * (ey * bz - ez * by) + b * bx;
    Jy = a * (ez * bx - ex * bz) + b * by;
    Jz = a * (ex * by - ey * bx) + b * bz;
  } else {
    // field_solver_EZ.cpp:187-196
    T inv = T(1) / B2;
    Jx = (divE * (ey * bz - ez * by) + Jp * bx) * inv;
    Jy = (divE * (ez * bx - ex * bz) + Jp * by) * inv;
    Jz = (divE * (ex * by - ey * bx) + Jp * bz) * inv;
  }

  T Px = 0, Py = 0, Pz = 0;
  if (k.divB_clean) {
    // tanh mask on grad P (field_solver_EZ.cpp:222-230), positions as the
    // reference CPU build evaluates them (lane offset on x, y and z)
    const int lane = (i - k.guard[0]) % k.mask_lanes;
    const T fl = T(lane);
    const T x = fl * k.delta[0] + node_pos(k, 0, i - lane);
    const T y = fl * k.delta[1] + node_pos(k, 1, j);
    const T z = fl * k.delta[2] + node_pos(k, 2, kk);
    const T r = sqrt(x * x + y * y + z * z);
    const T s = T(0.5) * (T(1) - tanh((r - T(3)) * T(2)));
    Px = d[6][0] * s;
    Py = d[6][1] * s;
    Pz = d[6][2] * s;
  }

  if (DIAG) {
    io.J[0][c] = Jx;
    io.J[1][c] = Jy;
    io.J[2][c] = Jz;
    io.divE[c] = divE;
    io.divB[c] = divB;
  }

  const T dt = k.dt;
  T du[7];
  du[0] = (rotBx - Jx) * dt;
  du[1] = (rotBy - Jy) * dt;
  du[2] = (rotBz - Jz) * dt;
  du[3] = -(rotEx + Px) * dt;
  du[4] = -(rotEy + Py) * dt;
  du[5] = -(rotEz + Pz) * dt;
  du[6] = -pc / k.tau - divB * k.ch2 * dt;
  if (!FIRST) {
#pragma unroll
    for (int q = 0; q < 7; q++) du[q] = fma(dprev[q], As, du[q]);
  }
  if (!LAST) {
#pragma unroll
    for (int q = 0; q < 7; q++) io.dU[q][c] = du[q];
  }

  if (FIRST && k.pml_any) {
    const T x = node_pos(k, 0, i), y = node_pos(k, 1, j), z = node_pos(k, 2, kk);
    if (in_layer(k, x, y, z)) {
#pragma unroll
      for (int q = 0; q < 7; q++) io.Un[q][c] = u[q];
    }
  }
#pragma unroll
  for (int q = 0; q < 7; q++) u[q] = fma(du[q], Bs, u[q]);

  if (DIAG) project_cell_account(io, c, u[0], u[1], u[2], u[3], u[4], u[5]);
  else project_cell(u[0], u[1], u[2], u[3], u[4], u[5]);
  if (STAR) {
    const T x = node_pos(k, 0, i), y = node_pos(k, 1, j), z = node_pos(k, 2, kk);
    star_boundary(k, x, y, z, tstage, u[0], u[1], u[2], u[3], u[4], u[5]);
  }
  if (LAST && k.pml_any) {
    // damping_boundary (damping_boundary.cpp:5-83): E toward En, B toward Bbg
    const T x = node_pos(k, 0, i), y = node_pos(k, 1, j), z = node_pos(k, 2, kk);
    if (in_layer(k, x, y, z)) {
      const T sig = pml_sigma(x, k.xl[0], k.xh[0], k.pmlscale[0], k.sigpml) +
                    pml_sigma(y, k.xl[1], k.xh[1], k.pmlscale[1], k.sigpml) +
                    pml_sigma(z, k.xl[2], k.xh[2], k.pmlscale[2], k.sigpml);
      if (sig > TinyOf<T>::v) {
        const T e = exp(-sig), f = (T(1) - e) / sig;
#pragma unroll
        for (int q = 0; q < 3; q++) {
          const T un = io.Un[q][c];
          u[q] = e * un + f * (u[q] - un);
        }
#pragma unroll
        for (int q = 0; q < 3; q++) {
          const T un = io.Un[3 + q][c], bg = io.Bbg[q][c];
          u[3 + q] = bg + e * (un - bg) + f * (u[3 + q] - un);
        }
      }
    }
  }
#pragma unroll
  for (int q = 0; q < 6; q++) io.out[q][c] = u[q];
  if (io.out[6]) io.out[6][c] = u[6];
}

// Per-cell end of the Kreiss-Oliger pass (field_solver_EZ.cpp:390-457 with
// the projection and boundary of .cpp:1057-1067): u[0..5] already hold
// f + (KOeps/64) sum_axes delta^6 f, u[6] = P (untouched on the CPU path).
// DIAG: e_before = |E|^2 + |B|^2 before the KO update; dU_KO = the change
// (field_solver_EZ.cpp:447-453), then the projection's accounting.
template <typename T, bool STAR, bool DIAG = false, bool O32 = false>
__device__ __forceinline__ void ko_cell_finish(const FlatK<T>& k, const FlatIO<T>& io, int i,
                                               int j, int kk, int c, T (&u)[7], T t,
                                               T e_before = T(0)) {
  if (DIAG) {
    const T e_after = (u[0] * u[0] + u[1] * u[1] + u[2] * u[2]) +
                      (u[3] * u[3] + u[4] * u[4] + u[5] * u[5]);
    if (io.acc[0]) io.acc[0][c] = e_after - e_before;
    if (io.acc[3]) io.acc[3][c] += e_after - e_before;
    project_cell_account(io, c, u[0], u[1], u[2], u[3], u[4], u[5]);
  } else {
    project_cell(u[0], u[1], u[2], u[3], u[4], u[5]);
  }
  if (STAR) {
    const T x = node_pos(k, 0, i), y = node_pos(k, 1, j), z = node_pos(k, 2, kk);
    star_boundary(k, x, y, z, t, u[0], u[1], u[2], u[3], u[4], u[5]);
  }
#pragma unroll
  for (int q = 0; q < 6; q++) stf<T, O32>(io.out[q], c, u[q]);
  if (io.out[6]) stf<T, O32>(io.out[6], c, u[6]);
}

// Simple stage kernel: one thread per cell, stencil neighbours read straight
// from global memory (periodic axes wrap into the interior).  Used for grids
// the tiled kernel does not take and as the second implementation in tests.
template <typename T, bool FIRST, bool LAST, bool DIAG, bool STAR>
__global__ void __launch_bounds__(256)
    ez_stage_kernel(const FlatK<T> k, const FlatIO<T> io, const T As, const T Bs, const T tstage) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + k.guard[0];
  const int j = blockIdx.y * blockDim.y + threadIdx.y + k.guard[1];
  const int kk = blockIdx.z + k.guard[2];
  if (i >= k.guard[0] + k.n[0] || j >= k.guard[1] + k.n[1]) return;
  const int c = i + (j + kk * k.dims[1]) * k.dims[0];
  int ox[4], oy[4], oz[4];
  const int dd[4] = {-2, -1, 1, 2};
#pragma unroll
  for (int q = 0; q < 4; q++) {
    ox[q] = wrap_idx(i + dd[q], k.guard[0], k.n[0], k.per[0]) - i;
    oy[q] = (wrap_idx(j + dd[q], k.guard[1], k.n[1], k.per[1]) - j) * k.sy;
    oz[q] = (wrap_idx(kk + dd[q], k.guard[2], k.n[2], k.per[2]) - kk) * k.sz;
  }
  T d[7][3];
  T u[7];
#pragma unroll
  for (int f = 0; f < 7; f++) {
    const T* p = io.in[f];
    d[f][0] = d1(p, c, ox[0], ox[1], ox[2], ox[3], k.c12[0]);
    d[f][1] = d1(p, c, oy[0], oy[1], oy[2], oy[3], k.c12[1]);
    d[f][2] = d1(p, c, oz[0], oz[1], oz[2], oz[3], k.c12[2]);
    u[f] = p[c];
  }
  T dp[7];
#pragma unroll
  for (int q = 0; q < 7; q++) dp[q] = FIRST ? T(0) : io.dU[q][c];
  stage_cell<T, FIRST, LAST, DIAG, STAR>(k, io, i, j, kk, c, d, u, dp, As, Bs, tstage);
}

// Simple Kreiss_Oliger kernel (one thread per cell); DIAG adds the energy
// accounting (diagnostics mode always runs this kernel).
template <typename T, bool STAR, bool DIAG = false>
__global__ void __launch_bounds__(256) ez_ko_kernel(const FlatK<T> k, const FlatIO<T> io, const T t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x + k.guard[0];
  const int j = blockIdx.y * blockDim.y + threadIdx.y + k.guard[1];
  const int kk = blockIdx.z + k.guard[2];
  if (i >= k.guard[0] + k.n[0] || j >= k.guard[1] + k.n[1]) return;
  const int c = i + (j + kk * k.dims[1]) * k.dims[0];
  int ox[6], oy[6], oz[6];
  const int dd[6] = {-3, -2, -1, 1, 2, 3};
#pragma unroll
  for (int q = 0; q < 6; q++) {
    ox[q] = wrap_idx(i + dd[q], k.guard[0], k.n[0], k.per[0]) - i;
    oy[q] = (wrap_idx(j + dd[q], k.guard[1], k.n[1], k.per[1]) - j) * k.sy;
    oz[q] = (wrap_idx(kk + dd[q], k.guard[2], k.n[2], k.per[2]) - kk) * k.sz;
  }
  T u[7];
#pragma unroll
  for (int q = 0; q < 6; q++) {
    const T* f = io.in[q];
    T ko = d6(f, c, ox) + d6(f, c, oy) + d6(f, c, oz);
    u[q] = fma(ko, k.KOfac, f[c]);
  }
  u[6] = io.in[6] ? io.in[6][c] : T(0);
  T e_before = T(0);
  if (DIAG) {
    T v[6];
#pragma unroll
    for (int q = 0; q < 6; q++) v[q] = io.in[q][c];
    e_before = (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]) + (v[3] * v[3] + v[4] * v[4] + v[5] * v[5]);
  }
  ko_cell_finish<T, STAR, DIAG>(k, io, i, j, kk, c, u, t, e_before);
}

// Guard slabs of one axis: periodic -> copy from the interior one period
// away (send_array_{x,y,z} on a single periodic rank, sim_env.cpp:365-438);
// COPY mode -> copy the same cell from src to dst (frozen guards of the
// ping-pong buffer).
template <typename T>
struct Ptr7 {
  T* p[7];
};

template <typename T, bool WRAP>
__global__ void __launch_bounds__(256)
    guard_slab_kernel(const FlatK<T> k, int ax, const Ptr7<T> dst, const Ptr7<T> src, int nfields) {
  const int g = k.guard[ax];
  // slab extent: 2g along ax, full dims along the other axes
  int ext[3] = {k.dims[0], k.dims[1], k.dims[2]};
  ext[ax] = 2 * g;
  const long total = (long)ext[0] * ext[1] * ext[2];
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    int c0 = (int)(t % ext[0]);
    long r = t / ext[0];
    int c1 = (int)(r % ext[1]);
    int c2 = (int)(r / ext[1]);
    int cc[3] = {c0, c1, c2};
    int m = cc[ax];
    if (m >= g) m = k.dims[ax] - 2 * g + m;  // upper slab
    cc[ax] = m;
    int s[3] = {cc[0], cc[1], cc[2]};
    if (WRAP) s[ax] = m < g ? m + k.n[ax] : m - k.n[ax];
    const int di = cc[0] + (cc[1] + cc[2] * k.dims[1]) * k.dims[0];
    const int si = s[0] + (s[1] + s[2] * k.dims[1]) * k.dims[0];
    for (int q = 0; q < nfields; q++) dst.p[q][di] = src.p[q][si];
  }
}

// Periodic guard fill of every guard cell in one pass: guard cell (i, j, k)
// takes the cell whose coordinate along each periodic axis is wrapped into the
// interior (other axes keep theirs).  This is the result of the reference's
// x, then y, then z slab sequence (sim_env.cpp:365-438: a corner is copied
// along x, the copy along y, that along z, ending at the fully wrapped
// interior cell; along a frozen axis the copies keep the frozen guard row).
// Guard cells are enumerated without overlap: the z slabs (whole x-y
// planes), then the y slabs of interior planes, then the x slabs of interior
// rows.
template <typename T>
__global__ void __launch_bounds__(256)
    guard_wrap_kernel(const FlatK<T> k, const Ptr7<T> a, int nfields) {
  const long d0 = k.dims[0], d1 = k.dims[1];
  const int g0 = k.guard[0], g1 = k.guard[1], g2 = k.guard[2];
  const long nzs = 2L * g2 * d0 * d1;
  const long nys = 2L * g1 * d0 * k.n[2];
  const long nxs = 2L * g0 * k.n[1] * k.n[2];
  const long total = nzs + nys + nxs;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    int i, j, kk;
    if (t < nzs) {
      i = (int)(t % d0);
      const long r = t / d0;
      j = (int)(r % d1);
      const int m = (int)(r / d1);
      kk = m < g2 ? m : k.dims[2] - 2 * g2 + m;
    } else if (t < nzs + nys) {
      const long u = t - nzs;
      i = (int)(u % d0);
      const long r = u / d0;
      const int m = (int)(r % (2 * g1));
      kk = g2 + (int)(r / (2 * g1));
      j = m < g1 ? m : k.dims[1] - 2 * g1 + m;
    } else {
      const long u = t - nzs - nys;
      const int m = (int)(u % (2 * g0));
      const long r = u / (2 * g0);
      j = g1 + (int)(r % k.n[1]);
      kk = g2 + (int)(r / k.n[1]);
      i = m < g0 ? m : k.dims[0] - 2 * g0 + m;
    }
    int si = i, sj = j, sk = kk;
    if (k.per[0]) si = i < g0 ? i + k.n[0] : (i >= g0 + k.n[0] ? i - k.n[0] : i);
    if (k.per[1]) sj = j < g1 ? j + k.n[1] : (j >= g1 + k.n[1] ? j - k.n[1] : j);
    if (k.per[2]) sk = kk < g2 ? kk + k.n[2] : (kk >= g2 + k.n[2] ? kk - k.n[2] : kk);
    if (si == i && sj == j && sk == kk) continue;  // frozen guard cell
    const long di = i + (j + (long)kk * d1) * d0;
    const long src = si + (sj + (long)sk * d1) * d0;
    for (int q = 0; q < nfields; q++) a.p[q][di] = a.p[q][src];
  }
}

// Block partial sums for total_energy / residuals.  Each block writes
// NRED doubles to part[blockIdx.x*NRED ...]; the host finishes the reduction.
constexpr int NRED = 5;

template <typename T>
__global__ void __launch_bounds__(256)
    residual_kernel(const FlatK<T> k, const FlatIO<T> io, double* __restrict__ part) {
  __shared__ double sh[NRED][256];
  double acc[NRED] = {0, 0, 0, 0, 0};
  const long total = (long)k.n[0] * k.n[1] * k.n[2];
  const int ox[4] = {-2, -1, 1, 2};
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    int i = (int)(t % k.n[0]) + k.guard[0];
    long r = t / k.n[0];
    int j = (int)(r % k.n[1]) + k.guard[1];
    int kk = (int)(r / k.n[1]) + k.guard[2];
    int c = i + (j + kk * k.dims[1]) * k.dims[0];
    T ex = io.in[0][c], ey = io.in[1][c], ez = io.in[2][c];
    T bx = io.in[3][c], by = io.in[4][c], bz = io.in[5][c];
    double E2 = (double)ex * ex + (double)ey * ey + (double)ez * ez;
    double B2 = (double)bx * bx + (double)by * by + (double)bz * bz;
    double EB = (double)ex * bx + (double)ey * by + (double)ez * bz;
    double B2c = B2 > (double)TinyOf<T>::v ? B2 : (double)TinyOf<T>::v;
    int o[3][4];
    for (int q = 0; q < 4; q++) {
      o[0][q] = wrap_idx(i + ox[q], k.guard[0], k.n[0], k.per[0]) - i;
      o[1][q] = (wrap_idx(j + ox[q], k.guard[1], k.n[1], k.per[1]) - j) * k.sy;
      o[2][q] = (wrap_idx(kk + ox[q], k.guard[2], k.n[2], k.per[2]) - kk) * k.sz;
    }
    double divB = (double)d1(io.in[3], c, o[0][0], o[0][1], o[0][2], o[0][3], k.c12[0]) +
                  (double)d1(io.in[4], c, o[1][0], o[1][1], o[1][2], o[1][3], k.c12[1]) +
                  (double)d1(io.in[5], c, o[2][0], o[2][1], o[2][2], o[2][3], k.c12[2]);
    bool finite = isfinite(E2) && isfinite(B2) && isfinite((double)io.in[6][c]);
    acc[0] += E2 + B2;
    acc[1] = fmax(acc[1], fabs(divB));
    acc[2] = fmax(acc[2], fabs(EB) / B2c);
    acc[3] = fmax(acc[3], (E2 - B2) / B2c);
    acc[4] += finite ? 0.0 : 1.0;
  }
  for (int q = 0; q < NRED; q++) sh[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      sh[0][threadIdx.x] += sh[0][threadIdx.x + s];
      sh[1][threadIdx.x] = fmax(sh[1][threadIdx.x], sh[1][threadIdx.x + s]);
      sh[2][threadIdx.x] = fmax(sh[2][threadIdx.x], sh[2][threadIdx.x + s]);
      sh[3][threadIdx.x] = fmax(sh[3][threadIdx.x], sh[3][threadIdx.x + s]);
      sh[4][threadIdx.x] += sh[4][threadIdx.x + s];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int q = 0; q < NRED; q++) part[blockIdx.x * NRED + q] = sh[q][0];
}

// field_solver_EZ::total_energy (field_solver_EZ.cu:755-795): strict
// inequalities on the layer bounds; r < radius excluded for problem 1.
template <typename T>
__global__ void __launch_bounds__(256)
    energy_kernel(const FlatK<T> k, const T* __restrict__ f0, const T* __restrict__ f1,
                  const T* __restrict__ f2, double* __restrict__ part) {
  __shared__ double sh[256];
  double acc = 0;
  const long total = (long)k.n[0] * k.n[1] * k.n[2];
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    int i = (int)(t % k.n[0]) + k.guard[0];
    long r = t / k.n[0];
    int j = (int)(r % k.n[1]) + k.guard[1];
    int kk = (int)(r / k.n[1]) + k.guard[2];
    int c = i + (j + kk * k.dims[1]) * k.dims[0];
    T x = node_pos(k, 0, i), y = node_pos(k, 1, j), z = node_pos(k, 2, kk);
    T rr = sqrt(x * x + y * y + z * z);
    bool ok = !(k.problem == 1 && rr < k.radius) && x < k.xh[0] && x > k.xl[0] && y < k.xh[1] &&
              y > k.xl[1] && z < k.xh[2] && z > k.xl[2];
    if (ok) acc += (double)f0[c] * f0[c] + (double)f1[c] * f1[c] + (double)f2[c] * f2[c];
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

}  // namespace coffee
