// C ABI of the gfx950 force-free field-update hot path (include/coffee_hip.h).
//
// A context holds the solver variant, the grid, the kernel parameter block in
// the context's Scalar precision and the scratch it owns: the second buffer of
// the ping-pong state (7 arrays of the full guarded extent), the reduction
// buffers and the timing events.  Caller arrays are never allocated or freed
// here.

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "coffee_hip.h"
#include "ctx.hpp"
#include "ez_flat.hpp"
#include "ez_flat_tiled.hpp"

using namespace coffee;

namespace coffee {
std::string& create_error() {
  static thread_local std::string e;
  return e;
}
}  // namespace coffee

namespace {

size_t cells(const coffee_grid& g) { return (size_t)g.dims[0] * g.dims[1] * g.dims[2]; }

template <typename T>
void fill_k(const coffee_grid& g, const coffee_params& p, FlatK<T>& k) {
  for (int a = 0; a < 3; a++) {
    k.dims[a] = g.dims[a];
    k.guard[a] = g.guard[a];
    k.n[a] = g.dims[a] - 2 * g.guard[a];
    k.per[a] = p.periodic[a] ? 1 : 0;
    k.delta[a] = (T)g.delta[a];
    k.lower[a] = (T)g.lower[a];
    // inv_delta = 1.0 / delta in Scalar (src/sim_env_impl.hpp:57-58)
    T inv = (T)(1.0 / (double)k.delta[a]);
    k.c12[a] = inv / T(12);
    T lo = (T)p.lower[a], sz = (T)p.size[a];
    k.xh[a] = lo + sz - (T)p.pml[a] * k.delta[a];
    k.xl[a] = lo + (T)p.pml[a] * k.delta[a];
    k.pmlscale[a] = (T)p.pmllen * k.delta[a];
  }
  k.sy = g.dims[0];
  k.sz = g.dims[0] * g.dims[1];
  k.dt = (T)p.dt;
  k.ch2 = (T)p.ch2;
  k.tau = (T)p.tau;
  // -KOeps * KO_const with KO_const = -1/64 (field_solver_EZ.cpp:379-388)
  k.KOfac = -(T)p.KOeps * (T)(-1.0 / 64.0);
  k.damp_over_dt = (T)p.damp_gamma / (T)p.dt;
  k.divB_clean = p.divB_clean;
  k.use_edotb = p.use_edotb_damping;
  k.problem = p.problem;
  k.pert_type = p.pert_type;
  k.mask_lanes = p.mask_lanes > 0 ? p.mask_lanes : 1;
  k.pml_any = (p.pml[0] > 0 || p.pml[1] > 0 || p.pml[2] > 0) ? 1 : 0;
  k.sigpml = (T)p.sigpml;
  k.radius = (T)p.radius;
  k.omega = (T)p.omega;
  k.b0 = (T)p.b0;
  k.p1 = (T)p.p1; k.p2 = (T)p.p2; k.p3 = (T)p.p3;
  k.q11 = (T)p.q11; k.q12 = (T)p.q12; k.q13 = (T)p.q13; k.q22 = (T)p.q22; k.q23 = (T)p.q23;
  k.qox = (T)p.q_offset_x; k.qoy = (T)p.q_offset_y; k.qoz = (T)p.q_offset_z;
  k.tp_start = (T)p.tp_start; k.tp_end = (T)p.tp_end; k.rpert1 = (T)p.rpert1;
  k.rpert2 = (T)p.rpert2; k.dw0 = (T)p.dw0; k.nT = (T)p.nT; k.ph1 = (T)p.ph1; k.ph2 = (T)p.ph2;
  k.dph = (T)p.dph; k.theta0 = (T)p.theta0; k.drpert = (T)p.drpert;
  k.tp_start1 = (T)p.tp_start1; k.tp_end1 = (T)p.tp_end1; k.rpert11 = (T)p.rpert11;
  k.rpert21 = (T)p.rpert21; k.dw1 = (T)p.dw1; k.nT1 = (T)p.nT1; k.ph11 = (T)p.ph11;
  k.ph21 = (T)p.ph21; k.dph1 = (T)p.dph1;
}

int validate(coffee_ctx* c, const coffee_grid* g, const coffee_params* p) {
  for (int a = 0; a < 3; a++) {
    if (g->dims[a] < 1 || g->guard[a] < 0 || g->dims[a] - 2 * g->guard[a] < 1)
      return fail(c, COFFEE_ERR_INVALID, "invalid grid dims/guard");
    if (g->guard[a] < 3)
      return fail(c, COFFEE_ERR_INVALID, "guard must be >= 3 on every axis (KO6 stencil radius)");
    if (p->periodic[a] && g->dims[a] - 2 * g->guard[a] < g->guard[a])
      return fail(c, COFFEE_ERR_INVALID, "periodic axis needs at least `guard` interior cells");
  }
  double total = (double)g->dims[0] * g->dims[1] * g->dims[2];
  if (total >= 2147483647.0) return fail(c, COFFEE_ERR_SIZE, "grid exceeds 2^31-1 cells per array");
  if (!(p->dt > 0)) return fail(c, COFFEE_ERR_INVALID, "dt must be positive");
  if (p->tau == 0) return fail(c, COFFEE_ERR_INVALID, "tau must be nonzero");
  if (p->shift_ghost != 0)
    return fail(c, COFFEE_ERR_INVALID, "shift_ghost != 0 is not supported (guards stay frozen)");
  return COFFEE_OK;
}

template <typename T>
FlatK<T>& kparams(coffee_ctx* c);
template <>
FlatK<float>& kparams<float>(coffee_ctx* c) { return c->kf; }
template <>
FlatK<double>& kparams<double>(coffee_ctx* c) { return c->kd; }

template <typename T>
int check_fields(coffee_ctx* c, const coffee_fields* f, bool need_du) {
  for (int q = 0; q < 3; q++)
    if (!f->E[q] || !f->B[q]) return fail(c, COFFEE_ERR_INVALID, "E/B pointers must be set");
  if (!f->P) return fail(c, COFFEE_ERR_INVALID, "P pointer must be set");
  if (need_du) {
    for (int q = 0; q < 3; q++)
      if (!f->dE[q] || !f->dB[q]) return fail(c, COFFEE_ERR_INVALID, "dE/dB pointers must be set");
    if (!f->dP) return fail(c, COFFEE_ERR_INVALID, "dP pointer must be set");
  }
  return COFFEE_OK;
}

template <typename T>
void state_ptrs(const coffee_fields* f, T* out[7]) {
  for (int q = 0; q < 3; q++) {
    out[q] = (T*)f->E[q];
    out[3 + q] = (T*)f->B[q];
  }
  out[6] = (T*)f->P;
}

template <typename T>
__global__ void fill_kernel(T* p, unsigned long long n, T v) {
  for (unsigned long long i = blockIdx.x * (unsigned long long)blockDim.x + threadIdx.x; i < n;
       i += (unsigned long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// Streaming copy of n 16-byte words: each workgroup copies one contiguous
// chunk of ceil(n / gridDim.x) words, every lane issuing U nontemporal loads
// (consecutive 4 KiB rows of the chunk) before its U nontemporal stores.  On
// MI355X this beat the grid-strided form by 3-15 % at every grid size.
typedef float stream_word __attribute__((ext_vector_type(4)));
template <int U>
__global__ void __launch_bounds__(256)
    stream_copy_kernel(const stream_word* __restrict__ src, stream_word* __restrict__ dst,
                       unsigned long long n) {
  const unsigned long long per = (n + gridDim.x - 1) / gridDim.x;
  const unsigned long long b = blockIdx.x * per;
  const unsigned long long e = b + per < n ? b + per : n;
  unsigned long long i = b + threadIdx.x;
  for (; i + (U - 1) * 256 < e; i += U * 256) {
    stream_word v[U];
#pragma unroll
    for (int u = 0; u < U; u++) v[u] = __builtin_nontemporal_load(&src[i + u * 256]);
#pragma unroll
    for (int u = 0; u < U; u++) __builtin_nontemporal_store(v[u], &dst[i + u * 256]);
  }
  for (; i < e; i += 256) __builtin_nontemporal_store(__builtin_nontemporal_load(&src[i]), &dst[i]);
}

template <typename T>
int guard_slabs(coffee_ctx* c, hipStream_t s, T* const dst[7], T* const src[7], bool wrap,
                int nfields = 7) {
  FlatK<T>& k = kparams<T>(c);
  Ptr7<T> d, r;
  for (int q = 0; q < 7; q++) {
    d.p[q] = dst[q];
    r.p[q] = src[q];
  }
  if (wrap) {
    // in place (dst == src): every periodic guard cell in one launch
    if (!k.per[0] && !k.per[1] && !k.per[2]) return COFFEE_OK;
    const long total = (long)k.dims[0] * k.dims[1] * k.dims[2] - (long)k.n[0] * k.n[1] * k.n[2];
    const int blocks = (int)std::min<long>((total + 255) / 256, 8192);
    return timed(c, "guard_fill", s, [&] {
      guard_wrap_kernel<T><<<blocks, 256, 0, s>>>(k, d, nfields);
    });
  }
  for (int ax = 0; ax < 3; ax++) {
    if (k.guard[ax] == 0) continue;
    if (wrap != (bool)k.per[ax]) continue;
    long ext = 2L * k.guard[ax];
    for (int a = 0; a < 3; a++)
      if (a != ax) ext *= k.dims[a];
    int blocks = (int)std::min<long>((ext + 255) / 256, 4096);
    int rc;
    if (wrap)
      rc = timed(c, "guard_fill", s, [&] {
        guard_slab_kernel<T, true><<<blocks, 256, 0, s>>>(k, ax, d, r, nfields);
      });
    else
      rc = timed(c, "guard_copy", s, [&] {
        guard_slab_kernel<T, false><<<blocks, 256, 0, s>>>(k, ax, d, r, nfields);
      });
    if (rc) return rc;
  }
  return COFFEE_OK;
}

// Planes per workgroup: one workgroup per CU (256 CUs) marching the whole
// z extent measured best at 512^3 (fewest warm-up planes, no chunk seams);
// grids with fewer tiles per plane split z so every CU still gets work.
int auto_zchunk(int tiles_per_plane, int nz) {
  int nzc = (256 + tiles_per_plane - 1) / tiles_per_plane;
  int zc = (nz + nzc - 1) / nzc;
  return zc < 8 ? 8 : zc;
}

// Tile (columns x rows) of the tiled kernels: variant 0 picks the measured
// best for the Scalar type on MI355X at 512^3 (FP32 128 x 8 = 1024 threads,
// 3.5 % faster per step than 64 x 16; FP64 64 x 8 = 512 threads, equal to
// 128 x 4); 2/3/4 force 64 x 4/8/16 (16 only in FP32), 5 and 6 wider rows
// (FP32 128 x 8, 256 x 4; FP64 128 x 4 for both: a 256 x 2 FP64 tile does not
// fit in LDS), 7 = 128 x 4; variant 1 is the thread-per-cell schedule.
template <typename T>
void tile_shape(int variant, int& tx, int& ty) {
  const bool f32 = sizeof(T) == 4;
  tx = f32 ? 128 : 64;
  ty = 8;
  if (variant == 2 || variant == 3 || variant == 4) {
    tx = 64;
    ty = variant == 2 ? 4 : (variant == 3 || !f32 ? 8 : 16);
  }
  if (variant == 5) {
    tx = 128;
    ty = f32 ? 8 : 4;
  }
  if (variant == 6) {
    tx = f32 ? 256 : 128;
    ty = 4;
  }
  if (variant == 7) {
    tx = 128;
    ty = 4;
  }
}

// The FP32 default tile (tile_shape variant 0) also has a Kreiss-Oliger
// kernel addressing the fields by 32-bit byte offsets, used by variant 0 when
// every array is below 4 GiB; other variants (5 is the same tile) and FP64
// use 64-bit addressing.
template <typename T, int TX, int TY>
constexpr bool has_o32() {
  return sizeof(T) == 4 && TX == 128 && TY == 8;
}

template <typename T, bool F, bool L, bool D, bool S, int TX, int TY>
void launch_stage(bool tiled, int tblocks, dim3 grid, dim3 block, hipStream_t s, const FlatK<T>& k,
                  const FlatIO<T>& io, T as, T bs, T ts, const TileMap& tm) {
  if (tiled)
    ez_stage_tiled<T, TX, TY, F, L, D, S><<<tblocks, TX * TY, 0, s>>>(k, io, as, bs, ts, tm);
  else
    ez_stage_kernel<T, F, L, D, S><<<grid, block, 0, s>>>(k, io, as, bs, ts);
}

template <typename T, bool F, bool L, int TX, int TY>
void launch_stage_fl(bool dg, bool star, bool tiled, int tblocks, dim3 grid, dim3 block,
                     hipStream_t s, const FlatK<T>& k, const FlatIO<T>& io, T as, T bs, T ts,
                     const TileMap& tm) {
  if (dg) {
    if (star) launch_stage<T, F, L, true, true, TX, TY>(tiled, tblocks, grid, block, s, k, io, as, bs, ts, tm);
    else launch_stage<T, F, L, true, false, TX, TY>(tiled, tblocks, grid, block, s, k, io, as, bs, ts, tm);
  } else {
    if (star) launch_stage<T, F, L, false, true, TX, TY>(tiled, tblocks, grid, block, s, k, io, as, bs, ts, tm);
    else launch_stage<T, F, L, false, false, TX, TY>(tiled, tblocks, grid, block, s, k, io, as, bs, ts, tm);
  }
}

template <typename T, int TX, int TY, bool S>
void launch_ko_tiled(bool o32, int tblocks, hipStream_t s, const FlatK<T>& k, const FlatIO<T>& io,
                     T t1, const TileMap& tm) {
  if constexpr (has_o32<T, TX, TY>()) {
    if (o32) {
      ez_ko_tiled<T, TX, TY, S, true><<<tblocks, TX * TY, 0, s>>>(k, io, t1, tm);
      return;
    }
  }
  ez_ko_tiled<T, TX, TY, S, false><<<tblocks, TX * TY, 0, s>>>(k, io, t1, tm);
}

// Buffers of P over the five stages: A -> B -> C -> B -> C -> A (fields 0-5
// alternate A -> B -> A -> B -> A -> B and Kreiss-Oliger brings them back to
// A).  Every stage reads P from another buffer than it writes, and the last
// stage writes P into the caller's array, so Kreiss-Oliger, which does not
// change P, neither reads nor writes it.
template <typename T>
void p_buffers(int st, T* a, T* b, T* cb, FlatIO<T>& io) {
  static const int from[5] = {0, 1, 2, 1, 2}, to[5] = {1, 2, 1, 2, 0};
  T* const buf[3] = {a, b, cb};
  io.in[6] = buf[from[st]];
  io.out[6] = buf[to[st]];
}

template <typename T, int TX, int TY>
int ez_launch_all(coffee_ctx* c, const FlatK<T>& k, FlatIO<T> io, T* const A[7], T* const B[7],
                  T* const Cp, double time, hipStream_t s, bool star) {
  static const double As[5] = {0, -0.4178904745, -1.192151694643, -1.697784692471,
                               -1.514183444257};
  static const double Bs[5] = {0.1496590219993, 0.3792103129999, 0.8229550293869,
                               0.6994504559488, 0.1530572479681};
  static const double cs[5] = {0, 0.1496590219993, 0.3704009573644, 0.6222557631345,
                               0.9582821306784};
  dim3 block(64, 4, 1);
  dim3 grid((k.n[0] + 63) / 64, (k.n[1] + 3) / 4, k.n[2]);
  TileMap tm;
  tm.ntx = (k.n[0] + TX - 1) / TX;
  tm.nty = (k.n[1] + TY - 1) / TY;
  tm.zc = c->zchunk > 0 ? c->zchunk : auto_zchunk(tm.ntx * tm.nty, k.n[2]);
  if (tm.zc > k.n[2]) tm.zc = k.n[2];
  tm.nzc = (k.n[2] + tm.zc - 1) / tm.zc;
  const int tblocks = tm.ntx * tm.nty * tm.nzc;
  const bool tiled = c->variant != 1;
  const bool dg = c->diagnostics != 0;
  // every field array (dims0 x dims1 x dims2 scalars) addressable by a
  // 32-bit byte offset (Kreiss-Oliger kernel)
  const bool o32 = c->variant == 0 &&
                   (unsigned long long)k.dims[0] * k.dims[1] * k.dims[2] * sizeof(T) < (1ull << 32);
  const T dt = k.dt;
  const T t0 = (T)time;
  int rc = COFFEE_OK;
  for (int st = 0; st < 5; st++) {
    T* const* in = (st % 2 == 0) ? A : B;
    T* const* out = (st % 2 == 0) ? B : A;
    for (int q = 0; q < 6; q++) {
      io.in[q] = in[q];
      io.out[q] = out[q];
    }
    p_buffers<T>(st, A[6], B[6], Cp, io);
    const T as = (T)As[st], bs = (T)Bs[st];
    const T ts = t0 + (T)cs[st] * dt;
    const char* name = st == 0 ? "ez_stage_first" : (st == 4 ? "ez_stage_last" : "ez_stage");
    rc = timed(c, name, s, [&] {
      if (st == 0)
        launch_stage_fl<T, true, false, TX, TY>(dg, star, tiled, tblocks, grid, block, s, k, io, as, bs, ts, tm);
      else if (st == 4)
        launch_stage_fl<T, false, true, TX, TY>(dg, star, tiled, tblocks, grid, block, s, k, io, as, bs, ts, tm);
      else
        launch_stage_fl<T, false, false, TX, TY>(dg, star, tiled, tblocks, grid, block, s, k, io, as, bs, ts, tm);
    });
    if (rc) return rc;
  }
  // Kreiss-Oliger on E, B: B -> A (P is already in A)
  for (int q = 0; q < 6; q++) {
    io.in[q] = B[q];
    io.out[q] = A[q];
  }
  io.in[6] = io.out[6] = nullptr;
  const T t1 = t0 + dt;
  return timed(c, "ez_ko", s, [&] {
    if (dg) {
      // diagnostics mode: the one-thread-per-cell KO kernel with the energy
      // accounting (bitwise the same field results as the tiled one)
      if (star) ez_ko_kernel<T, true, true><<<grid, block, 0, s>>>(k, io, t1);
      else ez_ko_kernel<T, false, true><<<grid, block, 0, s>>>(k, io, t1);
    } else if (tiled) {
      if (star) launch_ko_tiled<T, TX, TY, true>(o32, tblocks, s, k, io, t1, tm);
      else launch_ko_tiled<T, TX, TY, false>(o32, tblocks, s, k, io, t1, tm);
    } else {
      if (star) ez_ko_kernel<T, true><<<grid, block, 0, s>>>(k, io, t1);
      else ez_ko_kernel<T, false><<<grid, block, 0, s>>>(k, io, t1);
    }
  });
}

template <typename T>
int ez_step_impl(coffee_ctx* c, const coffee_fields* f, double time, hipStream_t s) {
  static const double As[5] = {0, -0.4178904745, -1.192151694643, -1.697784692471,
                               -1.514183444257};
  static const double Bs[5] = {0.1496590219993, 0.3792103129999, 0.8229550293869,
                               0.6994504559488, 0.1530572479681};
  static const double cs[5] = {0, 0.1496590219993, 0.3704009573644, 0.6222557631345,
                               0.9582821306784};
  int rc = check_fields<T>(c, f, true);
  if (rc) return rc;
  FlatK<T>& k = kparams<T>(c);
  if (k.pml_any) {
    for (int q = 0; q < 3; q++)
      if (!f->Etmp[q] || !f->Btmp[q] || !f->Bbg[q])
        return fail(c, COFFEE_ERR_INVALID, "absorbing layer needs Etmp, Btmp and Bbg");
    if (!f->Ptmp) return fail(c, COFFEE_ERR_INVALID, "absorbing layer needs Ptmp");
  }
  if (c->diagnostics) {
    for (int q = 0; q < 3; q++)
      if (!f->J[q]) return fail(c, COFFEE_ERR_INVALID, "diagnostics need J");
    if (!f->divE || !f->divB) return fail(c, COFFEE_ERR_INVALID, "diagnostics need divE/divB");
  }
  T* A[7];
  T* B[7];
  state_ptrs<T>(f, A);
  for (int q = 0; q < 7; q++) B[q] = (T*)c->scratch[q];
  T* const Cp = (T*)c->scratch[7];
  if (!Cp) return fail(c, COFFEE_ERR_STATE, "flat step needs the third P buffer");

  // Frozen (non-periodic) guards must be present in the partner buffers.
  rc = guard_slabs<T>(c, s, B, A, false);
  if (rc) return rc;
  {
    T* const cd[7] = {Cp, Cp, Cp, Cp, Cp, Cp, Cp};
    T* const cs[7] = {A[6], A[6], A[6], A[6], A[6], A[6], A[6]};
    rc = guard_slabs<T>(c, s, cd, cs, false, 1);
    if (rc) return rc;
  }

  FlatIO<T> io{};
  for (int q = 0; q < 3; q++) {
    io.dU[q] = (T*)f->dE[q];
    io.dU[3 + q] = (T*)f->dB[q];
    io.Un[q] = (T*)f->Etmp[q];
    io.Un[3 + q] = (T*)f->Btmp[q];
    io.Bbg[q] = (const T*)f->Bbg[q];
    io.J[q] = (T*)f->J[q];
  }
  io.dU[6] = (T*)f->dP;
  io.Un[6] = (T*)f->Ptmp;
  io.divE = (T*)f->divE;
  io.divB = (T*)f->divB;
  if (c->diagnostics) {
    void* const acc[6] = {f->dU_KO, f->dU_Epar, f->dU_EgtB, f->dU_KO_cum, f->dU_Epar_cum,
                          f->dU_EgtB_cum};
    for (int q = 0; q < 6; q++) io.acc[q] = (T*)acc[q];
    // the per-step arrays start every step at zero (field_solver_EZ.cpp:1020-1022)
    const unsigned long long n = cells(c->grid);
    for (int q = 0; q < 3; q++)
      if (io.acc[q]) {
        rc = timed(c, "fill", s, [&] {
          fill_kernel<T><<<(int)std::min<unsigned long long>((n + 255) / 256, 8192), 256, 0, s>>>(
              io.acc[q], n, T(0));
        });
        if (rc) return rc;
      }
  }

  const bool star = k.problem == 1 || k.problem == 2;
  int tx = 64, ty = 8;
  tile_shape<T>(c->variant, tx, ty);
  constexpr bool f32 = sizeof(T) == 4;
  if (tx == 128 && ty == 4) rc = ez_launch_all<T, 128, 4>(c, k, io, A, B, Cp, time, s, star);
  else if (tx == 128) rc = ez_launch_all<T, 128, f32 ? 8 : 4>(c, k, io, A, B, Cp, time, s, star);
  else if (tx == 256) rc = ez_launch_all<T, f32 ? 256 : 128, 4>(c, k, io, A, B, Cp, time, s, star);
  else if (ty == 4) rc = ez_launch_all<T, 64, 4>(c, k, io, A, B, Cp, time, s, star);
  else if (ty == 16) rc = ez_launch_all<T, 64, 16>(c, k, io, A, B, Cp, time, s, star);
  else rc = ez_launch_all<T, 64, 8>(c, k, io, A, B, Cp, time, s, star);
  if (rc) return rc;
  // single-GPU exchange: periodic guards of the caller's arrays
  return guard_slabs<T>(c, s, A, A, true);
}

template <typename T>
int residuals_impl(coffee_ctx* c, const coffee_fields* f, double out[5], hipStream_t s) {
  int rc = check_fields<T>(c, f, false);
  if (rc) return rc;
  FlatK<T>& k = kparams<T>(c);
  FlatIO<T> io{};
  T* A[7];
  state_ptrs<T>(f, A);
  for (int q = 0; q < 7; q++) io.in[q] = A[q];
  rc = timed(c, "residuals", s, [&] {
    residual_kernel<T><<<c->part_blocks, 256, 0, s>>>(k, io, c->d_part);
  });
  if (rc) return rc;
  HIP_TRY(c, hipMemcpyAsync(c->h_part.data(), c->d_part, sizeof(double) * NRED * c->part_blocks,
                            hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  fold_residuals(c->h_part.data(), c->part_blocks, out);
  return COFFEE_OK;
}

template <typename T>
int energy_impl(coffee_ctx* c, void* const comp[3], double* out, hipStream_t s) {
  for (int q = 0; q < 3; q++)
    if (!comp[q]) return fail(c, COFFEE_ERR_INVALID, "total_energy needs 3 component pointers");
  FlatK<T>& k = kparams<T>(c);
  int rc = timed(c, "total_energy", s, [&] {
    energy_kernel<T><<<c->part_blocks, 256, 0, s>>>(k, (const T*)comp[0], (const T*)comp[1],
                                                    (const T*)comp[2], c->d_part);
  });
  if (rc) return rc;
  HIP_TRY(c, hipMemcpyAsync(c->h_part.data(), c->d_part, sizeof(double) * c->part_blocks,
                            hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  double w = 0;
  for (int b = 0; b < c->part_blocks; b++) w += c->h_part[b];
  *out = w;
  return COFFEE_OK;
}

// Partner state of coffee_ez_step_noalloc.  Host code never uses this name
// as an address: each device's copy is reached through the pointer that
// hipGetSymbolAddress returns with that device current (noalloc_scratch).
static __device__ __align__(256) unsigned char scratch[COFFEE_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device `dev`, looked up on the
// first call on that device and kept for the process.
int noalloc_scratch(int dev, void** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(nullptr, COFFEE_ERR_INVALID, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(nullptr, COFFEE_ERR_HIP,
                  std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return COFFEE_OK;
}

}  // namespace

extern "C" {

int coffee_abi_version(void) { return COFFEE_ABI_VERSION; }

const char* coffee_last_error(const coffee_ctx* ctx) {
  return ctx ? ctx->err.c_str() : coffee::create_error().c_str();
}

int coffee_ctx_create(int solver, int scalar_bytes, const coffee_grid* grid,
                      const coffee_params* params, coffee_ctx** out) {
  if (!out || !grid || !params) return fail(nullptr, COFFEE_ERR_INVALID, "null argument");
  *out = nullptr;
  if (solver != COFFEE_SOLVER_EZ && solver != COFFEE_SOLVER_EZ_SPHERICAL &&
      solver != COFFEE_SOLVER_GR_EZ)
    return fail(nullptr, COFFEE_ERR_INVALID, "unknown solver variant");
  if (scalar_bytes != COFFEE_F32 && scalar_bytes != COFFEE_F64)
    return fail(nullptr, COFFEE_ERR_INVALID, "scalar_bytes must be 4 or 8");
  int rc = solver != COFFEE_SOLVER_EZ_SPHERICAL ? validate(nullptr, grid, params) : COFFEE_OK;
  if (rc) return rc;
  coffee_ctx* c = new coffee_ctx();
  c->solver = solver;
  c->scalar = scalar_bytes;
  c->grid = *grid;
  c->params = *params;
  if (hipGetDevice(&c->device) != hipSuccess) {
    delete c;
    return fail(nullptr, COFFEE_ERR_HIP, "hipGetDevice failed");
  }
  fill_k<float>(*grid, *params, c->kf);
  fill_k<double>(*grid, *params, c->kd);
  c->part_blocks = 1024;
  c->h_part.assign((size_t)NRED * c->part_blocks, 0.0);
  if (hipMalloc(&c->d_part, sizeof(double) * NRED * c->part_blocks) != hipSuccess) {
    coffee_ctx_destroy(c);
    return fail(nullptr, COFFEE_ERR_HIP, "hipMalloc reduction buffer failed");
  }
  if (solver == COFFEE_SOLVER_EZ_SPHERICAL) {
    int rs = coffee::sph_create(c);
    if (rs) {
      std::string m = c->err;
      coffee_ctx_destroy(c);
      return fail(nullptr, rs, m);
    }
    *out = c;
    return COFFEE_OK;
  }
  size_t bytes = cells(*grid) * (size_t)scalar_bytes;
  const int nscratch = solver == COFFEE_SOLVER_EZ ? 8 : 7;
  for (int q = 0; q < nscratch; q++) {
    hipError_t e = hipMalloc(&c->scratch[q], bytes);
    if (e != hipSuccess) {
      std::string m = std::string("hipMalloc scratch: ") + hipGetErrorString(e);
      coffee_ctx_destroy(c);
      return fail(nullptr, COFFEE_ERR_HIP, m);
    }
  }
  if (solver == COFFEE_SOLVER_GR_EZ) {
    int rg = coffee::gr_create(c);
    if (rg) {
      std::string m = c->err;
      coffee_ctx_destroy(c);
      return fail(nullptr, rg, m);
    }
  }
  *out = c;
  return COFFEE_OK;
}

int coffee_ctx_destroy(coffee_ctx* c) {
  if (!c) return COFFEE_OK;
  for (int q = 0; q < 8; q++)
    if (c->scratch[q]) (void)hipFree(c->scratch[q]);
  if (c->d_part) (void)hipFree(c->d_part);
  delete c->state;
  for (auto& r : c->trecs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto e : c->event_pool) (void)hipEventDestroy(e);
  delete c;
  return COFFEE_OK;
}

int coffee_set_params(coffee_ctx* c, const coffee_params* p) {
  if (!c || !p) return fail(c, COFFEE_ERR_INVALID, "null argument");
  if (c->solver != COFFEE_SOLVER_EZ)
    return fail(c, COFFEE_ERR_STATE, "set_params is available for the flat solver");
  int rc = validate(c, &c->grid, p);
  if (rc) return rc;
  c->params = *p;
  fill_k<float>(c->grid, *p, c->kf);
  fill_k<double>(c->grid, *p, c->kd);
  return COFFEE_OK;
}

int coffee_ez_step(coffee_ctx* c, const coffee_fields* f, double time, void* stream) {
  if (!c || !f) return fail(c, COFFEE_ERR_INVALID, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if (c->solver == COFFEE_SOLVER_EZ_SPHERICAL) return coffee::sph_step(c, f, time, s);
  if (c->solver == COFFEE_SOLVER_GR_EZ) return coffee::gr_step(c, f, time, s);
  return c->scalar == 4 ? ez_step_impl<float>(c, f, time, s) : ez_step_impl<double>(c, f, time, s);
}

int coffee_ez_step_noalloc(const coffee_grid* grid, const coffee_params* params, int scalar_bytes,
                           const coffee_fields* f, double time, void* stream) {
  if (!grid || !params || !f) return fail(nullptr, COFFEE_ERR_INVALID, "null argument");
  if (scalar_bytes != COFFEE_F32 && scalar_bytes != COFFEE_F64)
    return fail(nullptr, COFFEE_ERR_INVALID, "scalar_bytes must be 4 or 8");
  int rc = validate(nullptr, grid, params);
  if (rc) return rc;
  const unsigned long long per = (cells(*grid) * (unsigned long long)scalar_bytes + 255ull) & ~255ull;
  if (8ull * per > COFFEE_NOALLOC_SCRATCH_BYTES)
    return fail(nullptr, COFFEE_ERR_SIZE,
                "coffee_ez_step_noalloc: the partner state (8 x " + std::to_string(per) +
                    " bytes) exceeds COFFEE_NOALLOC_SCRATCH_BYTES (" +
                    std::to_string(COFFEE_NOALLOC_SCRATCH_BYTES) + ")");
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return fail(nullptr, COFFEE_ERR_HIP, "hipGetDevice failed");
  void* base = nullptr;
  rc = noalloc_scratch(dev, &base);
  if (rc) return rc;
  // A context on the stack: the parameter blocks and the partner pointers,
  // no device allocation, timing and diagnostics off, default schedule.
  coffee_ctx c;
  c.solver = COFFEE_SOLVER_EZ;
  c.scalar = scalar_bytes;
  c.device = dev;
  c.grid = *grid;
  c.params = *params;
  fill_k<float>(*grid, *params, c.kf);
  fill_k<double>(*grid, *params, c.kd);
  for (int q = 0; q < 8; q++) c.scratch[q] = (unsigned char*)base + q * per;
  hipStream_t s = (hipStream_t)stream;
  rc = scalar_bytes == 4 ? ez_step_impl<float>(&c, f, time, s) : ez_step_impl<double>(&c, f, time, s);
  if (rc) return fail(nullptr, rc, c.err);
  return COFFEE_OK;
}

int coffee_fill_guards(coffee_ctx* c, const coffee_fields* f, void* stream) {
  if (!c || !f) return fail(c, COFFEE_ERR_INVALID, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if (c->scalar == 4) {
    int rc = check_fields<float>(c, f, false);
    if (rc) return rc;
    float* A[7];
    state_ptrs<float>(f, A);
    return guard_slabs<float>(c, s, A, A, true);
  }
  int rc = check_fields<double>(c, f, false);
  if (rc) return rc;
  double* A[7];
  state_ptrs<double>(f, A);
  return guard_slabs<double>(c, s, A, A, true);
}

int coffee_total_energy(coffee_ctx* c, void* const comp[3], double* out, void* stream) {
  if (!c || !comp || !out) return fail(c, COFFEE_ERR_INVALID, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if (c->solver == COFFEE_SOLVER_EZ_SPHERICAL) return coffee::sph_energy(c, comp, out, s);
  if (c->solver == COFFEE_SOLVER_GR_EZ) return coffee::gr_energy(c, comp, out, s);
  return c->scalar == 4 ? energy_impl<float>(c, comp, out, s) : energy_impl<double>(c, comp, out, s);
}

int coffee_residuals(coffee_ctx* c, const coffee_fields* f, double out[5], void* stream) {
  if (!c || !f || !out) return fail(c, COFFEE_ERR_INVALID, "null argument");
  hipStream_t s = (hipStream_t)stream;
  if (c->solver == COFFEE_SOLVER_EZ_SPHERICAL) {
    if (c->is_batch)
      return fail(c, COFFEE_ERR_STATE, "batches report residuals per run: coffee_sph_batch_residuals");
    return coffee::sph_residuals(c, f, out, s);
  }
  if (c->solver == COFFEE_SOLVER_GR_EZ) return coffee::gr_residuals(c, f, out, s);
  return c->scalar == 4 ? residuals_impl<float>(c, f, out, s) : residuals_impl<double>(c, f, out, s);
}

int coffee_set_diagnostics(coffee_ctx* c, int enable) {
  if (!c) return fail(c, COFFEE_ERR_INVALID, "null argument");
  c->diagnostics = enable ? 1 : 0;
  return COFFEE_OK;
}

int coffee_fill(void* p, unsigned long long n, int scalar_bytes, double value, void* stream) {
  if (!p && n) return fail(nullptr, COFFEE_ERR_INVALID, "null pointer");
  if (scalar_bytes != 4 && scalar_bytes != 8)
    return fail(nullptr, COFFEE_ERR_INVALID, "scalar_bytes must be 4 or 8");
  if (n == 0) return COFFEE_OK;
  hipStream_t s = (hipStream_t)stream;
  const int blocks = (int)std::min<unsigned long long>((n + 255) / 256, 8192);
  if (scalar_bytes == 4)
    fill_kernel<float><<<blocks, 256, 0, s>>>((float*)p, n, (float)value);
  else
    fill_kernel<double><<<blocks, 256, 0, s>>>((double*)p, n, value);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(nullptr, COFFEE_ERR_HIP, std::string("fill: ") + hipGetErrorString(e));
  return COFFEE_OK;
}

int coffee_stream_copy(const void* src, void* dst, unsigned long long bytes, int unroll,
                       int blocks, void* stream) {
  if (!src || !dst) return fail(nullptr, COFFEE_ERR_INVALID, "null pointer");
  if (bytes % 16 || ((unsigned long long)src | (unsigned long long)dst) % 16)
    return fail(nullptr, COFFEE_ERR_INVALID, "bytes and pointers must be multiples of 16");
  if (blocks < 1 || blocks > 65536) return fail(nullptr, COFFEE_ERR_INVALID, "blocks out of range");
  hipStream_t s = (hipStream_t)stream;
  const unsigned long long n = bytes / 16;
  const stream_word* a = (const stream_word*)src;
  stream_word* b = (stream_word*)dst;
  switch (unroll) {
    case 1: stream_copy_kernel<1><<<blocks, 256, 0, s>>>(a, b, n); break;
    case 2: stream_copy_kernel<2><<<blocks, 256, 0, s>>>(a, b, n); break;
    case 4: stream_copy_kernel<4><<<blocks, 256, 0, s>>>(a, b, n); break;
    case 8: stream_copy_kernel<8><<<blocks, 256, 0, s>>>(a, b, n); break;
    default: return fail(nullptr, COFFEE_ERR_INVALID, "unroll must be 1, 2, 4 or 8");
  }
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(nullptr, COFFEE_ERR_HIP, std::string("stream_copy: ") + hipGetErrorString(e));
  return COFFEE_OK;
}

int coffee_fill_guards_arrays(const coffee_grid* grid, const int periodic[3], void* const* arrays,
                              int narrays, int scalar_bytes, void* stream) {
  if (!grid || !periodic || (!arrays && narrays > 0))
    return fail(nullptr, COFFEE_ERR_INVALID, "null argument");
  if (scalar_bytes != 4 && scalar_bytes != 8)
    return fail(nullptr, COFFEE_ERR_INVALID, "scalar_bytes must be 4 or 8");
  coffee_params p{};
  p.dt = 1;
  p.tau = 1;
  for (int a = 0; a < 3; a++) p.periodic[a] = periodic[a];
  for (int a = 0; a < 3; a++)
    if (periodic[a] && (grid->guard[a] > grid->dims[a] - 2 * grid->guard[a] ||
                        grid->dims[a] - 2 * grid->guard[a] < 1))
      return fail(nullptr, COFFEE_ERR_INVALID, "periodic axis shorter than its guard");
  coffee_ctx tmp;
  fill_k<float>(*grid, p, tmp.kf);
  fill_k<double>(*grid, p, tmp.kd);
  hipStream_t s = (hipStream_t)stream;
  for (int base = 0; base < narrays; base += 7) {
    const int m = std::min(7, narrays - base);
    int rc;
    if (scalar_bytes == 4) {
      float* A[7] = {};
      for (int q = 0; q < m; q++) A[q] = (float*)arrays[base + q];
      for (int q = m; q < 7; q++) A[q] = A[0];
      rc = guard_slabs<float>(&tmp, s, A, A, true, m);
    } else {
      double* A[7] = {};
      for (int q = 0; q < m; q++) A[q] = (double*)arrays[base + q];
      for (int q = m; q < 7; q++) A[q] = A[0];
      rc = guard_slabs<double>(&tmp, s, A, A, true, m);
    }
    if (rc) return fail(nullptr, rc, tmp.err);
  }
  return COFFEE_OK;
}

int coffee_set_schedule(coffee_ctx* c, int variant, int zchunk) {
  if (!c) return fail(c, COFFEE_ERR_INVALID, "null argument");
  if (variant < 0 || variant > 7) return fail(c, COFFEE_ERR_INVALID, "variant must be 0..7");
  if (zchunk < 0) return fail(c, COFFEE_ERR_INVALID, "zchunk must be >= 0");
  c->variant = variant;
  c->zchunk = zchunk;
  return COFFEE_OK;
}

int coffee_set_timing(coffee_ctx* c, int enable) {
  if (!c) return fail(c, COFFEE_ERR_INVALID, "null argument");
  c->timing = enable ? 1 : 0;
  return COFFEE_OK;
}

int coffee_reset_timing(coffee_ctx* c) {
  if (!c) return fail(c, COFFEE_ERR_INVALID, "null argument");
  for (auto& r : c->trecs) {
    c->event_pool.push_back(r.a);
    c->event_pool.push_back(r.b);
  }
  c->trecs.clear();
  return COFFEE_OK;
}

int coffee_get_timing(coffee_ctx* c, int max_entries, char (*names)[64], double* total_ms,
                      int* launches, int* n_entries) {
  if (!c || !n_entries) return fail(c, COFFEE_ERR_INVALID, "null argument");
  int n = (int)c->tnames.size();
  std::vector<double> ms(n, 0.0);
  std::vector<int> cnt(n, 0);
  if (!c->trecs.empty()) HIP_TRY(c, hipEventSynchronize(c->trecs.back().b));
  for (auto& r : c->trecs) {
    float t = 0;
    HIP_TRY(c, hipEventElapsedTime(&t, r.a, r.b));
    ms[r.name] += t;
    cnt[r.name] += 1;
  }
  int m = std::min(n, max_entries);
  for (int q = 0; q < m; q++) {
    if (names) {
      std::strncpy(names[q], c->tnames[q].c_str(), 63);
      names[q][63] = 0;
    }
    if (total_ms) total_ms[q] = ms[q];
    if (launches) launches[q] = cnt[q];
  }
  *n_entries = m;
  return COFFEE_OK;
}

}  // extern "C"
