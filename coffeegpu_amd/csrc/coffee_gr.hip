// Kerr-Schild GR EZ solver behind the C ABI (COFFEE_SOLVER_GR_EZ).
//
// field_solver_gr_EZ::evolve_fields (src/algorithms/field_solver_gr_EZ.cu:862-910)
// on one GPU: per stage a point sweep (E_i, H_i, sqrt(gamma) D, sqrt(gamma) B
// on the guarded grid) and a fused stage sweep; then a fused KO sweep.  U
// ping-pongs between the caller's arrays and the context's partner buffers
// (6 sweeps per step, so the result lands back in the caller's arrays).
// Non-periodic guards are frozen, as on one rank of the reference; a periodic
// axis is filled after every stage, as send_guard_cells does.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "coffee_hip.h"
#include "ctx.hpp"
#include "ez_gr.hpp"
#include "ez_gr_tiled.hpp"

namespace coffee {

namespace {

const double kAs[5] = {0, -0.4178904745, -1.192151694643, -1.697784692471, -1.514183444257};
const double kBs[5] = {0.1496590219993, 0.3792103129999, 0.8229550293869, 0.6994504559488,
                       0.1530572479681};

struct GrState : SolverState {
  void* W[12] = {};
  // allocated on first use by the schedules that need them: W[0..6] hold the
  // halo metric table of variant 2 until the two-sweep schedule (variant 1)
  // overwrites them
  bool mtab_valid = false;
  GrK<float> kf{};
  GrK<double> kd{};
  ~GrState() override {
    for (void* p : W)
      if (p) (void)hipFree(p);
  }
};

template <typename T>
void fill_gr(const coffee_grid& g, const coffee_params& p, GrK<T>& k) {
  for (int a = 0; a < 3; a++) {
    k.dims[a] = g.dims[a];
    k.guard[a] = g.guard[a];
    k.n[a] = g.dims[a] - 2 * g.guard[a];
    k.lower[a] = p.lower[a];
    k.delta[a] = p.size[a] / k.n[a];
    k.c12[a] = (T)(1.0 / k.delta[a]) / T(12);
    k.xh[a] = p.lower[a] + p.size[a] - p.pml[a] * k.delta[a];
    k.xl[a] = p.lower[a] + p.pml[a] * k.delta[a];
  }
  k.sy = g.dims[0];
  k.sz = g.dims[0] * g.dims[1];
  k.a = p.a;
  k.dt = (T)p.dt;
  k.tau = (T)p.tau;
  k.itau = (T)(1.0 / p.tau);
  k.KOfac = -(T)p.KOeps * (T)(-1.0 / 64.0);
  k.calc_current = p.calc_current;
  k.divB_clean = p.divB_clean;
  k.clean_ep = p.clean_ep;
  k.check_egb = p.check_egb;
  k.pml_any = (p.pml[0] > 0 || p.pml[1] > 0 || p.pml[2] > 0) ? 1 : 0;
  k.pmlscale = p.pmllen * k.delta[0];
  k.sigpml = p.sigpml;
  const double rH = 1.0 + std::sqrt(1.0 - p.a * p.a);
  k.r1 = 0.7 * rH;
  k.r2 = 0.01 * rH;
  k.dd = 0.2 * rH;
}

GrState* gr_state(coffee_ctx* c) { return static_cast<GrState*>(c->state); }

template <typename T>
GrK<T>& grk(GrState* S);
template <>
GrK<float>& grk<float>(GrState* S) { return S->kf; }
template <>
GrK<double>& grk<double>(GrState* S) { return S->kd; }

// Guard cells with the flat solver's kernels: periodic ones in place in one
// pass (wrap, dst == src), frozen ones copied slab by slab (src -> dst).
template <typename T>
int gr_guards(coffee_ctx* c, hipStream_t s, T* const dst[7], T* const src[7], bool wrap) {
  const FlatK<T>& k = flat_k<T>(c);
  Ptr7<T> d, r;
  for (int q = 0; q < 7; q++) {
    d.p[q] = dst[q];
    r.p[q] = src[q];
  }
  if (wrap) {
    if (!k.per[0] && !k.per[1] && !k.per[2]) return COFFEE_OK;
    const long total = (long)k.dims[0] * k.dims[1] * k.dims[2] - (long)k.n[0] * k.n[1] * k.n[2];
    const int blocks = (int)std::min<long>((total + 255) / 256, 8192);
    return timed(c, "guard_fill", s, [&] { guard_wrap_kernel<T><<<blocks, 256, 0, s>>>(k, d, 7); });
  }
  for (int ax = 0; ax < 3; ax++) {
    if (wrap != (bool)k.per[ax]) continue;
    long ext = 2L * k.guard[ax];
    for (int a = 0; a < 3; a++)
      if (a != ax) ext *= k.dims[a];
    int blocks = (int)std::min<long>((ext + 255) / 256, 4096);
    int rc = timed(c, wrap ? "guard_fill" : "guard_copy", s, [&] {
      if (wrap) guard_slab_kernel<T, true><<<blocks, 256, 0, s>>>(k, ax, d, r, 7);
      else guard_slab_kernel<T, false><<<blocks, 256, 0, s>>>(k, ax, d, r, 7);
    });
    if (rc) return rc;
  }
  return COFFEE_OK;
}

template <typename T, bool F, bool L>
void launch_gr_stage(bool dg, dim3 grid, dim3 block, hipStream_t s, const GrK<T>& k,
                     const GrIO<T>& io, T as, T bs) {
  if (dg) gr_stage_kernel<T, F, L, true><<<grid, block, 0, s>>>(k, io, as, bs);
  else gr_stage_kernel<T, F, L, false><<<grid, block, 0, s>>>(k, io, as, bs);
}

template <typename T, int TY, bool F, bool L, bool MT, int HD>
void launch_gr_tiled_mt(bool dg, int blocks, hipStream_t s, const GrK<T>& k, const GrIO<T>& io,
                        T as, T bs, const TileMap& tm) {
  if (dg) gr_stage_tiled<T, TY, F, L, true, MT, HD><<<blocks, 64 * TY, 0, s>>>(k, io, as, bs, tm);
  else gr_stage_tiled<T, TY, F, L, false, MT, HD><<<blocks, 64 * TY, 0, s>>>(k, io, as, bs, tm);
}

template <typename T, int TY, bool F, bool L>
void launch_gr_tiled(bool dg, bool mt, bool hd3, int blocks, hipStream_t s, const GrK<T>& k,
                     const GrIO<T>& io, T as, T bs, const TileMap& tm) {
  // the x-y halo three planes ahead in the FP32 middle stages (2.043 ->
  // 2.028 ms at 384^3; the first and last stage measured slower with it, and
  // in FP64 the deeper queue spills)
  constexpr int HD = sizeof(T) == 4 && !F && !L ? 3 : 1;
  if (mt) launch_gr_tiled_mt<T, TY, F, L, true, 1>(dg, blocks, s, k, io, as, bs, tm);
  else if (hd3 && HD > 1) launch_gr_tiled_mt<T, TY, F, L, false, HD>(dg, blocks, s, k, io, as, bs, tm);
  else launch_gr_tiled_mt<T, TY, F, L, false, 1>(dg, blocks, s, k, io, as, bs, tm);
}

// Rows per tile of the tiled stage kernel: 8 in FP32 (512 threads), 4 in FP64
// (256 threads; the register queues double in size).
template <typename T>
constexpr int gr_tile_rows() {
  return sizeof(T) == 4 ? 8 : 4;
}

template <typename T>
int gr_step_impl(coffee_ctx* c, const coffee_fields* f, hipStream_t s) {
  GrState* S = gr_state(c);
  const GrK<T>& k = grk<T>(S);
  for (int q = 0; q < 3; q++)
    if (!f->E[q] || !f->B[q] || !f->dE[q] || !f->dB[q])
      return fail(c, COFFEE_ERR_INVALID, "E (D), B, dE, dB must be set");
  if (!f->P || !f->dP) return fail(c, COFFEE_ERR_INVALID, "P, dP must be set");
  // the horizon excision (.cu:617-701) always runs, so the step-start copies are needed
  for (int q = 0; q < 3; q++)
    if (!f->Etmp[q] || !f->Btmp[q])
      return fail(c, COFFEE_ERR_INVALID, "the GR step needs Etmp (Dtmp), Btmp and Ptmp");
  if (!f->Ptmp) return fail(c, COFFEE_ERR_INVALID, "the GR step needs Ptmp");
  const bool dg = c->diagnostics != 0;
  if (dg) {
    for (int q = 0; q < 3; q++)
      if (!f->J[q]) return fail(c, COFFEE_ERR_INVALID, "diagnostics need J");
    if (!f->divE || !f->divB) return fail(c, COFFEE_ERR_INVALID, "diagnostics need divE/divB");
  }
  T* A[7];
  T* B[7];
  for (int q = 0; q < 3; q++) {
    A[q] = (T*)f->E[q];
    A[3 + q] = (T*)f->B[q];
  }
  A[6] = (T*)f->P;
  for (int q = 0; q < 7; q++) B[q] = (T*)c->scratch[q];
  int rc = gr_guards<T>(c, s, B, A, false);
  if (rc) return rc;

  GrIO<T> io{};
  for (int q = 0; q < 3; q++) {
    io.dU[q] = (T*)f->dE[q];
    io.dU[3 + q] = (T*)f->dB[q];
    io.Un[q] = (T*)f->Etmp[q];
    io.Un[3 + q] = (T*)f->Btmp[q];
    io.J[q] = (T*)f->J[q];
  }
  io.dU[6] = (T*)f->dP;
  io.Un[6] = (T*)f->Ptmp;
  io.divE = (T*)f->divE;
  io.divB = (T*)f->divB;
  for (int q = 0; q < 12; q++) io.W[q] = (T*)S->W[q];

  const int total = k.dims[0] * k.dims[1] * k.dims[2];
  const int pblocks = std::min((total + 255) / 256, 8192);
  dim3 block(64, 4, 1);
  dim3 grid((k.n[0] + 63) / 64, (k.n[1] + 3) / 4, k.n[2]);
  // variant 1: point sweep + stage sweep; others: tiled stage kernel, which
  // computes every point's metric on the fly, except variant 2, which reads
  // the halo points' metric in stages 2-5 from a table filled once (at 384^3
  // FP32 with the FP32 solver's Newton-refined metric: middle stage 2.06 ms
  // on the fly against 2.33 with the table, whose reads add 1.2 GB per
  // launch).  Stage 1 computes it always.
  // the tiled stage kernel addresses the fields by 32-bit byte offsets:
  // arrays of 4 GiB or more take the point schedule
  const bool fits32 = (size_t)k.dims[0] * k.dims[1] * k.dims[2] * sizeof(T) < (1ull << 32);
  const bool tiled = c->variant != 1 && fits32;
  const bool mt = tiled && c->variant == 2;
  if (!tiled || mt) {
    // the point sweep's E, H, sqrt(g) D, sqrt(g) B and the table live in
    // scratch arrays allocated on first use
    const size_t bytes = (size_t)k.dims[0] * k.dims[1] * k.dims[2] * sizeof(T);
    for (int q = 0; q < 12; q++)
      if (!S->W[q]) {
        hipError_t e = hipMalloc(&S->W[q], bytes);
        if (e != hipSuccess)
          return fail(c, COFFEE_ERR_HIP, std::string("hipMalloc GR scratch: ") + hipGetErrorString(e));
      }
    for (int q = 0; q < 12; q++) io.W[q] = (T*)S->W[q];
  }
  if (!tiled) S->mtab_valid = false;  // the point sweep overwrites W
  if (mt && !S->mtab_valid) {
    rc = timed(c, "gr_metric_table", s, [&] {
      gr_metric_table_kernel<T><<<pblocks, 256, 0, s>>>(k, io);
    });
    if (rc) return rc;
    S->mtab_valid = true;
  }
  constexpr int TY = gr_tile_rows<T>();
  TileMap tm;
  tm.ntx = (k.n[0] + 63) / 64;
  tm.nty = (k.n[1] + TY - 1) / TY;
  {
    // two workgroups per CU (256 CUs), whole columns when the plane has enough tiles
    const int tiles = tm.ntx * tm.nty;
    const int nzc = std::max(1, (512 + tiles - 1) / tiles);
    tm.zc = c->zchunk > 0 ? c->zchunk : std::max(4, (k.n[2] + nzc - 1) / nzc);
    if (tm.zc > k.n[2]) tm.zc = k.n[2];
    tm.nzc = (k.n[2] + tm.zc - 1) / tm.zc;
  }
  const int tblocks = tm.ntx * tm.nty * tm.nzc;
  for (int st = 0; st < 5; st++) {
    T* const* in = (st % 2 == 0) ? A : B;
    T* const* out = (st % 2 == 0) ? B : A;
    for (int q = 0; q < 7; q++) {
      io.in[q] = in[q];
      io.out[q] = out[q];
    }
    const T as = (T)kAs[st], bs = (T)kBs[st];
    const char* name = st == 0 ? "gr_stage_first" : (st == 4 ? "gr_stage_last" : "gr_stage");
    if (tiled) {
      rc = timed(c, name, s, [&] {
        const bool hd3 = c->variant != 3;  // variant 3: halo one plane ahead
        if (st == 0) launch_gr_tiled<T, TY, true, false>(dg, false, hd3, tblocks, s, k, io, as, bs, tm);
        else if (st == 4) launch_gr_tiled<T, TY, false, true>(dg, mt, hd3, tblocks, s, k, io, as, bs, tm);
        else launch_gr_tiled<T, TY, false, false>(dg, mt, hd3, tblocks, s, k, io, as, bs, tm);
      });
    } else {
      rc = timed(c, "gr_point", s, [&] { gr_point_kernel<T><<<pblocks, 256, 0, s>>>(k, io); });
      if (rc) return rc;
      rc = timed(c, name, s, [&] {
        if (st == 0) launch_gr_stage<T, true, false>(dg, grid, block, s, k, io, as, bs);
        else if (st == 4) launch_gr_stage<T, false, true>(dg, grid, block, s, k, io, as, bs);
        else launch_gr_stage<T, false, false>(dg, grid, block, s, k, io, as, bs);
      });
    }
    if (rc) return rc;
    rc = gr_guards<T>(c, s, out, out, true);
    if (rc) return rc;
  }
  for (int q = 0; q < 7; q++) {
    io.in[q] = B[q];
    io.out[q] = A[q];
  }
  rc = timed(c, "gr_ko", s, [&] {
    if (tiled) gr_ko_tiled<T, TY><<<tblocks, 64 * TY, 0, s>>>(k, io, tm);
    else gr_ko_kernel<T><<<grid, block, 0, s>>>(k, io);
  });
  if (rc) return rc;
  return gr_guards<T>(c, s, A, A, true);
}

template <typename T>
int gr_energy_impl(coffee_ctx* c, void* const comp[3], double* out, hipStream_t s) {
  GrState* S = gr_state(c);
  const GrK<T>& k = grk<T>(S);
  int rc = timed(c, "total_energy", s, [&] {
    gr_energy_kernel<T><<<c->part_blocks, 256, 0, s>>>(k, (const T*)comp[0], (const T*)comp[1],
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

template <typename T>
int gr_residuals_impl(coffee_ctx* c, const coffee_fields* f, double out[5], hipStream_t s) {
  GrState* S = gr_state(c);
  const GrK<T>& k = grk<T>(S);
  FlatIO<T> io{};
  for (int q = 0; q < 3; q++) {
    if (!f->E[q] || !f->B[q]) return fail(c, COFFEE_ERR_INVALID, "D/B pointers must be set");
    io.in[q] = (const T*)f->E[q];
    io.in[3 + q] = (const T*)f->B[q];
  }
  if (!f->P) return fail(c, COFFEE_ERR_INVALID, "P pointer must be set");
  io.in[6] = (const T*)f->P;
  int rc = timed(c, "residuals", s, [&] {
    gr_residual_kernel<T><<<c->part_blocks, 256, 0, s>>>(k, io, c->d_part);
  });
  if (rc) return rc;
  HIP_TRY(c, hipMemcpyAsync(c->h_part.data(), c->d_part, sizeof(double) * NRED * c->part_blocks,
                            hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  fold_residuals(c->h_part.data(), c->part_blocks, out);
  return COFFEE_OK;
}

}  // namespace

int gr_residuals(coffee_ctx* c, const coffee_fields* f, double out[5], hipStream_t s) {
  return c->scalar == 4 ? gr_residuals_impl<float>(c, f, out, s)
                        : gr_residuals_impl<double>(c, f, out, s);
}

int gr_create(coffee_ctx* c) {
  GrState* S = new GrState();
  c->state = S;
  fill_gr<float>(c->grid, c->params, S->kf);
  fill_gr<double>(c->grid, c->params, S->kd);
  if (!(c->params.a > -1.0 && c->params.a < 1.0))
    return fail(c, COFFEE_ERR_INVALID, "spin a must satisfy |a| < 1");
  return COFFEE_OK;
}

int gr_step(coffee_ctx* c, const coffee_fields* f, double time, hipStream_t s) {
  (void)time;  // the GR solver has no time-dependent boundary (.cu:905)
  return c->scalar == 4 ? gr_step_impl<float>(c, f, s) : gr_step_impl<double>(c, f, s);
}

int gr_energy(coffee_ctx* c, void* const comp[3], double* out, hipStream_t s) {
  for (int q = 0; q < 3; q++)
    if (!comp[q]) return fail(c, COFFEE_ERR_INVALID, "total_energy needs 3 component pointers");
  return c->scalar == 4 ? gr_energy_impl<float>(c, comp, out, s)
                        : gr_energy_impl<double>(c, comp, out, s);
}

}  // namespace coffee
