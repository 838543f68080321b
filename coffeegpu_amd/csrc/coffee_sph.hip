// Log-spherical 2D solver and batched runs behind the C ABI.
//
// A spherical context is a batch of one run; coffee_sph_batch_* exposes
// batches of many independent runs of different grids and parameters that
// advance together, one launch per substep, optionally replayed from a HIP
// graph, with the tile shape picked by a timing autotuner.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "coffee_hip.h"
#include "ctx.hpp"
#include "ez_sph.hpp"

namespace coffee {

namespace {

const double kAs[5] = {0, -0.4178904745, -1.192151694643, -1.697784692471, -1.514183444257};
const double kBs[5] = {0.1496590219993, 0.3792103129999, 0.8229550293869, 0.6994504559488,
                       0.1530572479681};
const double kCs[5] = {0, 0.1496590219993, 0.3704009573644, 0.6222557631345, 0.9582821306784};

// Tile shapes the stage/KO kernels are built for.  kind 0: TX x TY threads,
// one cell each, stencil taps read global memory through the caches; kind 1:
// the same with the tile and its halo staged in LDS; kind 2: TX threads march
// a band of TX radial cells over TY theta rows (register queue + LDS row).
// Every shape gives bitwise the same results.
struct Tile {
  int tx, ty, kind;
};
template <int TX, int TY, int KIND>
struct TileTag {
  static constexpr int tx = TX, ty = TY, kind = KIND;
};
#define COFFEE_SPH_TILES(X)                                                                      \
  X(0, 64, 4, 0) X(1, 64, 8, 0) X(2, 128, 2, 0) X(3, 32, 8, 0) X(4, 256, 1, 0) X(5, 64, 8, 1)    \
  X(6, 64, 16, 1) X(7, 128, 4, 1) X(8, 32, 16, 1) X(9, 64, 4, 1) X(10, 128, 64, 2)              \
  X(11, 256, 32, 2) X(12, 256, 64, 2) X(13, 256, 128, 2) X(14, 512, 64, 2) X(15, 128, 128, 2) \
  X(16, 512, 128, 2) X(17, 512, 256, 2)
#define COFFEE_SPH_TILE_ENTRY(n, tx, ty, k) {tx, ty, k},
const Tile kTiles[] = {COFFEE_SPH_TILES(COFFEE_SPH_TILE_ENTRY)};
#undef COFFEE_SPH_TILE_ENTRY
const int kNumTiles = sizeof(kTiles) / sizeof(kTiles[0]);

// Calls f(TileTag<tx, ty, kind>{}) for tile index `tile`.
template <typename F>
int with_tile(int tile, F&& f) {
  switch (tile) {
#define COFFEE_SPH_TILE_CASE(n, tx, ty, k) \
  case n:                                  \
    return f(TileTag<tx, ty, k>{});
    COFFEE_SPH_TILES(COFFEE_SPH_TILE_CASE)
#undef COFFEE_SPH_TILE_CASE
  }
  return COFFEE_ERR_INVALID;
}

// The first-stage / middle / last-stage kernels and the KO kernel of a tile.
template <typename T, typename Tag, bool F, bool L, bool D>
void launch_stage(int blocks, hipStream_t s, SphRun<T>* runs, const SphItem* items, int n,
                  int src, T as, T bs, T cs) {
  constexpr int TX = Tag::tx, TY = Tag::ty;
  if constexpr (Tag::kind == 2)
    sph_stage_march_kernel<T, TX, TY, F, L, D><<<blocks, TX, 0, s>>>(runs, items, n, src, as, bs, cs);
  else if constexpr (Tag::kind == 1)
    sph_stage_lds_kernel<T, TX, TY, F, L, D><<<blocks, TX * TY, 0, s>>>(runs, items, n, src, as, bs, cs);
  else
    sph_stage_kernel<T, TX, TY, F, L, D><<<blocks, TX * TY, 0, s>>>(runs, items, n, src, as, bs, cs);
}

template <typename T, typename Tag>
void launch_ko(int blocks, hipStream_t s, SphRun<T>* runs, const SphItem* items, int n, int src) {
  constexpr int TX = Tag::tx, TY = Tag::ty;
  if constexpr (Tag::kind == 2)
    sph_ko_march_kernel<T, TX, TY><<<blocks, TX, 0, s>>>(runs, items, n, src);
  else if constexpr (Tag::kind == 1)
    sph_ko_lds_kernel<T, TX, TY><<<blocks, TX * TY, 0, s>>>(runs, items, n, src);
  else
    sph_ko_kernel<T, TX, TY><<<blocks, TX * TY, 0, s>>>(runs, items, n, src);
}

// Workgroups of the middle-stage kernel of a tile resident per CU.
template <typename T, typename Tag>
int stage_occupancy() {
  constexpr int TX = Tag::tx, TY = Tag::ty;
  const void* k;
  if constexpr (Tag::kind == 2)
    k = reinterpret_cast<const void*>(sph_stage_march_kernel<T, TX, TY, false, false, false>);
  else if constexpr (Tag::kind == 1)
    k = reinterpret_cast<const void*>(sph_stage_lds_kernel<T, TX, TY, false, false, false>);
  else
    k = reinterpret_cast<const void*>(sph_stage_kernel<T, TX, TY, false, false, false>);
  const int threads = Tag::kind == 2 ? TX : TX * TY;
  int nb = 0;
  const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k, threads, 0);
  return e == hipSuccess ? nb : 2048 / threads;
}

struct SphStateBase : SolverState {
  virtual int step(coffee_ctx* c, hipStream_t s) = 0;
};

template <typename T>
struct SphState : SphStateBase {
  std::vector<SphRun<T>> h_runs;
  SphRun<T>* d_runs = nullptr;
  SphItem* d_items = nullptr;
  int nitems = 0;
  int tile = 0;
  int grid_blocks = 0;
  int max_n0 = 0;
  std::vector<void*> owned;
  bool use_graph = false;
  hipGraphExec_t graph_exec = nullptr;
  hipGraph_t graph = nullptr;
  int graph_tile = -1;

  ~SphState() override {
    drop_graph();
    for (void* p : owned) (void)hipFree(p);
    if (d_runs) (void)hipFree(d_runs);
    if (d_items) (void)hipFree(d_items);
  }

  // A captured graph holds the work-list pointer and the launch geometry of
  // the tile it was captured with: any rebuild of the work list drops it.
  void drop_graph() {
    if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
    if (graph) (void)hipGraphDestroy(graph);
    graph_exec = nullptr;
    graph = nullptr;
    graph_tile = -1;
  }

  int build_items(coffee_ctx* c) {
    drop_graph();
    const Tile t = kTiles[tile];
    std::vector<SphItem> items;
    for (int r = 0; r < (int)h_runs.size(); r++)
      for (int j0 = 0; j0 < h_runs[r].n1; j0 += t.ty)
        for (int i0 = 0; i0 < h_runs[r].n0; i0 += t.tx) items.push_back({r, i0, j0});
    if (d_items) (void)hipFree(d_items);
    d_items = nullptr;
    nitems = (int)items.size();
    HIP_TRY(c, hipMalloc(&d_items, sizeof(SphItem) * std::max(1, nitems)));
    HIP_TRY(c, hipMemcpy(d_items, items.data(), sizeof(SphItem) * nitems, hipMemcpyHostToDevice));
    int dev = 0, cus = 256;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    // persistent launch: exactly as many workgroups as can be resident at once
    // (a larger grid would leave the extra workgroups to start only when
    // resident ones have finished their whole share of the work list)
    const int per_cu = std::max(1, resident_per_cu());
    grid_blocks = std::max(1, std::min(nitems, cus * per_cu));
    return COFFEE_OK;
  }

  int resident_per_cu() const {
    return with_tile(tile, [](auto tag) { return stage_occupancy<T, decltype(tag)>(); });
  }

  template <typename Tag>
  int launch_step(coffee_ctx* c, hipStream_t s) {
    const int nr = (int)h_runs.size();
    const dim3 gb((max_n0 + 255) / 256, nr);
    int rc = timed(c, "sph_axis", s, [&] { sph_axis_kernel<T><<<gb, 256, 0, s>>>(d_runs, nr, 0); });
    if (rc) return rc;
    const dim3 gg(std::max(1, std::min(1024, (int)((long)max_n0 * 64 / 256))), nr);
    rc = timed(c, "sph_guard_copy", s,
               [&] { sph_guard_copy_kernel<T><<<gg, 256, 0, s>>>(d_runs, nr); });
    if (rc) return rc;
    const bool dg = c->diagnostics != 0;
    for (int st = 0; st < 5; st++) {
      const int src = st & 1;
      const T as = (T)kAs[st], bs = (T)kBs[st], cs = (T)kCs[st];
      const char* name = st == 0 ? "sph_stage_first" : (st == 4 ? "sph_stage_last" : "sph_stage");
      rc = timed(c, name, s, [&] {
        const int g = grid_blocks, n = nitems;
        if (st == 0) {
          if (dg) launch_stage<T, Tag, true, false, true>(g, s, d_runs, d_items, n, src, as, bs, cs);
          else launch_stage<T, Tag, true, false, false>(g, s, d_runs, d_items, n, src, as, bs, cs);
        } else if (st == 4) {
          if (dg) launch_stage<T, Tag, false, true, true>(g, s, d_runs, d_items, n, src, as, bs, cs);
          else launch_stage<T, Tag, false, true, false>(g, s, d_runs, d_items, n, src, as, bs, cs);
        } else {
          if (dg) launch_stage<T, Tag, false, false, true>(g, s, d_runs, d_items, n, src, as, bs, cs);
          else launch_stage<T, Tag, false, false, false>(g, s, d_runs, d_items, n, src, as, bs, cs);
        }
      });
      if (rc) return rc;
      rc = timed(c, "sph_axis", s,
                 [&] { sph_axis_kernel<T><<<gb, 256, 0, s>>>(d_runs, nr, src ^ 1); });
      if (rc) return rc;
    }
    rc = timed(c, "sph_ko", s,
               [&] { launch_ko<T, Tag>(grid_blocks, s, d_runs, d_items, nitems, 1); });
    if (rc) return rc;
    rc = timed(c, "sph_axis", s, [&] { sph_axis_kernel<T><<<gb, 256, 0, s>>>(d_runs, nr, 0); });
    if (rc) return rc;
    return timed(c, "sph_advance", s, [&] {
      sph_advance_kernel<T><<<(nr + 63) / 64, 64, 0, s>>>(d_runs, nr);
    });
  }

  int step_direct(coffee_ctx* c, hipStream_t s) {
    return with_tile(tile, [&](auto tag) { return launch_step<decltype(tag)>(c, s); });
  }

  int step(coffee_ctx* c, hipStream_t s) override {
    if (!use_graph || c->timing || c->diagnostics) return step_direct(c, s);
    if (!graph_exec || graph_tile != tile) {
      drop_graph();
      hipStream_t cs;
      HIP_TRY(c, hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
      HIP_TRY(c, hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      int rc = step_direct(c, cs);
      hipGraph_t g = nullptr;
      hipError_t e = hipStreamEndCapture(cs, &g);
      (void)hipStreamDestroy(cs);
      if (rc) return rc;
      if (e != hipSuccess) return fail(c, COFFEE_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
      graph = g;
      HIP_TRY(c, hipGraphInstantiate(&graph_exec, graph, nullptr, nullptr, 0));
      graph_tile = tile;
    }
    HIP_TRY(c, hipGraphLaunch(graph_exec, s));
    return COFFEE_OK;
  }

  int upload(coffee_ctx* c) {
    HIP_TRY(c, hipMemcpy(d_runs, h_runs.data(), sizeof(SphRun<T>) * h_runs.size(),
                         hipMemcpyHostToDevice));
    return COFFEE_OK;
  }
};

template <typename T>
int alloc_owned(coffee_ctx* c, SphState<T>* S, size_t n, T** out) {
  void* p = nullptr;
  HIP_TRY(c, hipMalloc(&p, n * sizeof(T)));
  HIP_TRY(c, hipMemset(p, 0, n * sizeof(T)));
  S->owned.push_back(p);
  *out = (T*)p;
  return COFFEE_OK;
}

int validate_sph(coffee_ctx* c, const coffee_grid& g) {
  if (g.dims[2] != 1 || g.guard[2] != 0)
    return fail(c, COFFEE_ERR_INVALID, "spherical grids are 2D: dims[2] = 1, guard[2] = 0");
  if (g.guard[0] < 3) return fail(c, COFFEE_ERR_INVALID, "radial guard must be >= 3 (KO6)");
  // the theta = pi axis row is the first upper guard row and the mirror
  // writes three rows beyond it
  if (g.guard[1] < 4) return fail(c, COFFEE_ERR_INVALID, "theta guard must be >= 4 (axis mirror)");
  if (g.dims[0] - 2 * g.guard[0] < 1 || g.dims[1] - 2 * g.guard[1] < 4)
    return fail(c, COFFEE_ERR_INVALID, "grid too small");
  // kernels address cells by 32-bit byte offsets
  if ((double)g.dims[0] * g.dims[1] * c->scalar >= 4294967296.0)
    return fail(c, COFFEE_ERR_SIZE, "grid exceeds 4 GiB per array");
  return COFFEE_OK;
}

// Fill one run descriptor (parameters in Scalar, metric tables, buffers).
template <typename T>
int make_run(coffee_ctx* c, SphState<T>* S, const coffee_grid& g, const coffee_params& p,
             const coffee_fields* f, int nsteps, double t0, SphRun<T>& R) {
  int rc = validate_sph(c, g);
  if (rc) return rc;
  if (!(p.dt > 0) || p.tau == 0) return fail(c, COFFEE_ERR_INVALID, "dt must be > 0 and tau != 0");
  if (p.shift_ghost != 0)
    return fail(c, COFFEE_ERR_INVALID, "shift_ghost != 0 is not supported (guards stay frozen)");
  std::memset(&R, 0, sizeof(R));
  R.dims0 = g.dims[0];
  R.dims1 = g.dims[1];
  R.g0 = g.guard[0];
  R.g1 = g.guard[1];
  R.n0 = g.dims[0] - 2 * g.guard[0];
  R.n1 = g.dims[1] - 2 * g.guard[1];
  R.dx = (T)g.delta[0];
  R.dy = (T)g.delta[1];
  R.x0 = (T)g.lower[0];
  R.y0 = (T)g.lower[1];
  R.inv_dx = (T)(1.0 / (double)R.dx);
  R.inv_dy = (T)(1.0 / (double)R.dy);
  R.idx12 = R.inv_dx / T(12);
  R.idy12 = R.inv_dy / T(12);
  R.dt = (T)p.dt;
  R.ch2 = (T)p.ch2;
  R.tau = (T)p.tau;
  R.KOfac = -(T)p.KOeps * (T)(-1.0 / 64.0);
  R.damp_gamma = (T)p.damp_gamma;
  R.radius = (T)p.radius;
  R.b0 = (T)p.b0;
  R.omega = (T)p.omega;
  R.sigpml = (T)p.sigpml;
  R.pmllen = (T)(p.pmllen * 1.0);
  R.xh = (T)p.lower[0] + (T)p.size[0] - (T)p.pml[0] * R.dx;
  R.rh = (T)std::exp((double)R.xh);
  R.tp_start = (T)p.tp_start;
  R.tp_end = (T)p.tp_end;
  R.dw0 = (T)p.dw0;
  R.nT = (T)p.nT;
  R.thp1 = (T)p.thp1;
  R.thp2 = (T)p.thp2;
  R.tp_start1 = (T)p.tp_start1;
  R.tp_end1 = (T)p.tp_end1;
  R.dw1 = (T)p.dw1;
  R.nT1 = (T)p.nT1;
  R.thp11 = (T)p.thp11;
  R.thp21 = (T)p.thp21;
  R.calc_current = p.calc_current;
  R.use_edotb = p.use_edotb_damping;
  R.divB_clean = p.divB_clean;
  R.clean_ep = p.clean_ep;
  R.check_egb = p.check_egb;
  R.KO_geometry = p.KO_geometry;
  R.pml_on = p.pml[0] > 0;
  // metric tables r(i) = exp(x_i), sin(theta_j) at the node positions
  std::vector<T> r(g.dims[0]), sth(g.dims[1]), r3(g.dims[0]), ir2(g.dims[0]);
  for (int i = 0; i < g.dims[0]; i++) {
    const double rd = std::exp((double)(R.x0 + R.dx * T(i - R.g0)));
    r[i] = (T)rd;
    r3[i] = (T)(rd * rd * rd);
    ir2[i] = (T)(1.0 / (rd * rd));
  }
  for (int j = 0; j < g.dims[1]; j++) sth[j] = (T)std::sin((double)(R.y0 + R.dy * T(j - R.g1)));
  T *dr = nullptr, *ds = nullptr, *dr3 = nullptr, *dir2 = nullptr;
  if ((rc = alloc_owned(c, S, r.size(), &dr))) return rc;
  if ((rc = alloc_owned(c, S, sth.size(), &ds))) return rc;
  if ((rc = alloc_owned(c, S, r3.size(), &dr3))) return rc;
  if ((rc = alloc_owned(c, S, ir2.size(), &dir2))) return rc;
  HIP_TRY(c, hipMemcpy(dr, r.data(), r.size() * sizeof(T), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(ds, sth.data(), sth.size() * sizeof(T), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(dr3, r3.data(), r3.size() * sizeof(T), hipMemcpyHostToDevice));
  HIP_TRY(c, hipMemcpy(dir2, ir2.data(), ir2.size() * sizeof(T), hipMemcpyHostToDevice));
  R.r = dr;
  R.sth = ds;
  R.r3 = dr3;
  R.ir2 = dir2;
  R.inv_tau = (T)(1.0 / p.tau);
  const size_t n = (size_t)g.dims[0] * g.dims[1];
  for (int q = 0; q < 7; q++)
    if ((rc = alloc_owned(c, S, n, &R.U[1][q]))) return rc;
  if (f) {
    for (int q = 0; q < 3; q++) {
      R.U[0][q] = (T*)f->E[q];
      R.U[0][3 + q] = (T*)f->B[q];
      R.dU[q] = (T*)f->dE[q];
      R.dU[3 + q] = (T*)f->dB[q];
      R.Un[q] = (T*)f->Etmp[q];
      R.Un[3 + q] = (T*)f->Btmp[q];
      R.J[q] = (T*)f->J[q];
    }
    R.U[0][6] = (T*)f->P;
    R.dU[6] = (T*)f->dP;
    R.Un[6] = (T*)f->Ptmp;
    R.divE = (T*)f->divE;
    R.divB = (T*)f->divB;
  }
  // A batch run brings its state arrays now; a single-run context gets all
  // caller arrays with each step instead.
  if (f)
    for (int q = 0; q < 7; q++) {
      if (!R.U[0][q]) return fail(c, COFFEE_ERR_INVALID, "E, B, P must be set for every run");
      if (!R.dU[q] && (rc = alloc_owned(c, S, n, &R.dU[q]))) return rc;
      if (R.pml_on && !R.Un[q] && (rc = alloc_owned(c, S, n, &R.Un[q]))) return rc;
    }
  R.t = t0;
  R.step = 0;
  R.nsteps = nsteps;
  R.active = 1;
  return COFFEE_OK;
}

template <typename T>
int batch_create(coffee_ctx* c, int nruns, const coffee_grid* grids, const coffee_params* params,
                 const coffee_fields* fields, const int* nsteps, const double* t0) {
  auto* S = new SphState<T>();
  c->state = S;
  S->h_runs.resize(nruns);
  for (int r = 0; r < nruns; r++) {
    int rc = make_run<T>(c, S, grids[r], params[r], fields ? &fields[r] : nullptr,
                         nsteps ? nsteps[r] : 0, t0 ? t0[r] : 0.0, S->h_runs[r]);
    if (rc) return rc;
    S->max_n0 = std::max(S->max_n0, S->h_runs[r].n0);
  }
  HIP_TRY(c, hipMalloc(&S->d_runs, sizeof(SphRun<T>) * nruns));
  int rc = S->upload(c);
  if (rc) return rc;
  return S->build_items(c);
}

template <typename T>
SphState<T>* sph_state(coffee_ctx* c) {
  return static_cast<SphState<T>*>(c->state);
}

template <typename T>
int luminosity_impl(coffee_ctx* c, const double* radius, double* out, hipStream_t s) {
  SphState<T>* S = sph_state<T>(c);
  const int nr = (int)S->h_runs.size();
  double* d = nullptr;
  HIP_TRY(c, hipMalloc(&d, sizeof(double) * 2 * nr));
  hipError_t e = hipMemcpyAsync(d, radius, sizeof(double) * nr, hipMemcpyHostToDevice, s);
  if (e == hipSuccess) {
    sph_luminosity_kernel<T><<<nr, 256, 0, s>>>(S->d_runs, d, d + nr);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipMemcpyAsync(out, d + nr, sizeof(double) * nr, hipMemcpyDeviceToHost, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(c, COFFEE_ERR_HIP, std::string("luminosity: ") + hipGetErrorString(e));
  return COFFEE_OK;
}

}  // namespace

int sph_create(coffee_ctx* c) {
  if (c->scalar == 4) return batch_create<float>(c, 1, &c->grid, &c->params, nullptr, nullptr, nullptr);
  return batch_create<double>(c, 1, &c->grid, &c->params, nullptr, nullptr, nullptr);
}

// Single run: the caller's arrays and time of this call.
template <typename T>
int sph_step_impl(coffee_ctx* c, const coffee_fields* f, double time, hipStream_t s) {
  SphState<T>* S = sph_state<T>(c);
  SphRun<T>& R = S->h_runs[0];
  for (int q = 0; q < 3; q++) {
    if (!f->E[q] || !f->B[q] || !f->dE[q] || !f->dB[q])
      return fail(c, COFFEE_ERR_INVALID, "E, B, dE, dB must be set");
    R.U[0][q] = (T*)f->E[q];
    R.U[0][3 + q] = (T*)f->B[q];
    R.dU[q] = (T*)f->dE[q];
    R.dU[3 + q] = (T*)f->dB[q];
    R.Un[q] = (T*)f->Etmp[q];
    R.Un[3 + q] = (T*)f->Btmp[q];
    R.J[q] = (T*)f->J[q];
  }
  if (!f->P || !f->dP) return fail(c, COFFEE_ERR_INVALID, "P, dP must be set");
  R.U[0][6] = (T*)f->P;
  R.dU[6] = (T*)f->dP;
  R.Un[6] = (T*)f->Ptmp;
  R.divE = (T*)f->divE;
  R.divB = (T*)f->divB;
  if (R.pml_on)
    for (int q = 0; q < 7; q++)
      if (!R.Un[q]) return fail(c, COFFEE_ERR_INVALID, "absorbing layer needs Etmp, Btmp, Ptmp");
  if (c->diagnostics && (!R.J[0] || !R.J[1] || !R.J[2] || !R.divE || !R.divB))
    return fail(c, COFFEE_ERR_INVALID, "diagnostics need J, divE, divB");
  R.t = time;
  R.step = 0;
  R.nsteps = 0;
  R.active = 1;
  HIP_TRY(c, hipMemcpyAsync(S->d_runs, &R, sizeof(SphRun<T>), hipMemcpyHostToDevice, s));
  // the host copy must outlive the asynchronous copy
  HIP_TRY(c, hipStreamSynchronize(s));
  return S->step_direct(c, s);
}

int sph_step(coffee_ctx* c, const coffee_fields* f, double time, hipStream_t s) {
  return c->scalar == 4 ? sph_step_impl<float>(c, f, time, s) : sph_step_impl<double>(c, f, time, s);
}

template <typename T>
__global__ void __launch_bounds__(256)
    sph_energy_kernel(const SphRun<T>* __restrict__ runs, const T* __restrict__ f0,
                      const T* __restrict__ f1, const T* __restrict__ f2, T xl, T xh, T yl, T yh,
                      double* __restrict__ part) {
  // field_solver_EZ_spherical::total_energy (.cu:1014-1052)
  __shared__ double sh[256];
  const SphRun<T>& R = runs[0];
  double acc = 0;
  const long total = (long)R.n0 * R.n1;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
       t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t % R.n0) + R.g0, j = (int)(t / R.n0) + R.g1;
    const int c = i + j * R.dims0;
    const T x = R.x0 + R.dx * T(i - R.g0), y = R.y0 + R.dy * T(j - R.g1);
    const T r = R.r[i], sth = R.sth[j];
    if (r >= R.radius && x < xh && x > xl && y < yh && y > yl) {
      const T g1 = r * r, g3 = (r * sth) * (r * sth);
      acc += (double)((g1 * f0[c] * f0[c] + g1 * f1[c] * f1[c] + g3 * f2[c] * f2[c]) *
                      sph_sqrtg(r, sth) * R.dx * R.dy);
    }
  }
  sh[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

template <typename T>
int sph_energy_impl(coffee_ctx* c, void* const comp[3], double* out, hipStream_t s) {
  SphState<T>* S = sph_state<T>(c);
  const SphRun<T>& R = S->h_runs[0];
  const coffee_params& p = c->params;
  const T xh = (T)p.lower[0] + (T)p.size[0] - (T)p.pml[0] * R.dx;
  const T xl = (T)p.lower[0];
  const T yh = (T)p.lower[1] + (T)p.size[1] - (T)p.pml[1] * R.dy;
  const T yl = (T)p.lower[1] + (T)p.pml[1] * R.dy;
  int rc = timed(c, "sph_energy", s, [&] {
    sph_energy_kernel<T><<<c->part_blocks, 256, 0, s>>>(S->d_runs, (const T*)comp[0],
                                                        (const T*)comp[1], (const T*)comp[2], xl,
                                                        xh, yl, yh, c->d_part);
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

int sph_energy(coffee_ctx* c, void* const comp[3], double* out, hipStream_t s) {
  return c->scalar == 4 ? sph_energy_impl<float>(c, comp, out, s)
                        : sph_energy_impl<double>(c, comp, out, s);
}

// Residual rows of every run: c->part_blocks workgroups, part_blocks / nruns
// per run, folded per run into out[run * 5 ...].
template <typename T>
int sph_residuals_runs(coffee_ctx* c, double* out, hipStream_t s) {
  SphState<T>* S = sph_state<T>(c);
  const int nr = (int)S->h_runs.size();
  const int bpr = c->part_blocks / nr;
  if (bpr < 1 || !c->d_part) return fail(c, COFFEE_ERR_STATE, "no residual buffer");
  int rc = timed(c, "sph_residuals", s, [&] {
    sph_residual_kernel<T><<<bpr * nr, 256, 0, s>>>(S->d_runs, bpr, c->d_part);
  });
  if (rc) return rc;
  HIP_TRY(c, hipMemcpyAsync(c->h_part.data(), c->d_part, sizeof(double) * NRED * bpr * nr,
                            hipMemcpyDeviceToHost, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  for (int r = 0; r < nr; r++) fold_residuals(c->h_part.data() + (size_t)r * bpr * NRED, bpr, out + 5 * r);
  return COFFEE_OK;
}

// Single run: the state arrays of this call.
template <typename T>
int sph_residuals_impl(coffee_ctx* c, const coffee_fields* f, double out[5], hipStream_t s) {
  SphState<T>* S = sph_state<T>(c);
  SphRun<T>& R = S->h_runs[0];
  for (int q = 0; q < 3; q++) {
    if (!f->E[q] || !f->B[q]) return fail(c, COFFEE_ERR_INVALID, "E, B must be set");
    R.U[0][q] = (T*)f->E[q];
    R.U[0][3 + q] = (T*)f->B[q];
  }
  if (!f->P) return fail(c, COFFEE_ERR_INVALID, "P must be set");
  R.U[0][6] = (T*)f->P;
  HIP_TRY(c, hipMemcpyAsync(S->d_runs, &R, sizeof(SphRun<T>), hipMemcpyHostToDevice, s));
  HIP_TRY(c, hipStreamSynchronize(s));
  return sph_residuals_runs<T>(c, out, s);
}

int sph_residuals(coffee_ctx* c, const coffee_fields* f, double out[5], hipStream_t s) {
  return c->scalar == 4 ? sph_residuals_impl<float>(c, f, out, s)
                        : sph_residuals_impl<double>(c, f, out, s);
}

}  // namespace coffee

using namespace coffee;

// Timing autotuner: times the first-stage kernel (which reads the state
// and writes only buffers the next step overwrites anyway) for every tile
// shape; the shapes within 15 % of the fastest are timed twice more,
// interleaved, and keep their minimum; shapes within 3 % of the fastest
// minimum count as ties and the lowest index wins, so run-to-run noise of a
// few per cent does not change the choice.
template <typename T>
int autotune_impl(coffee_ctx* c, int trials, hipStream_t s, int* best, double* ms_out) {
  auto* S = static_cast<SphState<T>*>(c->state);
  hipEvent_t a, b;
  HIP_TRY(c, hipEventCreate(&a));
  HIP_TRY(c, hipEventCreate(&b));
  std::vector<double> ms(kNumTiles, 0.0);
  auto time_tile = [&](int t, double* out) -> int {
    S->tile = t;
    int rc = S->build_items(c);
    if (rc) return rc;
    auto launch = [&] {
      with_tile(t, [&](auto tag) {
        launch_stage<T, decltype(tag), true, false, false>(S->grid_blocks, s, S->d_runs,
                                                           S->d_items, S->nitems, 0, T(0),
                                                           (T)kBs[0], T(0));
        return 0;
      });
    };
    launch();
    HIP_TRY(c, hipEventRecord(a, s));
    for (int k = 0; k < trials; k++) launch();
    HIP_TRY(c, hipEventRecord(b, s));
    HIP_TRY(c, hipEventSynchronize(b));
    float e = 0;
    HIP_TRY(c, hipEventElapsedTime(&e, a, b));
    *out = e / std::max(1, trials);
    return COFFEE_OK;
  };
  for (int t = 0; t < kNumTiles; t++) {
    int rc = time_tile(t, &ms[t]);
    if (rc) return rc;
  }
  double fastest = *std::min_element(ms.begin(), ms.end());
  std::vector<int> near;
  for (int t = 0; t < kNumTiles; t++)
    if (ms[t] <= fastest * 1.15) near.push_back(t);
  for (int round = 0; round < 2 && near.size() > 1; round++)
    for (int t : near) {
      double m = 0;
      int rc = time_tile(t, &m);
      if (rc) return rc;
      ms[t] = std::min(ms[t], m);
    }
  (void)hipEventDestroy(a);
  (void)hipEventDestroy(b);
  int pick = 0;
  fastest = *std::min_element(ms.begin(), ms.end());
  for (int t = 0; t < kNumTiles; t++)
    if (ms[t] <= fastest * 1.03) {
      pick = t;
      break;
    }
  S->tile = pick;
  if (best) *best = pick;
  if (ms_out)
    for (int t = 0; t < kNumTiles; t++) ms_out[t] = ms[t];
  return S->build_items(c);
}

// Every coffee_sph_batch_* entry takes only contexts made by
// coffee_sph_batch_create: a single-run spherical context has no run arrays
// until its first step, and other solvers hold other state.
static bool is_batch(const coffee_ctx* c) {
  return c && c->state && c->solver == COFFEE_SOLVER_EZ_SPHERICAL && c->is_batch;
}

extern "C" {

int coffee_sph_batch_create(int scalar_bytes, int nruns, const coffee_grid* grids,
                            const coffee_params* params, const coffee_fields* fields,
                            const int* nsteps, const double* t0, coffee_batch** out) {
  if (!out || !grids || !params || nruns < 1) return fail(nullptr, COFFEE_ERR_INVALID, "bad argument");
  *out = nullptr;
  if (scalar_bytes != 4 && scalar_bytes != 8)
    return fail(nullptr, COFFEE_ERR_INVALID, "scalar_bytes must be 4 or 8");
  if (!fields) return fail(nullptr, COFFEE_ERR_INVALID, "fields (E, B, P per run) must be given");
  coffee_ctx* c = new coffee_ctx();
  c->solver = COFFEE_SOLVER_EZ_SPHERICAL;
  c->is_batch = 1;
  c->scalar = scalar_bytes;
  c->grid = grids[0];
  c->params = params[0];
  (void)hipGetDevice(&c->device);
  int rc = scalar_bytes == 4 ? batch_create<float>(c, nruns, grids, params, fields, nsteps, t0)
                             : batch_create<double>(c, nruns, grids, params, fields, nsteps, t0);
  if (rc == COFFEE_OK) {
    // residual partial rows: about 1024 workgroups over the batch, at least
    // one per run
    c->part_blocks = std::max(1, 1024 / nruns) * nruns;
    c->h_part.assign((size_t)NRED * c->part_blocks, 0.0);
    if (hipMalloc(&c->d_part, sizeof(double) * NRED * c->part_blocks) != hipSuccess)
      rc = fail(c, COFFEE_ERR_HIP, "hipMalloc residual buffer failed");
  }
  if (rc) {
    std::string m = c->err;
    coffee_ctx_destroy(c);
    return fail(nullptr, rc, m);
  }
  *out = c;
  return COFFEE_OK;
}

int coffee_sph_batch_destroy(coffee_batch* b) {
  if (b && !is_batch(b)) return fail(b, COFFEE_ERR_STATE, "not a spherical batch");
  return coffee_ctx_destroy(b);
}

int coffee_sph_batch_step(coffee_batch* c, int nsteps, void* stream) {
  if (!is_batch(c)) return fail(c, COFFEE_ERR_STATE, "not a spherical batch");
  auto* S = static_cast<SphStateBase*>(c->state);
  for (int k = 0; k < nsteps; k++) {
    int rc = S->step(c, (hipStream_t)stream);
    if (rc) return rc;
  }
  return COFFEE_OK;
}

int coffee_sph_batch_set_tile(coffee_batch* c, int tile_index) {
  if (!is_batch(c)) return fail(c, COFFEE_ERR_STATE, "not a spherical batch");
  if (tile_index < 0 || tile_index >= kNumTiles) return fail(c, COFFEE_ERR_INVALID, "bad tile index");
  if (c->scalar == 4) {
    auto* S = static_cast<SphState<float>*>(c->state);
    S->tile = tile_index;
    return S->build_items(c);
  }
  auto* S = static_cast<SphState<double>*>(c->state);
  S->tile = tile_index;
  return S->build_items(c);
}

int coffee_sph_tile_shape(int tile_index, int* tx, int* ty) {
  if (tile_index < 0 || tile_index >= kNumTiles || !tx || !ty)
    return fail(nullptr, COFFEE_ERR_INVALID, "bad tile index");
  *tx = kTiles[tile_index].tx;
  *ty = kTiles[tile_index].ty;
  return COFFEE_OK;
}

int coffee_sph_num_tiles(void) { return kNumTiles; }

int coffee_sph_batch_autotune(coffee_batch* c, int trials, void* stream, int* best_tile,
                              double* ms_per_tile) {
  if (!is_batch(c)) return fail(c, COFFEE_ERR_STATE, "not a spherical batch");
  hipStream_t s = (hipStream_t)stream;
  return c->scalar == 4 ? autotune_impl<float>(c, trials, s, best_tile, ms_per_tile)
                        : autotune_impl<double>(c, trials, s, best_tile, ms_per_tile);
}

int coffee_sph_batch_use_graph(coffee_batch* c, int enable) {
  if (!is_batch(c)) return fail(c, COFFEE_ERR_STATE, "not a spherical batch");
  if (c->scalar == 4) static_cast<SphState<float>*>(c->state)->use_graph = enable != 0;
  else static_cast<SphState<double>*>(c->state)->use_graph = enable != 0;
  return COFFEE_OK;
}

int coffee_sph_batch_luminosity(coffee_batch* c, const double* radius, double* out, void* stream) {
  if (!is_batch(c)) return fail(c, COFFEE_ERR_STATE, "not a spherical batch");
  if (!radius || !out) return fail(c, COFFEE_ERR_INVALID, "null argument");
  hipStream_t s = (hipStream_t)stream;
  return c->scalar == 4 ? luminosity_impl<float>(c, radius, out, s)
                        : luminosity_impl<double>(c, radius, out, s);
}

int coffee_sph_batch_residuals(coffee_batch* c, double* out, void* stream) {
  if (!is_batch(c)) return fail(c, COFFEE_ERR_STATE, "not a spherical batch");
  if (!out) return fail(c, COFFEE_ERR_INVALID, "null argument");
  hipStream_t s = (hipStream_t)stream;
  return c->scalar == 4 ? sph_residuals_runs<float>(c, out, s) : sph_residuals_runs<double>(c, out, s);
}

int coffee_sph_batch_status(coffee_batch* c, int* active, double* t, int* step) {
  if (!is_batch(c)) return fail(c, COFFEE_ERR_STATE, "not a spherical batch");
  HIP_TRY(c, hipDeviceSynchronize());
  if (c->scalar == 4) {
    auto* S = static_cast<SphState<float>*>(c->state);
    std::vector<SphRun<float>> h(S->h_runs.size());
    HIP_TRY(c, hipMemcpy(h.data(), S->d_runs, sizeof(SphRun<float>) * h.size(), hipMemcpyDeviceToHost));
    for (size_t r = 0; r < h.size(); r++) {
      if (active) active[r] = h[r].active;
      if (t) t[r] = h[r].t;
      if (step) step[r] = h[r].step;
    }
  } else {
    auto* S = static_cast<SphState<double>*>(c->state);
    std::vector<SphRun<double>> h(S->h_runs.size());
    HIP_TRY(c, hipMemcpy(h.data(), S->d_runs, sizeof(SphRun<double>) * h.size(), hipMemcpyDeviceToHost));
    for (size_t r = 0; r < h.size(); r++) {
      if (active) active[r] = h[r].active;
      if (t) t[r] = h[r].t;
      if (step) step[r] = h[r].step;
    }
  }
  return COFFEE_OK;
}

}  // extern "C"
