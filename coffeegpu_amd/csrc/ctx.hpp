// Context of the C ABI and the launch helpers shared by the solver
// translation units (coffee_hip.hip: flat EZ and common entry points;
// coffee_sph.hip: log-spherical and batched runs; coffee_gr.hip: Kerr-Schild).
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "coffee_hip.h"
#include "ez_flat.hpp"

namespace coffee {

struct TimingRec {
  int name;
  hipEvent_t a, b;
};

// Solver-specific state held by a context (spherical batch, GR metric, ...).
struct SolverState {
  virtual ~SolverState() {}
};

std::string& create_error();

}  // namespace coffee

struct coffee_ctx {
  int solver = 0;
  int is_batch = 0;  // made by coffee_sph_batch_create
  int scalar = 8;
  int device = 0;
  coffee_grid grid{};
  coffee_params params{};
  coffee::FlatK<float> kf{};
  coffee::FlatK<double> kd{};
  // ping-pong partner buffers of the 7 state arrays; flat solver: [7] is
  // the third buffer of P (its five stages rotate over three buffers, so the
  // last stage writes P into the caller's array and Kreiss-Oliger, which
  // leaves P unchanged, does not copy it)
  void* scratch[8] = {};
  double* d_part = nullptr;
  int part_blocks = 0;
  std::vector<double> h_part;
  int diagnostics = 0;
  int timing = 0;
  int variant = 0;  // kernel schedule
  int zchunk = 0;   // planes per workgroup for the tiled kernels (0 = automatic)
  std::vector<std::string> tnames;
  std::map<std::string, int> tindex;
  std::vector<coffee::TimingRec> trecs;
  std::vector<hipEvent_t> event_pool;
  coffee::SolverState* state = nullptr;
  std::string err;
};

namespace coffee {

inline int fail(coffee_ctx* c, int code, const std::string& msg) {
  if (c)
    c->err = msg;
  else
    create_error() = msg;
  return code;
}

#define HIP_TRY(ctx, expr)                                                                  \
  do {                                                                                      \
    hipError_t e_ = (expr);                                                                 \
    if (e_ != hipSuccess)                                                                   \
      return ::coffee::fail(ctx, COFFEE_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

inline hipEvent_t take_event(coffee_ctx* c) {
  if (!c->event_pool.empty()) {
    hipEvent_t e = c->event_pool.back();
    c->event_pool.pop_back();
    return e;
  }
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

inline int tname(coffee_ctx* c, const char* n) {
  auto it = c->tindex.find(n);
  if (it != c->tindex.end()) return it->second;
  int id = (int)c->tnames.size();
  c->tnames.push_back(n);
  c->tindex[n] = id;
  return id;
}

// Launch helper: brackets the launch with events when timing is on.
template <typename F>
int timed(coffee_ctx* c, const char* name, hipStream_t s, F&& launch) {
  hipEvent_t a = nullptr, b = nullptr;
  if (c->timing) {
    a = take_event(c);
    b = take_event(c);
    if (!a || !b) return fail(c, COFFEE_ERR_HIP, "hipEventCreate failed");
    HIP_TRY(c, hipEventRecord(a, s));
  }
  launch();
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(c, COFFEE_ERR_HIP, std::string(name) + ": " + hipGetErrorString(e));
  if (c->timing) {
    HIP_TRY(c, hipEventRecord(b, s));
    c->trecs.push_back({tname(c, name), a, b});
  }
  return COFFEE_OK;
}

template <typename T>
inline FlatK<T>& flat_k(coffee_ctx* c);
template <>
inline FlatK<float>& flat_k<float>(coffee_ctx* c) { return c->kf; }
template <>
inline FlatK<double>& flat_k<double>(coffee_ctx* c) { return c->kd; }

// Entry points of the other translation units, dispatched from the ABI.
int gr_create(coffee_ctx* c);
int gr_step(coffee_ctx* c, const coffee_fields* f, double time, hipStream_t s);
int gr_energy(coffee_ctx* c, void* const comp[3], double* out, hipStream_t s);
int sph_create(coffee_ctx* c);
int sph_step(coffee_ctx* c, const coffee_fields* f, double time, hipStream_t s);
int sph_energy(coffee_ctx* c, void* const comp[3], double* out, hipStream_t s);
int sph_residuals(coffee_ctx* c, const coffee_fields* f, double out[5], hipStream_t s);
int gr_residuals(coffee_ctx* c, const coffee_fields* f, double out[5], hipStream_t s);

// Folds nblocks partial rows of NRED doubles (sum, max, max, max, sum) in
// workgroup order.
inline void fold_residuals(const double* part, int nblocks, double out[5]) {
  double r[5] = {0, 0, 0, 0, 0};
  for (int b = 0; b < nblocks; b++) {
    const double* p = part + b * 5;
    r[0] += p[0];
    r[1] = std::max(r[1], p[1]);
    r[2] = std::max(r[2], p[2]);
    r[3] = std::max(r[3], p[3]);
    r[4] += p[4];
  }
  for (int q = 0; q < 5; q++) out[q] = r[q];
}

}  // namespace coffee
