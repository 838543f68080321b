// Test-infrastructure harness around the reference's own CPU flat-EZ solver.
//
// Built only by oracle/Makefile, from the reference sources where they lie,
// into oracle/_ref/.  It is a checker: it never ships in the product and the
// product never calls it.
//
// The harness mirrors the driver loop of the reference apps
// (src/main-wave.cpp:56-81, src/main-alfven3d.cpp): sim_environment reads
// ./config.toml, sim_data is created, the initial fields are set, the solver is
// constructed after the initial data (src/main-alfven3d.cpp ordering, so the
// absorbing-layer background Bbg is evaluated on collocated nodes), then
// evolve_fields(time) runs with time += dt.
//
// Initial data is not computed here: it is read from a CFLD file written by the
// project's own initial-data generator, so the reference and the oracle start
// from bit-identical fields.  Buffers the CPU build never clears (dE, dB and the
// diagnostics, whose assign_dev() is a no-op on CPU: src/data/multi_array.cpp:44-54)
// are zeroed explicitly.
//
// CFLD layout (little endian): char[4] "CFLD", int32 version=1,
// int32 scalar_bytes, int32 dims[3], int32 nfields=7, then the 7 arrays
// Ex,Ey,Ez,Bx,By,Bz,P in Fortran order including guard cells.

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "algorithms/field_solver_EZ.h"
#include "data/sim_data.h"
#include "sim_env.h"

using namespace Coffee;

namespace {

multi_array<Scalar>* field_ptr(sim_data& d, int n) {
  if (n < 3) return &d.E.data(n);
  if (n < 6) return &d.B.data(n - 3);
  return &d.P;
}

// Reads the whole-domain CFLD file; with several MPI ranks each rank keeps
// the block of its own grid: local cell (i, j, k), guards included, is global
// cell (i, j, k) + Grid::offset (src/sim_env_impl.hpp:155-160).
bool read_cfld(const char* path, sim_data& d, const Grid& g) {
  FILE* f = fopen(path, "rb");
  if (!f) return false;
  char magic[4];
  int hdr[6];
  if (fread(magic, 1, 4, f) != 4 || memcmp(magic, "CFLD", 4) != 0) return false;
  if (fread(hdr, sizeof(int), 6, f) != 6) return false;
  if (hdr[1] != (int)sizeof(Scalar)) {
    fprintf(stderr, "scalar size mismatch %d vs %d\n", hdr[1], (int)sizeof(Scalar));
    return false;
  }
  const int gd[3] = {hdr[2], hdr[3], hdr[4]};
  for (int i = 0; i < 3; i++)
    if (g.offset[i] + g.dims[i] > gd[i]) {
      fprintf(stderr, "block beyond the file on axis %d: %d + %d > %d\n", i, g.offset[i],
              g.dims[i], gd[i]);
      return false;
    }
  if (hdr[5] != 7) return false;
  const size_t ng = (size_t)gd[0] * gd[1] * gd[2];
  std::vector<Scalar> buf(ng);
  for (int c = 0; c < 7; c++) {
    if (fread(buf.data(), sizeof(Scalar), ng, f) != ng) return false;
    Scalar* out = field_ptr(d, c)->host_ptr();
    for (int k = 0; k < g.dims[2]; k++)
      for (int j = 0; j < g.dims[1]; j++)
        for (int i = 0; i < g.dims[0]; i++)
          out[i + (j + (size_t)k * g.dims[1]) * g.dims[0]] =
              buf[(i + g.offset[0]) + ((j + g.offset[1]) + (size_t)(k + g.offset[2]) * gd[1]) * gd[0]];
  }
  fclose(f);
  return true;
}

bool write_cfld(const std::string& path, sim_data& d, const Grid& g) {
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return false;
  int hdr[6] = {1, (int)sizeof(Scalar), g.dims[0], g.dims[1], g.dims[2], 7};
  fwrite("CFLD", 1, 4, f);
  fwrite(hdr, sizeof(int), 6, f);
  size_t n = (size_t)g.dims[0] * g.dims[1] * g.dims[2];
  for (int c = 0; c < 7; c++) fwrite(field_ptr(d, c)->host_ptr(), sizeof(Scalar), n, f);
  fclose(f);
  return true;
}

// The energy-accounting arrays of sim_data (field_solver_EZ.cpp:447-453,
// 588-623) in the order dU_KO, dU_Epar, dU_EgtB and their _cum versions.
bool write_diag_cfld(const std::string& path, sim_data& d, const Grid& g) {
  multi_array<Scalar>* a[6] = {&d.dU_KO, &d.dU_Epar, &d.dU_EgtB,
                               &d.dU_KO_cum, &d.dU_Epar_cum, &d.dU_EgtB_cum};
  FILE* f = fopen(path.c_str(), "wb");
  if (!f) return false;
  int hdr[6] = {1, (int)sizeof(Scalar), g.dims[0], g.dims[1], g.dims[2], 6};
  fwrite("CFLD", 1, 4, f);
  fwrite(hdr, sizeof(int), 6, f);
  size_t n = (size_t)g.dims[0] * g.dims[1] * g.dims[2];
  for (int c = 0; c < 6; c++) fwrite(a[c]->host_ptr(), sizeof(Scalar), n, f);
  fclose(f);
  return true;
}

void zero(multi_array<Scalar>& a) { memset(a.host_ptr(), 0, a.size() * sizeof(Scalar)); }

// Interior sums of squares of the 7 fields, accumulated in double.
void norms(sim_data& d, const Grid& g, double out[7]) {
  for (int c = 0; c < 7; c++) {
    const Scalar* p = field_ptr(d, c)->host_ptr();
    double s = 0.0;
    for (int k = g.guard[2]; k < g.dims[2] - g.guard[2]; k++)
      for (int j = g.guard[1]; j < g.dims[1] - g.guard[1]; j++)
        for (int i = g.guard[0]; i < g.dims[0] - g.guard[0]; i++) {
          double v = p[i + (j + (size_t)k * g.dims[1]) * g.dims[0]];
          s += v * v;
        }
    out[c] = s;
  }
}

}  // namespace

int main(int argc, char* argv[]) {
  if (argc < 4) {
    fprintf(stderr,
            "usage: %s in.cfld out_prefix nsteps [dump_steps,comma,list|-] "
            "[norms.txt|-] [timing_warmup]\n",
            argv[0]);
    return 2;
  }
  const char* in_path = argv[1];
  std::string out_prefix = argv[2];
  int nsteps = atoi(argv[3]);
  std::set<int> dumps;
  if (argc > 4 && strcmp(argv[4], "-") != 0) {
    std::stringstream ss(argv[4]);
    std::string tok;
    while (std::getline(ss, tok, ',')) dumps.insert(atoi(tok.c_str()));
  }
  std::string norms_path = (argc > 5 && strcmp(argv[5], "-") != 0) ? argv[5] : "";
  int warmup = argc > 6 ? atoi(argv[6]) : -1;

  int env_argc = 1;
  char** env_argv = argv;
  sim_environment env(&env_argc, &env_argv);
  const Grid& g = env.grid();
  sim_data data(env);

  Stagger st[3] = {Stagger(0b111), Stagger(0b111), Stagger(0b111)};
  data.E.set_stagger(st);
  data.B.set_stagger(st);
  data.B0.set_stagger(st);
  for (int i = 0; i < 3; i++) {
    zero(data.E.data(i));
    zero(data.B.data(i));
    zero(data.B0.data(i));
  }
  zero(data.P);
  zero(data.divE);
  zero(data.divB);
  zero(data.dU_EgtB);
  zero(data.dU_Epar);
  zero(data.dU_KO);
  zero(data.dU_EgtB_cum);
  zero(data.dU_Epar_cum);
  zero(data.dU_KO_cum);
  if (!read_cfld(in_path, data, g)) {
    fprintf(stderr, "cannot read %s\n", in_path);
    return 3;
  }
  if (env.size() > 1 && (!dumps.empty() || argc > 5 && strcmp(argv[5], "-") != 0)) {
    fprintf(stderr, "dumps and norms are written by a single rank only\n");
    return 5;
  }

  field_solver_EZ solver(data, env);
  for (int i = 0; i < 3; i++) {
    zero(solver.dE.data(i));
    zero(solver.dB.data(i));
  }
  zero(solver.dP);

  FILE* nf = nullptr;
  if (!norms_path.empty()) nf = fopen(norms_path.c_str(), "w");
  Scalar time = 0.0;
  std::vector<double> step_ms;
  for (int step = 1; step <= nsteps; step++) {
    if (env.size() > 1) MPI_Barrier(MPI_COMM_WORLD);
    auto t0 = std::chrono::steady_clock::now();
    solver.evolve_fields(time);
    if (env.size() > 1) MPI_Barrier(MPI_COMM_WORLD);
    auto t1 = std::chrono::steady_clock::now();
    step_ms.push_back(std::chrono::duration<double, std::milli>(t1 - t0).count());
    time += env.params().dt;
    if (nf) {
      double s[7];
      norms(data, g, s);
      fprintf(nf, "%d", step);
      for (int c = 0; c < 7; c++) fprintf(nf, " %.17e", s[c]);
      fprintf(nf, "\n");
    }
    if (dumps.count(step)) {
      std::string p = out_prefix + "_s" + std::to_string(step) + ".cfld";
      std::string pd = out_prefix + "_dg_s" + std::to_string(step) + ".cfld";
      if (!write_cfld(p, data, g) || !write_diag_cfld(pd, data, g)) {
        fprintf(stderr, "cannot write %s\n", p.c_str());
        return 4;
      }
    }
  }
  if (nf) fclose(nf);
  if (warmup >= 0 && (int)step_ms.size() > warmup) {
    double tot = 0.0;
    for (size_t i = warmup; i < step_ms.size(); i++) tot += step_ms[i];
    double ms = tot / (step_ms.size() - warmup);
    double cells = (double)(g.dims[0] - 2 * g.guard[0]) * (g.dims[1] - 2 * g.guard[1]) *
                   (g.dims[2] - 2 * g.guard[2]) * env.size();
    if (env.rank() == 0)
      fprintf(stderr, "TIMING ms_per_step=%.6f cell_updates_per_s=%.6e ranks=%d\n", ms,
              cells / ms * 1e3, env.size());
  }
  return 0;
}
