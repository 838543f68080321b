// Kill-and-resume driver over the C++ host surface, shaped like the
// reference's restart path (src/main-alfven3d.cpp with is_restart: the
// exporter's load_snapshot before the time loop, save_snapshot every
// snapshot_interval steps, write_output every data_interval steps):
//
//   sim_environment env(&argc, &argv);           // ./config.toml
//   sim_data data(env);  <initial data from a CFLD file>
//   <solver> solver(data, env);  k steps;
//   exporter.save_snapshot(snapshot, data, k, time);
//   exporter.write_output(data, k, time);          // fld.%05d.h5
//   -- the run ends; a new one starts --
//   sim_data data2(env);  exporter.load_snapshot(snapshot, data2, step, time);
//   <solver> solver2(data2, env);  run to nsteps;  write the final state.
//
// usage: coffee_restart <in.cfld> <out.cfld> <nsteps> <k> <snapshot.h5> <data_dir> [ez|sph|gr]
// CFLD as coffee_drive: char[4] "CFLD", int32 {1, scalar_bytes, dims0, dims1,
// dims2, 7}, then Ex,Ey,Ez,Bx,By,Bz,P in Fortran order with guards.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>

#include "coffee/data_exporter.h"
#include "coffee/field_solver_EZ.h"
#include "coffee/field_solver_EZ_spherical.h"
#include "coffee/field_solver_gr_EZ.h"

using namespace Coffee;

static multi_array<Scalar>& field(sim_data& d, int c) {
  if (c < 3) return d.E.data(c);
  if (c < 6) return d.B.data(c - 3);
  return d.P;
}

// E and B on collocated nodes, as the reference's problem set-ups declare
// them (src/user_alfven.hpp:35-38, user_pulsar.hpp:6-9); B0 keeps sim_data's
// default stagger.
static void collocate(sim_data& d) {
  Stagger st[3] = {Stagger(0b111), Stagger(0b111), Stagger(0b111)};
  d.E.set_stagger(st);
  d.B.set_stagger(st);
}

template <typename Solver>
static Scalar advance(sim_data& data, sim_environment& env, int from, int to, Scalar time) {
  Solver solver(data, env);
  for (int step = from; step < to; step++) {
    solver.evolve_fields(time);
    time += env.params().dt;
  }
  return time;
}

template <typename Solver>
static int run(sim_environment& env, const char* in, const char* out, int nsteps, int k,
               const std::string& snap, const std::string& dir) {
  const Grid& g = env.grid();
  const size_t n = (size_t)g.size();
  int hdr[6];
  uint32_t step0 = 0;
  data_exporter exporter(env, step0, dir);
  {
    sim_data data(env);
    collocate(data);
    FILE* f = fopen(in, "rb");
    if (!f) return 3;
    char magic[4];
    if (fread(magic, 1, 4, f) != 4 || memcmp(magic, "CFLD", 4) != 0 || fread(hdr, 4, 6, f) != 6 ||
        hdr[1] != (int)sizeof(Scalar) || hdr[2] != g.dims[0] || hdr[3] != g.dims[1] ||
        hdr[4] != g.dims[2] || hdr[5] != 7) {
      fprintf(stderr, "bad input file %s\n", in);
      return 3;
    }
    for (int c = 0; c < 7; c++) {
      if (fread(field(data, c).host_ptr(), sizeof(Scalar), n, f) != n) return 3;
      field(data, c).sync_to_device();
    }
    fclose(f);
    Scalar time = advance<Solver>(data, env, 0, k, 0.0);
    exporter.save_snapshot(snap, data, (uint32_t)k, time);
    exporter.write_output(data, (uint32_t)k, time);
    printf("snapshot step %d time %.17g\n", k, (double)time);
  }
  sim_data data(env);
  collocate(data);
  uint32_t step = 0;
  Scalar time = 0;
  exporter.load_snapshot(snap, data, step, time);
  printf("resumed step %u time %.17g\n", step, (double)time);
  advance<Solver>(data, env, (int)step, nsteps, time);
  FILE* f = fopen(out, "wb");
  if (!f) return 4;
  fwrite("CFLD", 1, 4, f);
  fwrite(hdr, 4, 6, f);
  for (int c = 0; c < 7; c++) {
    field(data, c).sync_to_host();
    fwrite(field(data, c).host_ptr(), sizeof(Scalar), n, f);
  }
  fclose(f);
  return 0;
}

int main(int argc, char* argv[]) {
  if (argc < 7) {
    fprintf(stderr, "usage: %s in.cfld out.cfld nsteps k snapshot.h5 data_dir [ez|sph|gr]\n",
            argv[0]);
    return 2;
  }
  const std::string kind = argc > 7 ? argv[7] : "ez";
  const int nsteps = atoi(argv[3]), k = atoi(argv[4]);
  if (k < 0 || k > nsteps) {
    fprintf(stderr, "need 0 <= k <= nsteps\n");
    return 2;
  }
  int env_argc = 1;
  sim_environment env(&env_argc, &argv);
  try {
    if (kind == "sph") return run<field_solver_EZ_spherical>(env, argv[1], argv[2], nsteps, k, argv[5], argv[6]);
    if (kind == "gr") return run<field_solver_gr_EZ>(env, argv[1], argv[2], nsteps, k, argv[5], argv[6]);
    if (kind == "ez") return run<field_solver_EZ>(env, argv[1], argv[2], nsteps, k, argv[5], argv[6]);
  } catch (const std::exception& e) {
    fprintf(stderr, "%s\n", e.what());
    return 5;
  }
  fprintf(stderr, "unknown solver %s\n", kind.c_str());
  return 2;
}
