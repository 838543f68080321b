// Flat EZ driver over the C++ host surface, shaped like the reference's
// main loops (src/main-wave.cpp:20-90, src/main-alfven3d.cpp):
//   sim_environment env(&argc, &argv);   // reads ./config.toml
//   sim_data data(env);
//   <initial data>                       // here: read from a CFLD file
//   field_solver_EZ solver(data, env);
//   for (step...) { solver.evolve_fields(time); time += dt; }
//
// usage: coffee_drive <in.cfld> <out.cfld> <nsteps>   (run where config.toml is)
// CFLD: char[4] "CFLD", int32 {1, scalar_bytes, dims0, dims1, dims2, 7}, then
// Ex,Ey,Ez,Bx,By,Bz,P in Fortran order with guards.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "coffee/field_solver_EZ.h"

using namespace Coffee;

static multi_array<Scalar>& field(sim_data& d, int c) {
  if (c < 3) return d.E.data(c);
  if (c < 6) return d.B.data(c - 3);
  return d.P;
}

int main(int argc, char* argv[]) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s in.cfld out.cfld nsteps\n", argv[0]);
    return 2;
  }
  const char* in = argv[1];
  const char* out = argv[2];
  int nsteps = atoi(argv[3]);
  int env_argc = 1;
  sim_environment env(&env_argc, &argv);
  const Grid& g = env.grid();
  sim_data data(env);
  Stagger st[3] = {Stagger(0b111), Stagger(0b111), Stagger(0b111)};
  data.E.set_stagger(st);
  data.B.set_stagger(st);
  data.B0.set_stagger(st);

  FILE* f = fopen(in, "rb");
  if (!f) return 3;
  char magic[4];
  int hdr[6];
  if (fread(magic, 1, 4, f) != 4 || memcmp(magic, "CFLD", 4) != 0 || fread(hdr, 4, 6, f) != 6 ||
      hdr[1] != (int)sizeof(Scalar) || hdr[2] != g.dims[0] || hdr[3] != g.dims[1] ||
      hdr[4] != g.dims[2] || hdr[5] != 7) {
    fprintf(stderr, "bad input file %s\n", in);
    return 3;
  }
  size_t n = (size_t)g.size();
  for (int c = 0; c < 7; c++) {
    if (fread(field(data, c).host_ptr(), sizeof(Scalar), n, f) != n) return 3;
    field(data, c).sync_to_device();
  }
  fclose(f);

  field_solver_EZ solver(data, env);
  Scalar time = 0.0;
  for (int step = 0; step < nsteps; step++) {
    solver.evolve_fields(time);
    time += env.params().dt;
  }
  Scalar wB = solver.total_energy(data.B);
  Scalar wE = solver.total_energy(data.E);
  printf("energy B %.17g E %.17g\n", (double)wB, (double)wE);

  f = fopen(out, "wb");
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
