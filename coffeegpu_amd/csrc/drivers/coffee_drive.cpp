// Driver over the C++ host surface, shaped like the reference's main loops
// (src/main-wave.cpp:20-90, src/main-alfven3d.cpp, src/main-wald.cpp):
//   sim_environment env(&argc, &argv);   // reads ./config.toml
//   sim_data data(env);
//   <initial data>                       // here: read from a CFLD file
//   <solver> solver(data, env);          // field_solver_EZ / _EZ_spherical / _gr_EZ
//   for (step...) { solver.evolve_fields(time); time += dt; }
//
// usage: coffee_drive <in.cfld> <out.cfld> <nsteps> [ez|sph|gr]   (run where config.toml is)
// CFLD: char[4] "CFLD", int32 {1, scalar_bytes, dims0, dims1, dims2, 7}, then
// Ex,Ey,Ez,Bx,By,Bz,P in Fortran order with guards (Dx,Dy,Dz for gr).
// Prints the energies of B and E and the residual monitor of the final state.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>

#include "coffee/field_solver_EZ.h"
#include "coffee/field_solver_EZ_spherical.h"
#include "coffee/field_solver_gr_EZ.h"

using namespace Coffee;

static multi_array<Scalar>& field(sim_data& d, int c) {
  if (c < 3) return d.E.data(c);
  if (c < 6) return d.B.data(c - 3);
  return d.P;
}

template <typename Solver>
static void run(sim_data& data, sim_environment& env, int nsteps) {
  Solver solver(data, env);
  Scalar time = 0.0;
  for (int step = 0; step < nsteps; step++) {
    solver.evolve_fields(time);
    time += env.params().dt;
  }
  Scalar wB = solver.total_energy(data.B);
  Scalar wE = solver.total_energy(data.E);
  double res[5];
  solver.residuals(res);
  printf("energy B %.17g E %.17g\n", (double)wB, (double)wE);
  printf("residuals energy %.17g max_divB %.17g max_EdotB %.17g max_E2mB2 %.17g nonfinite %.0f\n",
         res[0], res[1], res[2], res[3], res[4]);
}

int main(int argc, char* argv[]) {
  if (argc < 4) {
    fprintf(stderr, "usage: %s in.cfld out.cfld nsteps [ez|sph|gr]\n", argv[0]);
    return 2;
  }
  const std::string kind = argc > 4 ? argv[4] : "ez";
  if (kind != "ez" && kind != "sph" && kind != "gr") {
    fprintf(stderr, "unknown solver %s\n", kind.c_str());
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

  try {
    if (kind == "sph") run<field_solver_EZ_spherical>(data, env, nsteps);
    else if (kind == "gr") run<field_solver_gr_EZ>(data, env, nsteps);
    else run<field_solver_EZ>(data, env, nsteps);
  } catch (const std::exception& e) {
    fprintf(stderr, "%s\n", e.what());
    return 5;
  }

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
