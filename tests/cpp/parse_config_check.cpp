// Prints every sim_params field read by parse_config from ./config.toml as
// "name value" lines, for comparison with the Python parser.
#include <cstdio>

#include "coffee/sim_params.h"

using namespace Coffee;

int main(int argc, char** argv) {
  sim_params p = parse_config(argc > 1 ? argv[1] : "config.toml");
  printf("dt %.17g\nmax_steps %llu\nproblem %d\npmllen %d\nsigpml %.17g\n", (double)p.dt,
         (unsigned long long)p.max_steps, p.problem, p.pmllen, (double)p.sigpml);
  for (int i = 0; i < 3; i++)
    printf("N%d %d\nguard%d %d\nlower%d %.17g\nsize%d %.17g\nperiodic%d %d\npml%d %d\nnodes%d %d\n",
           i, p.N[i], i, p.guard[i], i, (double)p.lower[i], i, (double)p.size[i], i,
           (int)p.periodic_boundary[i], i, p.pml[i], i, p.nodes[i]);
  printf("KOeps %.17g\ntau %.17g\nch2 %.17g\nuse_edotb_damping %d\ndivB_clean %d\n",
         (double)p.KOeps, (double)p.tau, (double)p.ch2, (int)p.use_edotb_damping,
         (int)p.divB_clean);
  printf("radius %.17g\nomega %.17g\nb0 %.17g\npert_type %d\ndw0 %.17g\na %.17g\n",
         (double)p.radius, (double)p.omega, (double)p.b0, p.pert_type, (double)p.dw0, (double)p.a);
  return 0;
}
