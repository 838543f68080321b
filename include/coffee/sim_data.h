// sim_data (src/data/sim_data.h:11-24, sim_data.cpp:6-44): the state arrays
// E, B, P, the current J (stored in B0, as the reference does), divE, divB
// and the per-cell energy-accounting arrays.
#ifndef COFFEE_SIM_DATA_H
#define COFFEE_SIM_DATA_H

#include "coffee/fields.h"
#include "coffee/sim_env.h"

namespace Coffee {

struct sim_data {
  vector_field<Scalar> E, B, B0;
  multi_array<Scalar> P, divB, divE;
  multi_array<Scalar> dU_EgtB, dU_Epar, dU_KO;
  multi_array<Scalar> dU_EgtB_cum, dU_Epar_cum, dU_KO_cum;
  const sim_environment& env;

  explicit sim_data(const sim_environment& e)
      : E(e.grid()), B(e.grid()), B0(e.grid()), P(e.grid().extent()), divB(e.grid().extent()),
        divE(e.grid().extent()), dU_EgtB(e.grid().extent()), dU_Epar(e.grid().extent()),
        dU_KO(e.grid().extent()), dU_EgtB_cum(e.grid().extent()), dU_Epar_cum(e.grid().extent()),
        dU_KO_cum(e.grid().extent()), env(e) {
    Stagger st_e[3] = {Stagger(0b110), Stagger(0b101), Stagger(0b011)};
    Stagger st_b[3] = {Stagger(0b001), Stagger(0b010), Stagger(0b100)};
    B.set_stagger(st_b);
    B0.set_stagger(st_b);
    E.set_stagger(st_e);
  }
  void sync_to_host() {
    E.sync_to_host();
    B.sync_to_host();
    B0.sync_to_host();
    P.sync_to_host();
    divB.sync_to_host();
    divE.sync_to_host();
  }
  void sync_to_device() {
    E.sync_to_device();
    B.sync_to_device();
    B0.sync_to_device();
    P.sync_to_device();
    divB.sync_to_device();
    divE.sync_to_device();
  }
};

inline void sim_environment::send_guard_cells(sim_data& d) {
  void* a[7] = {d.E.dev_ptr(0), d.E.dev_ptr(1), d.E.dev_ptr(2), d.B.dev_ptr(0),
                d.B.dev_ptr(1), d.B.dev_ptr(2), d.P.dev_ptr()};
  fill(a, 7);
}

}  // namespace Coffee

#endif
