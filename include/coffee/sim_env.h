// sim_environment (src/sim_env.h:17-87, sim_env_impl.hpp:12-177) on one GPU:
// reads ./config.toml (or takes a sim_params), builds the one-domain Grid and
// replaces the MPI guard exchange by the single-GPU periodic fill
// (coffee_fill_guards_arrays).  rank() is 0 and size() is 1.
#ifndef COFFEE_SIM_ENV_H
#define COFFEE_SIM_ENV_H

#include <string>

#include "coffee/grid.h"
#include "coffee/multi_array.h"
#include "coffee/sim_params.h"

namespace Coffee {

struct sim_data;

class sim_environment {
 public:
  // Reference signature: command-line arguments are accepted and ignored
  // except -r (restart flag); parameters come from ./config.toml.
  sim_environment(int* argc, char*** argv) {
    if (argc && argv)
      for (int i = 1; i < *argc; i++) {
        std::string a = (*argv)[i];
        if (a == "-r" || a.rfind("--restart-file", 0) == 0) m_is_restart = true;
      }
    init(parse_config("config.toml"));
  }
  explicit sim_environment(const sim_params& p) { init(p); }
  sim_environment(const sim_environment&) = delete;
  sim_environment& operator=(const sim_environment&) = delete;

  const Grid& grid() const { return m_grid; }
  const sim_params& params() const { return m_params; }
  int size() const { return 1; }
  int rank() const { return 0; }
  int mpi_dims(int) const { return 1; }
  int mpi_coord(int) const { return 0; }
  bool is_periodic(int n) const { return m_params.periodic_boundary[n]; }
  bool is_boundary(int) const { return true; }
  bool is_restart() const { return m_is_restart; }
  const std::string& restart_file() const { return m_restart_file; }

  // Guard cells of E, B, P: periodic axes take the interior one period away.
  inline void send_guard_cells(sim_data& data);
  void send_guard_cell_array(multi_array<Scalar>& array) {
    void* a[1] = {array.dev_ptr()};
    fill(a, 1);
  }

  coffee_grid c_grid() const {
    coffee_grid g{};
    for (int i = 0; i < 3; i++) {
      g.dims[i] = m_grid.dims[i];
      g.guard[i] = m_grid.guard[i];
      g.delta[i] = m_grid.delta[i];
      g.lower[i] = m_grid.lower[i];
      g.sizes[i] = m_grid.sizes[i];
    }
    return g;
  }

 private:
  void init(const sim_params& p) {
    m_params = p;
    for (int i = 0; i < 3; i++) {
      m_grid.guard[i] = p.guard[i];
      m_grid.sizes[i] = p.size[i];
      m_grid.lower[i] = p.lower[i];
      m_grid.dims[i] = p.N[i] + 2 * p.guard[i];
      m_grid.delta[i] = p.size[i] / p.N[i];
      m_grid.inv_delta[i] = 1.0 / m_grid.delta[i];
    }
  }
  void fill(void** arrays, int n) {
    coffee_grid g = c_grid();
    int per[3] = {m_params.periodic_boundary[0], m_params.periodic_boundary[1],
                  m_params.periodic_boundary[2]};
    coffee_check(coffee_fill_guards_arrays(&g, per, arrays, n, (int)sizeof(Scalar), nullptr),
                 "send_guard_cells");
  }

  sim_params m_params;
  Grid m_grid;
  bool m_is_restart = false;
  std::string m_restart_file = "snapshot.h5";
};

}  // namespace Coffee

#endif
