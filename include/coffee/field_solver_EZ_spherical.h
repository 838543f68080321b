// field_solver_EZ_spherical (src/algorithms/field_solver_EZ_spherical.h:6-32)
// on one MI355X: same constructor, evolve_fields(Scalar time),
// total_energy(vector_field&) and public scratch members (Etmp, Btmp, dE, dB,
// El, Bl, dP, Ptmp), computed by the gfx950 kernels behind the C ABI
// (include/coffee_hip.h).  Semantics are the reference CUDA source's: 2D
// log-spherical grid (x = ln r, theta; N[2] = 1), upper components stored.
// El and Bl (the lowered fields of kernel_compute_ElBl, .cu:175-198) are
// formed on the fly inside the stage kernels; the members exist for interface
// parity and stay zero.  The theta guard must be >= 4 (the axis mirror writes
// three rows beyond the theta = pi row).
#ifndef COFFEE_FIELD_SOLVER_EZ_SPHERICAL_H
#define COFFEE_FIELD_SOLVER_EZ_SPHERICAL_H

#include "coffee/field_solver_EZ.h"

namespace Coffee {

class field_solver_EZ_spherical {
 public:
  vector_field<Scalar> Etmp, Btmp, dE, dB, El, Bl;
  multi_array<Scalar> dP, Ptmp;

  field_solver_EZ_spherical(sim_data& mydata, sim_environment& env)
      : Etmp(env.grid()), Btmp(env.grid()), dE(env.grid()), dB(env.grid()), El(env.grid()),
        Bl(env.grid()), dP(env.grid().extent()), Ptmp(env.grid().extent()), m_data(mydata),
        m_env(env) {
    coffee_grid g = env.c_grid();
    coffee_params p = make_coffee_params(env.params(), 1);
    coffee_check(coffee_ctx_create(COFFEE_SOLVER_EZ_SPHERICAL, (int)sizeof(Scalar), &g, &p, &m_ctx),
                 "field_solver_EZ_spherical");
  }
  field_solver_EZ_spherical(const field_solver_EZ_spherical&) = delete;
  field_solver_EZ_spherical& operator=(const field_solver_EZ_spherical&) = delete;
  ~field_solver_EZ_spherical() { coffee_ctx_destroy(m_ctx); }

  // One time step (field_solver_EZ_spherical.cu:958-1012); synchronises the
  // default stream at the end, as the reference does.
  void evolve_fields(Scalar time) {
    coffee_fields f = fields();
    coffee_check(coffee_ez_step(m_ctx, &f, (double)time, nullptr), "evolve_fields", m_ctx);
    hip_check(hipStreamSynchronize(nullptr), "evolve_fields");
  }

  // sum gamma_ii f^i f^i sqrt(gamma) dx dtheta over r >= radius outside the
  // absorbing layer (.cu:1014-1052)
  Scalar total_energy(vector_field<Scalar>& fld) {
    void* comp[3] = {fld.dev_ptr(0), fld.dev_ptr(1), fld.dev_ptr(2)};
    double w = 0;
    coffee_check(coffee_total_energy(m_ctx, comp, &w, nullptr), "total_energy", m_ctx);
    return (Scalar)w;
  }

  // Residual monitor (coffee_residuals): energy, max |div B|, max |E.B|/B^2,
  // max (E^2 - B^2)/B^2, non-finite cells.
  void residuals(double out[5]) {
    coffee_fields f = fields();
    coffee_check(coffee_residuals(m_ctx, &f, out, nullptr), "residuals", m_ctx);
  }

  coffee_ctx* context() { return m_ctx; }

 private:
  coffee_fields fields() {
    coffee_fields f{};
    for (int i = 0; i < 3; i++) {
      f.E[i] = m_data.E.dev_ptr(i);
      f.B[i] = m_data.B.dev_ptr(i);
      f.dE[i] = dE.dev_ptr(i);
      f.dB[i] = dB.dev_ptr(i);
      f.Etmp[i] = Etmp.dev_ptr(i);
      f.Btmp[i] = Btmp.dev_ptr(i);
      f.J[i] = m_data.B0.dev_ptr(i);
    }
    f.P = m_data.P.dev_ptr();
    f.dP = dP.dev_ptr();
    f.Ptmp = Ptmp.dev_ptr();
    f.divE = m_data.divE.dev_ptr();
    f.divB = m_data.divB.dev_ptr();
    return f;
  }

  sim_data& m_data;
  sim_environment& m_env;
  coffee_ctx* m_ctx = nullptr;
};

}  // namespace Coffee

#endif
