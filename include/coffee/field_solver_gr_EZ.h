// field_solver_gr_EZ (src/algorithms/field_solver_gr_EZ.h:6-32) on one
// MI355X: Cartesian Kerr-Schild black hole (M = 1, spin sim_params::a).  Same
// constructor, evolve_fields(Scalar time), total_energy(vector_field&) and
// public scratch members (Dtmp, Btmp, dD, dB, Ed, Hd, dP, Ptmp); the
// sim_data "E" arrays hold D (upper components), as in the reference.  Ed, Hd
// (kernel_compute_E_gr / _H_gr, .cu:90-186) are formed on the fly inside the
// stage kernels; the members exist for interface parity and stay zero.
// total_energy is declared by the reference but never defined there; here it
// is sum gamma_ij f^i f^j sqrt(gamma) outside the absorbing layer and the
// excision region.
#ifndef COFFEE_FIELD_SOLVER_GR_EZ_H
#define COFFEE_FIELD_SOLVER_GR_EZ_H

#include "coffee/field_solver_EZ.h"

namespace Coffee {

class field_solver_gr_EZ {
 public:
  vector_field<Scalar> Dtmp, Btmp, dD, dB, Ed, Hd;
  multi_array<Scalar> dP, Ptmp;

  field_solver_gr_EZ(sim_data& mydata, sim_environment& env)
      : Dtmp(env.grid()), Btmp(env.grid()), dD(env.grid()), dB(env.grid()), Ed(env.grid()),
        Hd(env.grid()), dP(env.grid().extent()), Ptmp(env.grid().extent()), m_data(mydata),
        m_env(env) {
    coffee_grid g = env.c_grid();
    coffee_params p = make_coffee_params(env.params(), 1);
    coffee_check(coffee_ctx_create(COFFEE_SOLVER_GR_EZ, (int)sizeof(Scalar), &g, &p, &m_ctx),
                 "field_solver_gr_EZ");
  }
  field_solver_gr_EZ(const field_solver_gr_EZ&) = delete;
  field_solver_gr_EZ& operator=(const field_solver_gr_EZ&) = delete;
  ~field_solver_gr_EZ() { coffee_ctx_destroy(m_ctx); }

  // One time step (field_solver_gr_EZ.cu:862-910; the time argument is
  // unused there too); synchronises the default stream at the end.
  void evolve_fields(Scalar time) {
    coffee_fields f = fields();
    coffee_check(coffee_ez_step(m_ctx, &f, (double)time, nullptr), "evolve_fields", m_ctx);
    hip_check(hipStreamSynchronize(nullptr), "evolve_fields");
  }

  Scalar total_energy(vector_field<Scalar>& fld) {
    void* comp[3] = {fld.dev_ptr(0), fld.dev_ptr(1), fld.dev_ptr(2)};
    double w = 0;
    coffee_check(coffee_total_energy(m_ctx, comp, &w, nullptr), "total_energy", m_ctx);
    return (Scalar)w;
  }

  // Residual monitor (coffee_residuals) with the 3-metric: energy,
  // max |div B|, max |D.B|/B^2, max (D^2 - B^2)/B^2, non-finite cells.
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
      f.dE[i] = dD.dev_ptr(i);
      f.dB[i] = dB.dev_ptr(i);
      f.Etmp[i] = Dtmp.dev_ptr(i);
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
