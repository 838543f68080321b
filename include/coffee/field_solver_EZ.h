// field_solver_EZ (src/algorithms/field_solver_EZ.h:6-33) on one MI355X:
// same constructor, evolve_fields(Scalar time), total_energy(vector_field&)
// and public scratch members (Etmp, Btmp, dE, dB, Bbg, dP, Ptmp), computed by
// the gfx950 kernels behind the C ABI (include/coffee_hip.h).  Semantics are
// the reference CPU build's (the normative ones for this path).
#ifndef COFFEE_FIELD_SOLVER_EZ_H
#define COFFEE_FIELD_SOLVER_EZ_H

#include <cmath>

#include "coffee/sim_data.h"

namespace Coffee {

// b0 * quadru_dipole(x, y, z, ..., phase, n) (src/algorithms/pulsar_impl.hpp:57-129)
inline double coffee_quadru_dipole(const sim_params& p, double x, double y, double z,
                                   double phase, int n) {
  const double tiny = TINY;
  double c = std::cos(phase), s = std::sin(phase);
  double p1t = p.p1 * c - p.p2 * s, p2t = p.p1 * s + p.p2 * c, p3t = p.p3;
  double r2 = x * x + y * y + z * z;
  double pr = p1t * x + p2t * y + p3t * z;
  double comp = n == 0 ? x : (n == 1 ? y : z);
  double pt = n == 0 ? p1t : (n == 1 ? p2t : p3t);
  double dip = (3.0 * comp * pr / (r2 + tiny) - pt) / (std::sqrt(r2 * r2 * r2) + tiny);
  double q11t = p.q11 * c * c - 2.0 * p.q12 * c * s + p.q22 * s * s;
  double q12t = p.q12 * std::cos(2.0 * phase) + (p.q11 - p.q22) * c * s;
  double q22t = p.q22 * c * c + p.q11 * s * s + 2.0 * p.q12 * c * s;
  double q13t = p.q13 * c - p.q23 * s, q23t = p.q23 * c + p.q13 * s, q33t = -p.q11 - p.q22;
  double ox = p.q_offset_x * c - p.q_offset_y * s, oy = p.q_offset_x * s + p.q_offset_y * c;
  double x1 = x - ox, y1 = y - oy, z1 = z - p.q_offset_z;
  double q2 = x1 * x1 + y1 * y1 + z1 * z1;
  double r7 = q2 * q2 * q2 * std::sqrt(q2);
  double xqx = q11t * x1 * x1 + 2.0 * q12t * x1 * y1 + q22t * y1 * y1 +
               2.0 * (q13t * x1 + q23t * y1) * z1 + q33t * z1 * z1;
  double a0 = n == 0 ? q11t : (n == 1 ? q12t : q13t);
  double a1 = n == 0 ? q12t : (n == 1 ? q22t : q23t);
  double a2 = n == 0 ? q13t : (n == 1 ? q23t : q33t);
  double cq = n == 0 ? x1 : (n == 1 ? y1 : z1);
  double quad = (-2.0 * (a0 * x1 + a1 * y1 + a2 * z1) * q2 + 5.0 * cq * xqx) / (r7 + tiny);
  return p.b0 * (dip + quad);
}

inline coffee_params make_coffee_params(const sim_params& p, int mask_lanes) {
  coffee_params c{};
  c.dt = p.dt;
  for (int i = 0; i < 3; i++) {
    c.periodic[i] = p.periodic_boundary[i];
    c.pml[i] = p.pml[i];
    c.lower[i] = p.lower[i];
    c.size[i] = p.size[i];
  }
  c.pmllen = p.pmllen;
  c.sigpml = p.sigpml;
  c.divB_clean = p.divB_clean;
  c.ch2 = p.ch2;
  c.tau = p.tau;
  c.KOeps = p.KOeps;
  c.KO_geometry = p.KO_geometry;
  c.use_edotb_damping = p.use_edotb_damping;
  c.damp_gamma = p.damp_gamma;
  c.calc_current = p.calc_current;
  c.clean_ep = p.clean_ep;
  c.check_egb = p.check_egb;
  c.problem = p.problem;
  c.radius = p.radius;
  c.omega = p.omega;
  c.b0 = p.b0;
  c.p1 = p.p1; c.p2 = p.p2; c.p3 = p.p3;
  c.q11 = p.q11; c.q12 = p.q12; c.q13 = p.q13; c.q22 = p.q22; c.q23 = p.q23;
  c.q_offset_x = p.q_offset_x; c.q_offset_y = p.q_offset_y; c.q_offset_z = p.q_offset_z;
  c.tp_start = p.tp_start; c.tp_end = p.tp_end; c.rpert1 = p.rpert1; c.rpert2 = p.rpert2;
  c.dw0 = p.dw0; c.nT = p.nT; c.ph1 = p.ph1; c.ph2 = p.ph2; c.dph = p.dph;
  c.theta0 = p.theta0; c.drpert = p.drpert;
  c.tp_start1 = p.tp_start1; c.tp_end1 = p.tp_end1; c.rpert11 = p.rpert11;
  c.rpert21 = p.rpert21; c.dw1 = p.dw1; c.nT1 = p.nT1; c.ph11 = p.ph11; c.ph21 = p.ph21;
  c.dph1 = p.dph1;
  c.pert_type = p.pert_type;
  c.a = p.a;
  c.mask_lanes = mask_lanes;
  c.shift_ghost = p.shift_ghost;
  return c;
}

class field_solver_EZ {
 public:
  vector_field<Scalar> Etmp, Btmp, dE, dB, Bbg;
  multi_array<Scalar> dP, Ptmp;

  // mask_lanes: SIMD width of the reference CPU build whose tanh-mask
  // positions are reproduced (16 floats / 8 doubles on AVX-512).
  field_solver_EZ(sim_data& mydata, sim_environment& env,
                  int mask_lanes = sizeof(Scalar) == 4 ? 16 : 8)
      : Etmp(env.grid()), Btmp(env.grid()), dE(env.grid()), dB(env.grid()), Bbg(env.grid()),
        dP(env.grid().extent()), Ptmp(env.grid().extent()), m_data(mydata), m_env(env) {
    // Bbg = b0 * quadru_dipole at phase 0 (field_solver_EZ.cpp:66-85)
    Bbg.copy_stagger(m_data.B);
    for (int n = 0; n < 3; n++)
      Bbg.initialize(n, [&](Scalar x, Scalar y, Scalar z) {
        return (Scalar)coffee_quadru_dipole(env.params(), x, y, z, 0.0, n);
      });
    coffee_grid g = env.c_grid();
    coffee_params p = make_coffee_params(env.params(), mask_lanes);
    coffee_check(coffee_ctx_create(COFFEE_SOLVER_EZ, (int)sizeof(Scalar), &g, &p, &m_ctx),
                 "field_solver_EZ");
  }
  field_solver_EZ(const field_solver_EZ&) = delete;
  field_solver_EZ& operator=(const field_solver_EZ&) = delete;
  ~field_solver_EZ() { coffee_ctx_destroy(m_ctx); }

  // One time step; synchronises the default stream at the end, as the
  // reference GPU build does.
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
      f.Bbg[i] = Bbg.dev_ptr(i);
      f.J[i] = m_data.B0.dev_ptr(i);
    }
    f.P = m_data.P.dev_ptr();
    f.dP = dP.dev_ptr();
    f.Ptmp = Ptmp.dev_ptr();
    f.divE = m_data.divE.dev_ptr();
    f.divB = m_data.divB.dev_ptr();
    f.dU_KO = m_data.dU_KO.dev_ptr();
    f.dU_Epar = m_data.dU_Epar.dev_ptr();
    f.dU_EgtB = m_data.dU_EgtB.dev_ptr();
    f.dU_KO_cum = m_data.dU_KO_cum.dev_ptr();
    f.dU_Epar_cum = m_data.dU_Epar_cum.dev_ptr();
    f.dU_EgtB_cum = m_data.dU_EgtB_cum.dev_ptr();
    return f;
  }

  sim_data& m_data;
  sim_environment& m_env;
  coffee_ctx* m_ctx = nullptr;
};

}  // namespace Coffee

#endif
