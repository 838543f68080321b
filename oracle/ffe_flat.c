/* Flat Cartesian EZ oracle (TEST INFRASTRUCTURE ONLY; see ffe_oracle.h).
 *
 * Restates src/algorithms/field_solver_EZ.cpp, damping_boundary.cpp,
 * pulsar_impl.hpp and the single-rank periodic exchange of sim_env.cpp in
 * FP64 and FP32 (ffe_flat_impl.inc, included once per precision).
 *
 * Pinning (tests/test_oracle_flat.py):
 *   - whole steps: against the reference's own CPU build (oracle/_ref, built
 *     from the reference sources) on the golden fixtures in tests/golden/
 *     (periodic wave, Alfven packets, pulsar with absorbing layer, Alfven
 *     boundary), FP64 and FP32;
 *   - diff1_4 / diff6_2: exactness on polynomials, Fourier symbols, Nyquist;
 *   - projections: E.B = 0, E^2 <= B^2, idempotence;
 *   - KO: Nyquist decay factor, constants and linear fields untouched;
 *   - LSRK(5,4): stability polynomial vs exp(z) to 4th order;
 *   - dipole2: closed-form divergence-free field and axis values.
 */
#define _GNU_SOURCE
#include <math.h>
#include <string.h>
#include "ffe_oracle.h"

int ffe_plant = FFE_PLANT_NONE;
void ffe_set_plant(int which) { ffe_plant = (which > 0 && which < FFE_PLANT_COUNT) ? which : 0; }
int ffe_get_plant(void) { return ffe_plant; }

#define REAL double
#define SFX _f64
#define DIAG_T ffe_diag_f64
#define TINY_R 1e-11
#define SQRT_R sqrt
#define TANH_R tanh
#define FMA_R fma
#include "ffe_flat_impl.inc"
#undef REAL
#undef SFX
#undef DIAG_T
#undef TINY_R
#undef SQRT_R
#undef TANH_R
#undef FMA_R

#define REAL float
#define SFX _f32
#define DIAG_T ffe_diag_f32
#define TINY_R ((float)(1e-7))
#define SQRT_R sqrtf
#define TANH_R tanhf
#define FMA_R fmaf
#include "ffe_flat_impl.inc"

/* ---- single building blocks in FP64 for unit tests ----------------------- */

int ffe_flat_rk_stage_f64(const ffe_params *p, double *const U[7], double *const dU[7],
                          const ffe_diag_f64 *diag, double As, double Bs) {
  grid_t_f64 g; sp_t_f64 sp;
  make_grid_f64(p, &g); make_sp_f64(p, &sp);
  rk_stage_f64(p, &g, &sp, U, dU, diag, As, Bs);
  return 0;
}

int ffe_flat_ko_f64(const ffe_params *p, double *const U[7], double *const tmp[7],
                    const ffe_diag_f64 *diag) {
  grid_t_f64 g; sp_t_f64 sp;
  make_grid_f64(p, &g); make_sp_f64(p, &sp);
  ko_f64(p, &g, &sp, U, tmp, diag);
  return 0;
}

int ffe_flat_project_f64(const ffe_params *p, double *const U[7], const ffe_diag_f64 *diag) {
  grid_t_f64 g;
  make_grid_f64(p, &g);
  project_f64(&g, U, diag);
  return 0;
}

int ffe_flat_fill_guards_f64(const ffe_params *p, double *const U[7]) {
  grid_t_f64 g;
  make_grid_f64(p, &g);
  fill_guards_f64(p, &g, U);
  return 0;
}

int ffe_flat_absorb_f64(const ffe_params *p, double *const U[7], double *const Un[7],
                        double *const Bbg[3]) {
  grid_t_f64 g; sp_t_f64 sp;
  make_grid_f64(p, &g); make_sp_f64(p, &sp);
  absorb_f64(p, &g, &sp, U, Un, Bbg);
  return 0;
}

int ffe_flat_boundary_f64(const ffe_params *p, double *const U[7], double t) {
  grid_t_f64 g; sp_t_f64 sp;
  make_grid_f64(p, &g); make_sp_f64(p, &sp);
  boundary_f64(p, &g, &sp, U, t);
  return 0;
}

double ffe_dipole2(double x, double y, double z, double p1, double p2, double p3,
                   double phase, int n) {
  return dipole2_f64(x, y, z, p1, p2, p3, phase, n);
}

double ffe_quadru_dipole(const ffe_params *p, double x, double y, double z, double phase,
                         int n) {
  sp_t_f64 sp;
  make_sp_f64(p, &sp);
  return quadru_dipole_f64(&sp, x, y, z, phase, n);
}

double ffe_diff1_4(const double *f, long ijk, long s) { return df1_f64(f, ijk, s, 1.0); }
double ffe_diff6_2(const double *f, long ijk, long s) { return diff6_f64(f, ijk, s); }
