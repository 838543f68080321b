/* Log-spherical 2D EZ oracle (TEST INFRASTRUCTURE ONLY; see ffe_oracle.h).
 *
 * Restates src/algorithms/field_solver_EZ_spherical.cu and metric_sph.h in
 * FP64 and FP32 (ffe_sph_impl.inc).  The reference has no CPU build of this
 * path (its .cpp is a stub, field_solver_EZ_spherical.cpp:1-19), so no
 * reference-build vector exists: parity of this restatement with the CUDA
 * source is unpinned by the reference itself.  It is pinned instead by what
 * the mathematics fixes, stated in the tests independently of this code:
 *   tests/test_oracle_pins_sph.py
 *     - sph_elbl + sph_rk (curl, div, force-free current in both branches,
 *       grad P, dP): one stage's dU/dt on manufactured fields with E.B != 0,
 *       div E != 0, P != 0 converges at 4th order to the textbook
 *       orthonormal-component right-hand side (complex-step derivatives);
 *     - the Michel monopole (J != 0) is stationary to rounding;
 *     - sph_project: closed form in orthonormal components, E.B = 0 and
 *       E^2 <= B^2 in the metric norm, idempotence;
 *     - sph_absorb: closed form of the radial layer (sigma in r units);
 *     - sph_pulsar, wpert_angle: star interior / surface values and the twist
 *       envelope in closed form;
 *     - ko2d6: (-1)^j / sqrt(gamma) decays by exactly (1 - KOeps);
 *     - each planted mistake (FFE_PLANT_*) fails one of these;
 *   tests/test_oracle_sph.py
 *     - the static vacuum dipole (Omega = 0) stays steady to truncation;
 *     - sph_axis: mirror parity of the guard rows, on-axis zeros;
 *     - flat KO (KO_2d0): theta-Nyquist decay by (1 - KOeps);
 *     - ffe_sph_energy: dipole energy against 4/3 (R^-3 - r_out^-3);
 *     - FP32 tracking FP64.
 */
#define _GNU_SOURCE
#include <math.h>
#include <string.h>
#include "ffe_oracle.h"

#define REAL double
#define SFX _f64
#define TINY_R 1e-11
#define SQRT_R sqrt
#include "ffe_sph_impl.inc"
#undef REAL
#undef SFX
#undef TINY_R
#undef SQRT_R

#define REAL float
#define SFX _f32
#define TINY_R ((float)(1e-7))
#define SQRT_R sqrtf
#include "ffe_sph_impl.inc"

int ffe_sph_axis_f64(const ffe_params *p, double *const U[7]) {
  sgrid_t_f64 g;
  make_sgrid_f64(p, &g);
  sph_axis_f64(&g, U);
  return 0;
}

int ffe_sph_ko_f64(const ffe_params *p, double *const U[7], double *const tmp[7]) {
  sgrid_t_f64 g;
  make_sgrid_f64(p, &g);
  sph_ko_f64(p, &g, U, tmp);
  return 0;
}

int ffe_sph_rk_stage_f64(const ffe_params *p, double *const U[7], double *const dU[7],
                         double *const L[6], double As, double Bs) {
  sgrid_t_f64 g;
  make_sgrid_f64(p, &g);
  sph_elbl_f64(&g, U, L);
  sph_rk_f64(p, &g, U, dU, L, As, Bs);
  return 0;
}

int ffe_sph_project_f64(const ffe_params *p, double *const U[7]) {
  sgrid_t_f64 g;
  make_sgrid_f64(p, &g);
  sph_project_f64(p, &g, U);
  return 0;
}

int ffe_sph_boundary_f64(const ffe_params *p, double *const U[7], double t) {
  sgrid_t_f64 g;
  make_sgrid_f64(p, &g);
  sph_pulsar_f64(p, &g, U, t);
  return 0;
}

int ffe_sph_absorb_f64(const ffe_params *p, double *const U[7], double *const Un[7]) {
  sgrid_t_f64 g;
  make_sgrid_f64(p, &g);
  sph_absorb_f64(p, &g, U, Un);
  return 0;
}
