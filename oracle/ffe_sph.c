/* Log-spherical 2D EZ oracle (TEST INFRASTRUCTURE ONLY; see ffe_oracle.h).
 *
 * Restates src/algorithms/field_solver_EZ_spherical.cu and metric_sph.h in
 * FP64 and FP32 (ffe_sph_impl.inc).  The reference has no CPU build of this
 * path (its .cpp is a stub, field_solver_EZ_spherical.cpp:1-19), so no
 * reference-build vector exists: parity of this restatement with the CUDA
 * source is unpinned by the reference itself.  It is pinned instead by
 * (tests/test_oracle_sph.py): the static vacuum dipole (Omega = 0) staying
 * steady to truncation error, metric identities (sqrt(gamma) = r^3 sin(theta),
 * gamma_ii gamma^ii = 1), the axis parity of the mirror boundary, the KO
 * Nyquist decay in theta, the current vanishing for a curl-free field,
 * divergence of the dipole converging at 4th order, and the energy integral
 * of a dipole against its closed form.
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
