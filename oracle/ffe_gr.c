/* Kerr-Schild GR EZ oracle (TEST INFRASTRUCTURE ONLY; see ffe_oracle.h).
 *
 * Restates, in FP64, src/algorithms/field_solver_gr_EZ.cu (the reference's
 * only implementation of this path: CUDA, no CPU build), the outer absorbing
 * layer of src/algorithms/boundary.cu:7-79 and the Cartesian Kerr-Schild
 * metric of src/algorithms/metric_cks_impl.hpp (M = 1).  Every metric
 * component is evaluated with the reference's own expression, in the same
 * order of operations, at the same points (including the x +- k*delta points
 * of div4, .cu:44-88).
 *
 * Parity of this restatement with the CUDA source is unpinned by the
 * reference itself (no runnable build).  It is pinned instead by:
 *   tests/test_oracle_gr.py
 *     - cks_at / cks_r: metric identities (gamma_ij gamma^jk = delta,
 *       det gamma_ij = gamma, alpha sqrt(gamma) = 1, beta^i = gamma^ij beta_j)
 *       and the a = 0 Schwarzschild form;
 *     - gr_eh: the constitutive relations (E.B = alpha D_i B^i, ...);
 *     - gr_rk1 (vacuum part): the exact Wald solution is stationary to 4th
 *       order; a wrong spin is not;
 *     - gr_absorb: cells with r < 0.01 r_H are zeroed;
 *   tests/test_oracle_pins_gr.py
 *     - gr_eh + gr_rk1 + div4 (curl, div, current, cleaning with its tanh
 *       weight, dP): one stage's dU/dt on manufactured D, B, P with E.B != 0,
 *       div D != 0, P != 0 converges at 4th order to the 3+1 force-free
 *       right-hand side built from the Kerr-Schild 4-metric's own definition
 *       (complex-step derivatives), at a = 0, 0.7, 0.99;
 *     - gr_project: closed form with metric dot products, D.B = 0,
 *       D^2 <= B^2, idempotence;
 *     - gr_absorb: closed form of the outer layer and the horizon excision;
 *     - gr_ko: checkerboard decay by (1 - 3 KOeps), linear fields untouched;
 *     - each planted mistake (FFE_PLANT_*) fails one of these.
 */
#define _GNU_SOURCE
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "ffe_oracle.h"

#define TINY_G 1e-11 /* src/data/typedefs.h:11 (FP64 build) */

static inline double sq(double v) { return v * v; }
static inline double cube(double v) { return v * v * v; }
static inline double mx(double a, double b) { return fmax(a, b); }

typedef struct {
  double gd11, gd12, gd13, gd22, gd23, gd33;
  double gu11, gu12, gu13, gu22, gu23, gu33;
  double bd1, bd2, bd3, bu1, bu2, bu3;
  double alpha, sg, r;
} cks_t;

/* metric_cks_impl.hpp:17-24 */
static inline double cks_r(double a, double x, double y, double z) {
  double R2 = x * x + y * y + z * z;
  return sqrt((R2 - a * a + sqrt(mx(sq(R2 - a * a) + 4.0 * sq(a * z), TINY_G))) / 2.0);
}

/* metric_cks_impl.hpp:31-181, each component as written there. */
static void cks_at(double a, double x, double y, double z, cks_t *m) {
  const double r = cks_r(a, x, y, z);
  const double Q = mx(r * r * r * r + a * a * z * z, TINY_G);
  const double Qu = mx(2.0 * r * r * r + r * r * r * r + a * a * z * z, TINY_G);
  const double A1 = mx(a * a + r * r, TINY_G);
  const double A2 = mx(sq(a * a + r * r), TINY_G);
  m->r = r;
  m->bd1 = 2.0 * r * r * r * (r * x + a * y) / A1 / Q;
  m->bd2 = 2.0 * r * r * r * (-a * x + r * y) / A1 / Q;
  m->bd3 = 2.0 * r * r * z / Q;
  m->bu1 = 2.0 * r * r * r * (r * x + a * y) / A1 / Qu;
  m->bu2 = 2 * r * r * r * (-a * x + r * y) / A1 / Qu;
  m->bu3 = 2.0 * r * r * z / Qu;
  m->sg = sqrt(1.0 + 2.0 * r * r * r / Q);
  m->alpha = sqrt(1.0 / (1.0 + 2.0 * r * r * r / Q));
  m->gd11 = 1.0 + 2.0 * r * r * r * sq(r * x + a * y) / A2 / Q;
  m->gd12 = 2.0 * r * r * r * (r * x + a * y) * (-a * x + r * y) / A2 / Q;
  m->gd13 = 2.0 * r * r * (r * x + a * y) * z / A1 / Q;
  m->gd22 = 1.0 + 2.0 * r * r * r * sq(a * x - r * y) / A2 / Q;
  m->gd23 = 2.0 * r * r * (-a * x + r * y) * z / A1 / Q;
  m->gd33 = 1.0 + 2.0 * r * z * z / Q;
  m->gu11 = 1.0 - 2.0 * r * r * r * sq(r * x + a * y) / A2 / Qu;
  m->gu12 = -2.0 * r * r * r * (r * x + a * y) * (-a * x + r * y) / A2 / Qu;
  m->gu13 = -2.0 * r * r * (r * x + a * y) * z / A1 / Qu;
  m->gu22 = 1.0 - 2.0 * r * r * r * sq(a * x - r * y) / A2 / Qu;
  m->gu23 = 2.0 * r * r * (a * x - r * y) * z / A1 / Qu;
  m->gu33 = 1.0 - 2.0 * r * z * z / Qu;
}

/* get_sqrt_gamma (metric_cks_impl.hpp:72-81) alone, for div4's neighbours */
static inline double cks_sg(double a, double x, double y, double z) {
  const double r = cks_r(a, x, y, z);
  return sqrt(1.0 + 2.0 * r * r * r / mx(r * r * r * r + a * a * z * z, TINY_G));
}

void ffe_gr_metric_f64(double a, double x, double y, double z, double out[21]) {
  cks_t m;
  cks_at(a, x, y, z, &m);
  const double v[21] = {m.gd11, m.gd12, m.gd13, m.gd22, m.gd23, m.gd33, m.gu11,
                        m.gu12, m.gu13, m.gu22, m.gu23, m.gu33, m.bd1,  m.bd2,
                        m.bd3,  m.bu1,  m.bu2,  m.bu3,  m.alpha, m.sg,  m.r};
  memcpy(out, v, sizeof(v));
}

typedef struct {
  int dims[3], guard[3], N[3];
  double delta[3], inv_delta[3], lower[3];
  long sy, sz;
} ggrid_t;

static void make_ggrid(const ffe_params *p, ggrid_t *g) {
  for (int i = 0; i < 3; i++) {
    g->N[i] = p->N[i];
    g->guard[i] = p->guard[i];
    g->dims[i] = p->N[i] + 2 * p->guard[i];
    g->lower[i] = p->lower[i];
    g->delta[i] = p->size[i] / p->N[i];
    g->inv_delta[i] = 1.0 / g->delta[i];
  }
  g->sy = g->dims[0];
  g->sz = (long)g->dims[0] * g->dims[1];
}

/* Grid::pos(i, n, stagger = 1) (src/data/grid.h:116-134) */
static inline double gpos(const ggrid_t *g, int ax, int n) {
  return g->lower[ax] + g->delta[ax] * ((double)(n - g->guard[ax]) + 0.0);
}

/* finite_diff.h:13-18, 85-92 */
static inline double df1(const double *f, long ijk, long s, double inv) {
  return (f[ijk - 2 * s] - 8.0 * f[ijk - s] + 8.0 * f[ijk + s] - f[ijk + 2 * s]) / 12.0 * inv;
}
static inline double diff6_2(const double *f, long ijk, long s) {
  return (f[ijk - 3 * s] - 6.0 * f[ijk - 2 * s] + 15.0 * f[ijk - s] - 20.0 * f[ijk] +
          15.0 * f[ijk + s] - 6.0 * f[ijk + 2 * s] + f[ijk + 3 * s]);
}

/* div4 (.cu:44-88).  SGT: sqrt(gamma) at a tap (at the centre under the
 * DIV_CENTRE planted mistake). */
#define SGT(xx, yy, zz) \
  (ffe_plant == FFE_PLANT_DIV_CENTRE ? cks_sg(a, x, y, z) : cks_sg(a, xx, yy, zz))
static double div4(const ggrid_t *g, double a, const double *fx, const double *fy,
                   const double *fz, long ijk, double x, double y, double z) {
  const double d0 = g->delta[0], d1 = g->delta[1], d2 = g->delta[2];
  double tx = (fx[ijk - 2] * SGT(x - 2.0 * d0, y, z) -
               8.0 * fx[ijk - 1] * SGT(x - 1.0 * d0, y, z) +
               8.0 * fx[ijk + 1] * SGT(x + 1.0 * d0, y, z) -
               fx[ijk + 2] * SGT(x + 2.0 * d0, y, z)) /
              12.0 / d0;
  long s = g->sy;
  double ty = (fy[ijk - 2 * s] * SGT(x, y - 2.0 * d1, z) -
               8.0 * fy[ijk - 1 * s] * SGT(x, y - 1.0 * d1, z) +
               8.0 * fy[ijk + 1 * s] * SGT(x, y + 1.0 * d1, z) -
               fy[ijk + 2 * s] * SGT(x, y + 2.0 * d1, z)) /
              12.0 / d1;
  s = g->sz;
  double tz = (fz[ijk - 2 * s] * SGT(x, y, z - 2.0 * d2) -
               8.0 * fz[ijk - 1 * s] * SGT(x, y, z - 1.0 * d2) +
               8.0 * fz[ijk + 1 * s] * SGT(x, y, z + 1.0 * d2) -
               fz[ijk + 2 * s] * SGT(x, y, z + 2.0 * d2)) /
              12.0 / d2;
  return (tx + ty + tz) / cks_sg(a, x, y, z);
}
#undef SGT

#define ALL_CELLS(g)                                            \
  _Pragma("omp parallel for schedule(static)")                  \
  for (int k = 0; k < (g)->dims[2]; k++)                        \
    for (int j = 0; j < (g)->dims[1]; j++)                      \
      for (int i = 0; i < (g)->dims[0]; i++)

#define INTERIOR(g)                                                           \
  _Pragma("omp parallel for schedule(static)")                                \
  for (int k = (g)->guard[2]; k < (g)->dims[2] - (g)->guard[2]; k++)          \
    for (int j = (g)->guard[1]; j < (g)->dims[1] - (g)->guard[1]; j++)        \
      for (int i = (g)->guard[0]; i < (g)->dims[0] - (g)->guard[0]; i++)

/* kernel_compute_E_gr / kernel_compute_H_gr (.cu:90-186), every cell
 * including guards.  U = D (the evolved "E" arrays), B, P; EH = Ed, Hd. */
static void gr_eh(const ggrid_t *g, double a, double *const U[7], double *const EH[6]) {
  ALL_CELLS(g) {
    long ijk = i + j * g->sy + k * g->sz;
    double x = gpos(g, 0, i), y = gpos(g, 1, j), z = gpos(g, 2, k);
    cks_t m;
    cks_at(a, x, y, z, &m);
    const double Dx = U[0][ijk], Dy = U[1][ijk], Dz = U[2][ijk];
    const double Bx = U[3][ijk], By = U[4][ijk], Bz = U[5][ijk];
    double Ddx = m.gd11 * Dx + m.gd12 * Dy + m.gd13 * Dz;
    double Ddy = m.gd12 * Dx + m.gd22 * Dy + m.gd23 * Dz;
    double Ddz = m.gd13 * Dx + m.gd23 * Dy + m.gd33 * Dz;
    EH[0][ijk] = m.alpha * Ddx + m.sg * (m.bu2 * Bz - m.bu3 * By);
    EH[1][ijk] = m.alpha * Ddy + m.sg * (m.bu3 * Bx - m.bu1 * Bz);
    EH[2][ijk] = m.alpha * Ddz + m.sg * (m.bu1 * By - m.bu2 * Bx);
    double Bdx = m.gd11 * Bx + m.gd12 * By + m.gd13 * Bz;
    double Bdy = m.gd12 * Bx + m.gd22 * By + m.gd23 * Bz;
    double Bdz = m.gd13 * Bx + m.gd23 * By + m.gd33 * Bz;
    EH[3][ijk] = m.alpha * Bdx - m.sg * (m.bu2 * Dz - m.bu3 * Dy);
    EH[4][ijk] = m.alpha * Bdy - m.sg * (m.bu3 * Dx - m.bu1 * Dz);
    EH[5][ijk] = m.alpha * Bdz - m.sg * (m.bu1 * Dy - m.bu2 * Dx);
  }
}

/* kernel_rk_step1_gr (.cu:188-373): writes dU (and J, divB, divD when the
 * diagnostics pointers are given). */
static void gr_rk1(const ffe_params *p, const ggrid_t *g, double *const U[7],
                   double *const dU[7], double *const EH[6], const ffe_diag_f64 *diag,
                   double As) {
  const double a = p->a, dt = p->dt, tau = p->tau;
  const long sy = g->sy, sz = g->sz;
  INTERIOR(g) {
    long ijk = i + j * sy + k * sz;
    double x = gpos(g, 0, i), y = gpos(g, 1, j), z = gpos(g, 2, k);
    cks_t m;
    cks_at(a, x, y, z, &m);
    const double r = m.r, gmsqrt = m.sg, alpha = m.alpha;
    const double gmr = ffe_plant == FFE_PLANT_ROT_NO_SQRTG ? 1.0 : gmsqrt;
    const double gmj = ffe_plant == FFE_PLANT_J_NO_SQRTG ? 1.0 : gmsqrt;
    const double *Ex = EH[0], *Ey = EH[1], *Ez = EH[2], *Hx = EH[3], *Hy = EH[4], *Hz = EH[5];
    const double *Dx = U[0], *Dy = U[1], *Dz = U[2], *Bx = U[3], *By = U[4], *Bz = U[5];
    const double *P = U[6];
    const double i0 = g->inv_delta[0], i1 = g->inv_delta[1], i2 = g->inv_delta[2];

    double rotHx = (df1(Hz, ijk, sy, i1) - df1(Hy, ijk, sz, i2)) / gmr;
    double rotHy = (df1(Hx, ijk, sz, i2) - df1(Hz, ijk, 1, i0)) / gmr;
    double rotHz = (df1(Hy, ijk, 1, i0) - df1(Hx, ijk, sy, i1)) / gmr;
    double rotEx = (df1(Ez, ijk, sy, i1) - df1(Ey, ijk, sz, i2)) / gmr;
    double rotEy = (df1(Ex, ijk, sz, i2) - df1(Ez, ijk, 1, i0)) / gmr;
    double rotEz = (df1(Ey, ijk, 1, i0) - df1(Ex, ijk, sy, i1)) / gmr;

    double divD = div4(g, a, Dx, Dy, Dz, ijk, x, y, z);
    double divB = div4(g, a, Bx, By, Bz, ijk, x, y, z);

    double Bdx = m.gd11 * Bx[ijk] + m.gd12 * By[ijk] + m.gd13 * Bz[ijk];
    double Bdy = m.gd12 * Bx[ijk] + m.gd22 * By[ijk] + m.gd23 * Bz[ijk];
    double Bdz = m.gd13 * Bx[ijk] + m.gd23 * By[ijk] + m.gd33 * Bz[ijk];
    double Ddx = m.gd11 * Dx[ijk] + m.gd12 * Dy[ijk] + m.gd13 * Dz[ijk];
    double Ddy = m.gd12 * Dx[ijk] + m.gd22 * Dy[ijk] + m.gd23 * Dz[ijk];
    double Ddz = m.gd13 * Dx[ijk] + m.gd23 * Dy[ijk] + m.gd33 * Dz[ijk];

    double Jx, Jy, Jz;
    if (p->calc_current) {
      double B2 = Bx[ijk] * Bdx + By[ijk] * Bdy + Bz[ijk] * Bdz;
      if (B2 < TINY_G) B2 = TINY_G;
      double Jp = (Bdx * rotHx + Bdy * rotHy + Bdz * rotHz) - (Ddx * rotEx + Ddy * rotEy + Ddz * rotEz);
      Jx = (divD * (Ey[ijk] * Bdz - Ez[ijk] * Bdy) / gmj + Jp * Bx[ijk]) / B2;
      Jy = (divD * (Ez[ijk] * Bdx - Ex[ijk] * Bdz) / gmj + Jp * By[ijk]) / B2;
      Jz = (divD * (Ex[ijk] * Bdy - Ey[ijk] * Bdx) / gmj + Jp * Bz[ijk]) / B2;
      if (ffe_plant == FFE_PLANT_J_SIGN) {
        Jx = -Jx;
        Jy = -Jy;
        Jz = -Jz;
      }
    } else {
      Jx = 0.0;
      Jy = 0.0;
      Jz = 0.0;
    }

    double Pxd = df1(P, ijk, 1, i0);
    double Pyd = df1(P, ijk, sy, i1);
    double Pzd = df1(P, ijk, sz, i2);
    double Pxu = m.gu11 * Pxd + m.gu12 * Pyd + m.gu13 * Pzd;
    double Pyu = m.gu12 * Pxd + m.gu22 * Pyd + m.gu23 * Pzd;
    double Pzu = m.gu13 * Pxd + m.gu23 * Pyd + m.gu33 * Pzd;
    if (ffe_plant == FFE_PLANT_GRADP_LOWER) {
      Pxu = m.gd11 * Pxd + m.gd12 * Pyd + m.gd13 * Pzd;
      Pyu = m.gd12 * Pxd + m.gd22 * Pyd + m.gd23 * Pzd;
      Pzu = m.gd13 * Pxd + m.gd23 * Pyd + m.gd33 * Pzd;
    }

    double r0 = 5.0, del = 1.0;
    double shape = 0.5 * (1.0 - tanh((r - r0) / del));
    if (ffe_plant == FFE_PLANT_CLEAN_SHAPE) shape = 0.5 * (1.0 + tanh((r - r0) / del));

    if (p->divB_clean) {
      dU[3][ijk] = As * dU[3][ijk] + dt * (-rotEx + (-alpha * Pxu + m.bu1 * divB) * shape);
      dU[4][ijk] = As * dU[4][ijk] + dt * (-rotEy + (-alpha * Pyu + m.bu2 * divB) * shape);
      dU[5][ijk] = As * dU[5][ijk] + dt * (-rotEz + (-alpha * Pzu + m.bu3 * divB) * shape);
    } else {
      dU[3][ijk] = As * dU[3][ijk] - dt * rotEx;
      dU[4][ijk] = As * dU[4][ijk] - dt * rotEy;
      dU[5][ijk] = As * dU[5][ijk] - dt * rotEz;
    }
    dU[0][ijk] = As * dU[0][ijk] + dt * (rotHx - Jx);
    dU[1][ijk] = As * dU[1][ijk] + dt * (rotHy - Jy);
    dU[2][ijk] = As * dU[2][ijk] + dt * (rotHz - Jz);
    dU[6][ijk] = As * dU[6][ijk] +
                 dt * (-alpha * divB + m.bu1 * Pxd + m.bu2 * Pyd + m.bu3 * Pzd) -
                 alpha * (ffe_plant == FFE_PLANT_TAU_DT ? dt * P[ijk] : P[ijk]) / tau;
    if (diag) {
      if (diag->J[0]) diag->J[0][ijk] = Jx;
      if (diag->J[1]) diag->J[1][ijk] = Jy;
      if (diag->J[2]) diag->J[2][ijk] = Jz;
      if (diag->divB) diag->divB[ijk] = divB;
      if (diag->divE) diag->divE[ijk] = divD;
    }
  }
}

/* kernel_rk_step2_gr (.cu:375-438) */
static void gr_rk2(const ggrid_t *g, double *const U[7], double *const dU[7], double Bs) {
  INTERIOR(g) {
    long ijk = i + j * g->sy + k * g->sz;
    for (int c = 0; c < 7; c++) U[c][ijk] = U[c][ijk] + Bs * dU[c][ijk];
  }
}

/* kernel_clean_epar_gr then kernel_check_eGTb_gr (.cu:440-535); both are
 * pointwise, so one pass per cell in the reference's order is the same. */
static void gr_project(const ffe_params *p, const ggrid_t *g, double *const U[7]) {
  const double a = p->a;
  INTERIOR(g) {
    long ijk = i + j * g->sy + k * g->sz;
    double x = gpos(g, 0, i), y = gpos(g, 1, j), z = gpos(g, 2, k);
    cks_t m;
    cks_at(a, x, y, z, &m);
    if (ffe_plant == FFE_PLANT_PROJ_FLAT) {
      m.gd11 = m.gd22 = m.gd33 = 1.0;
      m.gd12 = m.gd13 = m.gd23 = 0.0;
    }
    double *Dx = U[0], *Dy = U[1], *Dz = U[2];
    const double *Bx = U[3], *By = U[4], *Bz = U[5];
    if (p->clean_ep) {
      double Bdx = m.gd11 * Bx[ijk] + m.gd12 * By[ijk] + m.gd13 * Bz[ijk];
      double Bdy = m.gd12 * Bx[ijk] + m.gd22 * By[ijk] + m.gd23 * Bz[ijk];
      double Bdz = m.gd13 * Bx[ijk] + m.gd23 * By[ijk] + m.gd33 * Bz[ijk];
      double B2 = Bx[ijk] * Bdx + By[ijk] * Bdy + Bz[ijk] * Bdz;
      if (B2 < TINY_G) B2 = TINY_G;
      double DB = Dx[ijk] * Bdx + Dy[ijk] * Bdy + Dz[ijk] * Bdz;
      Dx[ijk] = Dx[ijk] - DB * Bx[ijk] / B2;
      Dy[ijk] = Dy[ijk] - DB * By[ijk] / B2;
      Dz[ijk] = Dz[ijk] - DB * Bz[ijk] / B2;
    }
    if (p->check_egb) {
      double Ddx = m.gd11 * Dx[ijk] + m.gd12 * Dy[ijk] + m.gd13 * Dz[ijk];
      double Ddy = m.gd12 * Dx[ijk] + m.gd22 * Dy[ijk] + m.gd23 * Dz[ijk];
      double Ddz = m.gd13 * Dx[ijk] + m.gd23 * Dy[ijk] + m.gd33 * Dz[ijk];
      double D2 = Dx[ijk] * Ddx + Dy[ijk] * Ddy + Dz[ijk] * Ddz;
      if (D2 < TINY_G) D2 = TINY_G;
      double Bdx = m.gd11 * Bx[ijk] + m.gd12 * By[ijk] + m.gd13 * Bz[ijk];
      double Bdy = m.gd12 * Bx[ijk] + m.gd22 * By[ijk] + m.gd23 * Bz[ijk];
      double Bdz = m.gd13 * Bx[ijk] + m.gd23 * By[ijk] + m.gd33 * Bz[ijk];
      double B2 = Bx[ijk] * Bdx + By[ijk] * Bdy + Bz[ijk] * Bdz;
      double temp = D2 > B2 ? sqrt(B2 / D2) : 1.0;
      Dx[ijk] = temp * Dx[ijk];
      Dy[ijk] = temp * Dy[ijk];
      Dz[ijk] = temp * Dz[ijk];
    }
  }
}

/* pmlsigma (boundary.cu:7-15) */
static inline double pmlsigma(double x, double xl, double xh, double pmlscale, double sig0) {
  const int sq_only = ffe_plant == FFE_PLANT_SIGMA_SQUARED;
  if (x > xh) return sig0 * (sq_only ? sq((x - xh) / pmlscale) : cube((x - xh) / pmlscale));
  if (x < xl) return sig0 * (sq_only ? sq((xl - x) / pmlscale) : cube((xl - x) / pmlscale));
  return 0.0;
}

/* kernel_boundary_absorbing_thread (boundary.cu:17-79) followed by
 * kernel_absorbing_inner (.cu:610-701), as boundary_absorbing (.cu:810-832)
 * runs them; Un = Dtmp, Btmp (step-start copies). */
static void gr_absorb(const ffe_params *p, const ggrid_t *g, double *const U[7],
                      double *const Un[7]) {
  const double a = p->a, sig0 = p->sigpml, dt = p->dt;
  double lo[3], hi[3];
  for (int d = 0; d < 3; d++) {
    hi[d] = p->lower[d] + p->size[d] - p->pml[d] * g->delta[d];
    lo[d] = p->lower[d] + p->pml[d] * g->delta[d];
  }
  const double scale = p->pmllen * g->delta[0];
  const double rH = 1.0 + sqrt(1.0 - sq(a));
  const double r1 = 0.7 * rH, r2 = 0.01 * rH, dd = 0.2 * rH;
  INTERIOR(g) {
    long ijk = i + j * g->sy + k * g->sz;
    double x = gpos(g, 0, i), y = gpos(g, 1, j), z = gpos(g, 2, k);
    if (x > hi[0] || x < lo[0] || y > hi[1] || y < lo[1] || z > hi[2] || z < lo[2]) {
      double sig = pmlsigma(x, lo[0], hi[0], scale, sig0) + pmlsigma(y, lo[1], hi[1], scale, sig0) +
                   pmlsigma(z, lo[2], hi[2], scale, sig0);
      if (sig > TINY_G) {
        for (int c = 0; c < 6; c++)
          U[c][ijk] = exp(-sig) * Un[c][ijk] + (1.0 - exp(-sig)) / sig * (U[c][ijk] - Un[c][ijk]);
      }
    }
  }
  INTERIOR(g) {
    long ijk = i + j * g->sy + k * g->sz;
    double x = gpos(g, 0, i), y = gpos(g, 1, j), z = gpos(g, 2, k);
    double r = cks_r(a, x, y, z);
    if (r < r1) {
      double sig = sig0 * cube((r1 - r) / dd) * (ffe_plant == FFE_PLANT_EXCISION_NODT ? 1.0 : dt);
      if (sig > TINY_G) {
        for (int c = 0; c < 6; c++)
          U[c][ijk] = exp(-sig) * Un[c][ijk] + (1.0 - exp(-sig)) / sig * (U[c][ijk] - Un[c][ijk]);
      }
    }
    if (r < r2) {
      for (int c = 0; c < 6; c++) U[c][ijk] = 0.0;
    }
  }
}

/* kernel_KO_step1_gr / step2_gr (.cu:537-608): D, B and P. */
static void gr_ko(const ffe_params *p, const ggrid_t *g, double *const U[7], double *const tmp[7]) {
  const long sy = g->sy, sz = g->sz;
  INTERIOR(g) {
    long ijk = i + j * sy + k * sz;
    for (int c = 0; c < 7; c++)
      tmp[c][ijk] = diff6_2(U[c], ijk, 1) + diff6_2(U[c], ijk, sy) + diff6_2(U[c], ijk, sz);
  }
  const double KO_const = -1. / 64;
  const double eps = p->KOeps;
  INTERIOR(g) {
    long ijk = i + j * sy + k * sz;
    for (int c = 0; c < 7; c++) U[c][ijk] -= eps * KO_const * tmp[c][ijk];
  }
}

static void gr_fill(const ffe_params *p, double *const U[7]) {
  if (p->periodic[0] || p->periodic[1] || p->periodic[2]) ffe_flat_fill_guards_f64(p, U);
}

/* field_solver_gr_EZ::evolve_fields (.cu:862-910) on one rank. */
int ffe_gr_evolve_f64(const ffe_params *p, double *const U[7], double *const dU[7],
                      double *const Un[7], double *const EH[6], const ffe_diag_f64 *diag,
                      double time, int nsteps) {
  static const double As[5] = {0, -0.4178904745, -1.192151694643, -1.697784692471,
                               -1.514183444257};
  static const double Bs[5] = {0.1496590219993, 0.3792103129999, 0.8229550293869,
                               0.6994504559488, 0.1530572479681};
  (void)time;
  ggrid_t g;
  make_ggrid(p, &g);
  for (int d = 0; d < 3; d++)
    if (g.guard[d] < 3) return -1;
  const size_t cnt = (size_t)g.dims[0] * g.dims[1] * g.dims[2];
  for (int step = 0; step < nsteps; step++) {
    for (int c = 0; c < 7; c++) memcpy(Un[c], U[c], cnt * sizeof(double));
    for (int s = 0; s < 5; s++) {
      gr_eh(&g, p->a, U, EH);
      gr_rk1(p, &g, U, dU, EH, diag, As[s]);
      gr_rk2(&g, U, dU, Bs[s]);
      gr_project(p, &g, U);
      if (s == 4) gr_absorb(p, &g, U, Un);
      gr_fill(p, U);
    }
    gr_ko(p, &g, U, Un);
    gr_project(p, &g, U);
    gr_fill(p, U);
  }
  return 0;
}

/* One call of rk_step (.cu:753-772) with the given coefficients, for tests. */
int ffe_gr_rk_stage_f64(const ffe_params *p, double *const U[7], double *const dU[7],
                        double *const EH[6], const ffe_diag_f64 *diag, double As, double Bs) {
  ggrid_t g;
  make_ggrid(p, &g);
  gr_eh(&g, p->a, U, EH);
  gr_rk1(p, &g, U, dU, EH, diag, As);
  gr_rk2(&g, U, dU, Bs);
  return 0;
}

/* Ed, Hd of kernel_compute_E_gr / _H_gr, for tests. */
int ffe_gr_eh_f64(const ffe_params *p, double *const U[7], double *const EH[6]) {
  ggrid_t g;
  make_ggrid(p, &g);
  gr_eh(&g, p->a, U, EH);
  return 0;
}

/* The projections alone, for tests. */
int ffe_gr_project_f64(const ffe_params *p, double *const U[7]) {
  ggrid_t g;
  make_ggrid(p, &g);
  gr_project(p, &g, U);
  return 0;
}

/* The outer layer and horizon excision alone, for tests. */
int ffe_gr_absorb_f64(const ffe_params *p, double *const U[7], double *const Un[7]) {
  ggrid_t g;
  make_ggrid(p, &g);
  gr_absorb(p, &g, U, Un);
  return 0;
}

/* Kreiss-Oliger alone (tmp: 7 scratch arrays), for tests. */
int ffe_gr_ko_f64(const ffe_params *p, double *const U[7], double *const tmp[7]) {
  ggrid_t g;
  make_ggrid(p, &g);
  gr_ko(p, &g, U, tmp);
  return 0;
}
