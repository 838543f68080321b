/* CPU oracle for the COFFEE force-free field-update hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP product
 * path is compared against; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py load it.  The product never links or calls it.
 *
 * It restates, in plain C with OpenMP over k, the algorithms of:
 *   flat EZ   src/algorithms/field_solver_EZ.cpp (CPU semantics, use_cuda=0),
 *             src/algorithms/damping_boundary.cpp, src/algorithms/pulsar_impl.hpp,
 *             src/sim_env.cpp:365-438 (single-rank periodic guard fill);
 *   spherical src/algorithms/field_solver_EZ_spherical.cu, metric_sph.h;
 *   GR        src/algorithms/field_solver_gr_EZ.cu, boundary.cu, metric_cks_impl.hpp.
 * Every function cites the reference lines it follows.
 *
 * Pinning status (what checks each function) is listed in the header comment
 * of each .c file.
 *
 * Arrays are Fortran order with guards: idx = i + (j + k*dims1)*dims0
 * (src/data/multi_array_impl.hpp:79-91).  Field order in the 7-array state
 * U is Ex,Ey,Ez,Bx,By,Bz,P.
 */
#ifndef FFE_ORACLE_H
#define FFE_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Subset of Coffee::sim_params (src/sim_params.h:15-124) that the hot path
 * reads, held in double/int; each solver rounds to its Scalar on entry exactly
 * as the reference's parse_config stores params in Scalar. */
typedef struct ffe_params {
  int N[3];
  int guard[3];
  double lower[3];
  double size[3];
  int periodic[3];
  double dt;
  int pml[3];
  int pmllen;
  double sigpml;
  int divB_clean;
  double ch2;
  double tau;
  double KOeps;
  int KO_geometry;
  int use_edotb_damping;
  double damp_gamma;
  int calc_current;
  int clean_ep;
  int check_egb;
  int problem;
  double radius, omega, b0;
  double p1, p2, p3;
  double q11, q12, q13, q22, q23;
  double q_offset_x, q_offset_y, q_offset_z;
  double tp_start, tp_end, rpert1, rpert2, dw0, nT, ph1, ph2, dph, theta0, drpert;
  double tp_start1, tp_end1, rpert11, rpert21, dw1, nT1, ph11, ph21, dph1;
  int pert_type;
  double a;
  double thp1, thp2, thp11, thp21;
  /* SIMD width of the reference CPU build whose tanh-mask positions are
   * reproduced (field_solver_EZ.cpp:222-224 adds the lane offset to x, y and
   * z alike); 1 evaluates the mask at the node itself. */
  int mask_lanes;
} ffe_params;

/* Optional per-cell diagnostics written by the flat solver, as the reference
 * does every stage (field_solver_EZ.cpp:123-125,149-150,198-200) and in its
 * energy accounting (.cpp:447-453,588-623).  Any pointer may be NULL. */
typedef struct ffe_diag_f64 {
  double *J[3];
  double *divE, *divB;
  double *dU_KO, *dU_Epar, *dU_EgtB;
  double *dU_KO_cum, *dU_Epar_cum, *dU_EgtB_cum;
} ffe_diag_f64;

typedef struct ffe_diag_f32 {
  float *J[3];
  float *divE, *divB;
  float *dU_KO, *dU_Epar, *dU_EgtB;
  float *dU_KO_cum, *dU_Epar_cum, *dU_EgtB_cum;
} ffe_diag_f32;

/* ---- planted mistakes ------------------------------------------------------
 * A test sets one of these to switch a single deliberate error into the
 * spherical / Kerr-Schild restatement and shows that its check then fails
 * (mutation check of the test itself).  0 = none, the only value outside such
 * tests; with 0 every result is bitwise what the unmodified restatement
 * computes. */
enum {
  FFE_PLANT_NONE = 0,
  FFE_PLANT_J_SIGN = 1,         /* force-free current with the wrong sign */
  FFE_PLANT_J_NO_SQRTG = 2,     /* rho E x B term without its 1/sqrt(gamma) */
  FFE_PLANT_GRADP_LOWER = 3,    /* grad P raised with gamma_ij instead of gamma^ij */
  FFE_PLANT_TAU_DT = 4,         /* P damping rate P/tau taken per unit time, not per step */
  FFE_PLANT_SIGMA_SQUARED = 5,  /* absorbing-layer sigma ~ depth^2 instead of depth^3 */
  FFE_PLANT_PROJ_FLAT = 6,      /* E.B and |E|,|B| of the projections without the metric */
  FFE_PLANT_ROT_NO_SQRTG = 7,   /* curl without its 1/sqrt(gamma) */
  FFE_PLANT_DIV_CENTRE = 8,     /* divergence with sqrt(gamma) of the centre, not the taps */
  FFE_PLANT_STAR_SIGN = 9,      /* corotation field E = +(Omega x r) x B inside the star */
  FFE_PLANT_CLEAN_SHAPE = 10,   /* Kerr-Schild cleaning weight 1/2(1 + tanh) */
  FFE_PLANT_EXCISION_NODT = 11, /* horizon excision sigma without its factor dt */
  FFE_PLANT_COUNT = 12
};
void ffe_set_plant(int which);
int ffe_get_plant(void);
extern int ffe_plant;

/* ---- flat Cartesian EZ (field_solver_EZ.cpp) ---------------------------- */

/* Advance nsteps steps from `time`.  U: 7 state arrays; dU: 7 LSRK registers
 * (dE, dB, dP); Un: 7 step-start copies (Etmp, Btmp, Ptmp; also KO scratch);
 * Bbg: 3 absorbing-layer background arrays (may be NULL when pml==0). */
int ffe_flat_evolve_f64(const ffe_params *p, double *const U[7], double *const dU[7],
                        double *const Un[7], double *const Bbg[3], const ffe_diag_f64 *diag,
                        double time, int nsteps);
int ffe_flat_evolve_f32(const ffe_params *p, float *const U[7], float *const dU[7],
                        float *const Un[7], float *const Bbg[3], const ffe_diag_f32 *diag,
                        double time, int nsteps);

/* Building blocks exposed for unit tests (single calls, interior only). */
int ffe_flat_rk_stage_f64(const ffe_params *p, double *const U[7], double *const dU[7],
                          const ffe_diag_f64 *diag, double As, double Bs);
int ffe_flat_ko_f64(const ffe_params *p, double *const U[7], double *const tmp[7],
                    const ffe_diag_f64 *diag);
int ffe_flat_project_f64(const ffe_params *p, double *const U[7], const ffe_diag_f64 *diag);
int ffe_flat_fill_guards_f64(const ffe_params *p, double *const U[7]);
int ffe_flat_absorb_f64(const ffe_params *p, double *const U[7], double *const Un[7],
                        double *const Bbg[3]);
int ffe_flat_boundary_f64(const ffe_params *p, double *const U[7], double t);
/* Bbg / initial dipole: b0 * quadru_dipole(x,y,z,...,phase=0,n) on collocated
 * nodes including guards (field_solver_EZ.cpp:66-85). */
int ffe_flat_bbg_f64(const ffe_params *p, double *const Bbg[3]);
int ffe_flat_bbg_f32(const ffe_params *p, float *const Bbg[3]);

/* ---- log-spherical 2D EZ (field_solver_EZ_spherical.cu) ------------------ */

/* Arrays are dims0 x dims1 (x = ln r fastest, then theta); p->guard[1] must
 * be >= 4.  L: 6 scratch arrays for the lowered El, Bl.  Un: step-start copy
 * and KO scratch. */
int ffe_sph_evolve_f64(const ffe_params *p, double *const U[7], double *const dU[7],
                       double *const Un[7], double *const L[6], double time, int nsteps);
int ffe_sph_evolve_f32(const ffe_params *p, float *const U[7], float *const dU[7],
                       float *const Un[7], float *const L[6], double time, int nsteps);
double ffe_sph_energy_f64(const ffe_params *p, double *const F[3]);
double ffe_sph_energy_f32(const ffe_params *p, float *const F[3]);
int ffe_sph_axis_f64(const ffe_params *p, double *const U[7]);
int ffe_sph_ko_f64(const ffe_params *p, double *const U[7], double *const tmp[7]);
int ffe_sph_rk_stage_f64(const ffe_params *p, double *const U[7], double *const dU[7],
                         double *const L[6], double As, double Bs);
/* kernel_Epar_sph + kernel_EgtB_sph, kernel_boundary_pulsar_sph at time t,
 * kernel_boundary_absorbing_sph (Un: step-start copy), interior cells. */
int ffe_sph_project_f64(const ffe_params *p, double *const U[7]);
int ffe_sph_boundary_f64(const ffe_params *p, double *const U[7], double t);
int ffe_sph_absorb_f64(const ffe_params *p, double *const U[7], double *const Un[7]);

/* ---- Kerr-Schild GR EZ (field_solver_gr_EZ.cu), FP64 ------------------- */

/* U = D (the evolved "E" arrays of sim_data), B, P; dU = dD, dB, dP;
 * Un = Dtmp, Btmp, Ptmp (step-start copies, then KO scratch);
 * EH = Ed, Hd scratch (6 arrays).  `time` is unused, as in the reference. */
int ffe_gr_evolve_f64(const ffe_params *p, double *const U[7], double *const dU[7],
                      double *const Un[7], double *const EH[6], const ffe_diag_f64 *diag,
                      double time, int nsteps);
int ffe_gr_rk_stage_f64(const ffe_params *p, double *const U[7], double *const dU[7],
                        double *const EH[6], const ffe_diag_f64 *diag, double As, double Bs);
int ffe_gr_eh_f64(const ffe_params *p, double *const U[7], double *const EH[6]);
/* kernel_clean_epar_gr + kernel_check_eGTb_gr; the outer layer and horizon
 * excision of boundary_absorbing (Un = Dtmp, Btmp), interior cells. */
int ffe_gr_project_f64(const ffe_params *p, double *const U[7]);
int ffe_gr_absorb_f64(const ffe_params *p, double *const U[7], double *const Un[7]);
int ffe_gr_ko_f64(const ffe_params *p, double *const U[7], double *const tmp[7]);
/* gd11 gd12 gd13 gd22 gd23 gd33, gu11 .. gu33, bd1..3, bu1..3, alpha,
 * sqrt(gamma), r at one point (metric_cks_impl.hpp). */
void ffe_gr_metric_f64(double a, double x, double y, double z, double out[21]);

/* Closed-form helpers (src/algorithms/pulsar_impl.hpp) in double. */
double ffe_dipole2(double x, double y, double z, double p1, double p2, double p3,
                   double phase, int n);
double ffe_quadru_dipole(const ffe_params *p, double x, double y, double z, double phase,
                         int n);
double ffe_diff1_4(const double *f, long ijk, long s);
double ffe_diff6_2(const double *f, long ijk, long s);

#ifdef __cplusplus
}
#endif

#endif
