/* coffee_hip.h — C ABI of the MI355X (gfx950) force-free field-update hot path.
 *
 * Plain C: pointers, sizes and PODs only.  Every entry point is asynchronous
 * on the hipStream_t passed as `stream` (a `void*`, NULL = the default
 * stream) unless stated otherwise, returns 0 on success and a negative code on
 * failure (never throws, never exits), and records a message retrievable with
 * coffee_last_error().  Arrays are device pointers owned by the caller, in
 * Fortran order with guard cells: idx = i + (j + k*dims[1])*dims[0]
 * (src/data/multi_array_impl.hpp:79-91).  A context owns only its scratch
 * (the second buffer of the ping-pong state, timing events, reduction
 * buffers); it never frees caller memory.
 *
 * Each declaration names the reference interface it replaces.
 */
#ifndef COFFEE_HIP_H
#define COFFEE_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define COFFEE_ABI_VERSION 1

/* Status codes. */
#define COFFEE_OK 0
#define COFFEE_ERR_INVALID (-1)  /* bad argument or configuration */
#define COFFEE_ERR_HIP (-2)      /* HIP runtime error (message has the hipError string) */
#define COFFEE_ERR_SIZE (-3)     /* grid too large for this build's index type */
#define COFFEE_ERR_STATE (-4)    /* call not valid for this context's solver */

/* Solver variants (the three field_solver classes of the reference). */
#define COFFEE_SOLVER_EZ 0           /* field_solver_EZ (src/algorithms/field_solver_EZ.h:6-33) */
#define COFFEE_SOLVER_EZ_SPHERICAL 1 /* field_solver_EZ_spherical (field_solver_EZ_spherical.h:6-32) */
#define COFFEE_SOLVER_GR_EZ 2        /* field_solver_gr_EZ (field_solver_gr_EZ.h:6-32) */

/* Scalar type of the arrays. */
#define COFFEE_F32 4
#define COFFEE_F64 8

/* Grid geometry: Coffee::Grid (src/data/grid.h:11-22), held in double. */
typedef struct coffee_grid {
  int dims[3];      /* N + 2*guard per axis (1 for an absent axis) */
  int guard[3];
  double delta[3];
  double lower[3];
  double sizes[3];
} coffee_grid;

/* Hot-path subset of Coffee::sim_params (src/sim_params.h:15-124), same
 * names and meaning; dt and tau as in the reference (tau in units of dt). */
typedef struct coffee_params {
  double dt;
  int periodic[3];
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
  double a; /* black-hole spin (GR) */
  double thp1, thp2, thp11, thp21; /* spherical twist pulses (wpert_sph_angle) */
  double lower[3];
  double size[3];
  /* Flat solver: evaluate the tanh mask of the cleaning term at the
   * positions the reference CPU build uses (its SIMD lane offset is added to
   * x, y and z alike, field_solver_EZ.cpp:222-224).  Set to the reference
   * build's vector width (16 for FP32, 8 for FP64 on AVX-512) for parity with
   * it; 1 evaluates the mask at the node. */
  int mask_lanes;
  /* sim_params::shift_ghost: guard layers the reference's flat CPU loops and
   * GR kernels also update (field_solver_EZ.cpp:128-136).  Only 0 is
   * supported; other values are rejected with COFFEE_ERR_INVALID. */
  int shift_ghost;
} coffee_params;

/* Device arrays of one solver call: the reference's sim_data members and
 * the solver's public scratch members, all caller-owned.  Unused entries
 * may be NULL (see each entry point).  For the GR solver E/B hold D/B. */
typedef struct coffee_fields {
  void *E[3];      /* sim_data::E */
  void *B[3];      /* sim_data::B */
  void *P;         /* sim_data::P */
  void *dE[3];     /* solver dE  (LSRK register) */
  void *dB[3];     /* solver dB */
  void *dP;        /* solver dP */
  void *Etmp[3];   /* solver Etmp: step-start E in the absorbing layer */
  void *Btmp[3];   /* solver Btmp */
  void *Ptmp;      /* solver Ptmp */
  void *Bbg[3];    /* solver Bbg: layer background (flat CPU semantics) */
  void *J[3];      /* sim_data::B0 (holds J) — written only when diagnostics are on */
  void *divE;      /* sim_data::divE — diagnostics */
  void *divB;      /* sim_data::divB — diagnostics */
  /* sim_data energy accounting (field_solver_EZ.cpp:447-453,588-623), flat
   * solver, diagnostics mode only, each optional: per step dU_KO (energy
   * change of E, B by Kreiss-Oliger), dU_Epar, dU_EgtB (summed over the step's
   * six projections; dU_EgtB counted from the pre-projection energy, so it
   * includes dU_Epar, as the reference counts it), and the _cum arrays
   * accumulating over all steps (never cleared by the solver). */
  void *dU_KO, *dU_Epar, *dU_EgtB;
  void *dU_KO_cum, *dU_Epar_cum, *dU_EgtB_cum;
} coffee_fields;

typedef struct coffee_ctx coffee_ctx;

/* Create a solver context.  Replaces the constructors
 * field_solver_EZ(sim_data&, sim_environment&) (field_solver_EZ.cpp:50-101,
 * field_solver_EZ.cu:709-751) and the constant-memory upload of sim_params /
 * Grid (src/cuda/constant_mem.cu:7-21, src/cuda/sim_env.cpp:250-268).
 * Allocates the context's scratch on the current HIP device. */
int coffee_ctx_create(int solver, int scalar_bytes, const coffee_grid *grid,
                      const coffee_params *params, coffee_ctx **out);

/* Destroy a context and free its scratch (never caller arrays). */
int coffee_ctx_destroy(coffee_ctx *ctx);

/* Replace the parameters (e.g. dt, omega) between steps: the reference's
 * init_dev_params (src/cuda/constant_mem_func.h). Grid is fixed. */
int coffee_set_params(coffee_ctx *ctx, const coffee_params *params);

/* One full time step: 5 LSRK(5,4) stages (stencil, current, cleaning, dU
 * and U update, E.B / E<B projection, problem boundary, stage-5 absorbing
 * layer), Kreiss-Oliger, projection, boundary and the single-GPU guard fill.
 * Replaces field_solver_EZ::evolve_fields(Scalar time)
 * (field_solver_EZ.cpp:1007-1073, field_solver_EZ.cu:656-707) together with
 * the sim_environment::send_guard_cells calls it makes
 * (src/sim_env_impl.hpp:179-195).  On return (stream order) the state is back
 * in the caller's E, B, P.  Needs E, B, P, dE, dB, dP; Etmp/Btmp/Ptmp and Bbg
 * when an absorbing layer is active. */
int coffee_ez_step(coffee_ctx *ctx, const coffee_fields *f, double time, void *stream);

/* Largest partner state coffee_ez_step_noalloc accepts: 8 arrays (the
 * second buffer of the 7 state arrays and a third one of P) of
 * dims[0]*dims[1]*dims[2] scalars, each rounded up to 256 bytes (80 MiB: a
 * 128^3 FP32 grid with guard 3 needs 77.0 MB; FP64 fits up to about
 * 1.31e6 cells including guards, e.g. 103^3 interior with guard 3). */
#define COFFEE_NOALLOC_SCRATCH_BYTES (80ull << 20)

/* No counterpart in the reference.  coffee_ez_step without a context and
 * without any allocation, for callers that cannot call hipMalloc (code under
 * HIP graph capture, tight serving loops, harnesses that forbid allocation):
 * the same kernels and launch sequence as coffee_ez_step on a fresh flat
 * context (default schedule, diagnostics and timing off), so the results are
 * bitwise those of coffee_ez_step.  The ping-pong partner state lives in one
 * static device array of COFFEE_NOALLOC_SCRATCH_BYTES per GPU, held until the
 * process ends; grids needing more return COFFEE_ERR_SIZE.  Nothing in it
 * has to start at zero (every step writes the partner's interior and copies
 * its frozen guards first).  All calls on one GPU share that array: calls
 * that overlap on one GPU (two streams or two threads) race and must be
 * serialised by the caller.  Arguments as coffee_ctx_create + coffee_ez_step
 * with solver COFFEE_SOLVER_EZ. */
int coffee_ez_step_noalloc(const coffee_grid *grid, const coffee_params *params,
                           int scalar_bytes, const coffee_fields *f, double time,
                           void *stream);

/* Guard-cell fill on one GPU: periodic axes take the interior one period
 * away (x, then y, then z); other axes are left untouched (frozen), as
 * send_guard_cells does on a single rank (src/sim_env.cpp:365-438). */
int coffee_fill_guards(coffee_ctx *ctx, const coffee_fields *f, void *stream);

/* Sum over interior cells of f0^2+f1^2+f2^2, excluding the absorbing band
 * and (problem 1) r < radius: field_solver_EZ::total_energy
 * (field_solver_EZ.cu:755-795).  Synchronous on `stream`; result in *out. */
int coffee_total_energy(coffee_ctx *ctx, void *const comp[3], double *out, void *stream);

/* Residual monitor over interior cells (no reference counterpart; the
 * reference writes per-cell divE/divB/J arrays each stage instead,
 * field_solver_EZ.cpp:123-125, field_solver_EZ_spherical.cu:325-329,
 * field_solver_gr_EZ.cu:336-338): out[0] = field energy, out[1] = max |div B|,
 * out[2] = max |E.B|/max(B^2,TINY), out[3] = max (E^2-B^2)/max(B^2,TINY)
 * (> 0 means E > B somewhere), out[4] = number of cells with a non-finite
 * E, B or P.  Flat: plain sums, sum(E^2+B^2).  Spherical and Kerr-Schild:
 * dot products with the 3-metric (E_i B^i, ...), div B = d_i(sqrt(gamma) B^i)/
 * sqrt(gamma) with the solver's 4th-order stencil, energy
 * sum (E_i E^i + B_i B^i) sqrt(gamma) (times dx dtheta in the spherical case;
 * axis cells, where sqrt(gamma) = 0, are left out of div B); for the GR
 * solver E means D.  A spherical batch uses coffee_sph_batch_residuals.
 * Synchronous on `stream`. */
int coffee_residuals(coffee_ctx *ctx, const coffee_fields *f, double out[5], void *stream);

/* Diagnostics mode: when on, each stage also writes J (into f->J), divE and
 * divB exactly as the reference does every stage (field_solver_EZ.cpp:123-125,
 * 149-150,198-200); off (default) in timed runs. */
int coffee_set_diagnostics(coffee_ctx *ctx, int enable);

/* Kernel schedule of the solver (no reference counterpart: the reference
 * has one thread-per-cell launch geometry, field_solver_EZ.cu:10-12).
 * variant 0 = LDS-tiled z-marching kernels with the default tile for the
 * Scalar type (FP32 128 x 8, FP64 64 x 8), 1 = one thread per cell, 2/3/4 =
 * tiled 64 wide with 4/8/16 rows per tile (16 in FP32, 8 in FP64), 5/6 = tiled
 * 128/256 wide (FP32 128 x 8 and 256 x 4; FP64 128 x 4 for both), 7 = 128 x 4;
 * variant 0 in FP32 runs Kreiss-Oliger with 32-bit field offsets when every
 * array is below 4 GiB, variant 5 is the same tile with 64-bit addressing;
 * the GR solver runs its tiled schedule, every metric point computed on the
 * fly, for every variant but 1 (point sweep + stage sweep) and 2 (the halo
 * points' metric in stages 2-5 read from a table filled once); in FP32 its
 * middle stages load the x-y halo three planes ahead except in variant 3
 * (one plane ahead); variants 1
 * and 2 allocate 12 grid-sized scratch arrays on first use, and GR arrays of
 * 4 GiB or more always take the point schedule; zchunk = z planes per
 * workgroup of the tiled kernels, 0 = automatic.
 * Schedules change speed only: results are bitwise identical. */
int coffee_set_schedule(coffee_ctx *ctx, int variant, int zchunk);

/* Per-kernel timing with HIP events recorded on the launch stream.  When
 * enabled, every launch is bracketed by events; coffee_get_timing returns,
 * per kernel name, the number of launches and the summed milliseconds since
 * the last reset (synchronises the stream of the last launch). */
int coffee_set_timing(coffee_ctx *ctx, int enable);
int coffee_get_timing(coffee_ctx *ctx, int max_entries, char (*names)[64], double *total_ms,
                      int *launches, int *n_entries);
int coffee_reset_timing(coffee_ctx *ctx);

/* ---- batched log-spherical runs (no reference counterpart) --------------
 * Many independent axisymmetric runs (each its own grid, parameters and
 * caller-owned E, B, P) advance together: one launch per substep covers a
 * work list of (run, tile) items; each run keeps its time and step counter on
 * the device and stops advancing after nsteps[r] steps (0 = never stops).
 * dE/dB/dP and the absorbing-layer copies are allocated by the batch when
 * the corresponding fields[r] pointers are NULL.  Results of a run in a
 * batch are bitwise those of the same run in its own context. */
typedef struct coffee_ctx coffee_batch;
int coffee_sph_batch_create(int scalar_bytes, int nruns, const coffee_grid *grids,
                            const coffee_params *params, const coffee_fields *fields,
                            const int *nsteps, const double *t0, coffee_batch **out);
int coffee_sph_batch_destroy(coffee_batch *b);
/* Advance every active run by `nsteps` steps (asynchronous on `stream`). */
int coffee_sph_batch_step(coffee_batch *b, int nsteps, void *stream);
/* Tile shapes the kernels are built for: coffee_sph_num_tiles() of them.
 * Tile i has tx x ty threads, one cell each, or, for the theta-marching
 * tiles, tx threads that walk a band of tx radial cells over ty theta rows.
 * Every tile gives bitwise the same results. */
int coffee_sph_num_tiles(void);
int coffee_sph_tile_shape(int tile_index, int *tx, int *ty);
int coffee_sph_batch_set_tile(coffee_batch *b, int tile_index);
/* Time every tile shape on the live batch (first-stage kernel, `trials`
 * launches each; state is not advanced), keep the fastest (ties within 3%
 * go to the lower index); ms_per_tile may be NULL. */
int coffee_sph_batch_autotune(coffee_batch *b, int trials, void *stream, int *best_tile,
                              double *ms_per_tile);
/* Capture one step's launch sequence into a HIP graph and replay it. */
int coffee_sph_batch_use_graph(coffee_batch *b, int enable);
/* Residual monitor of every run (the quantities of coffee_residuals for the
 * spherical solver), out[5*run .. 5*run+4].  Synchronous on `stream`. */
int coffee_sph_batch_residuals(coffee_batch *b, double *out, void *stream);
/* Per-run active flag, time and step count (synchronises the device). */
int coffee_sph_batch_status(coffee_batch *b, int *active, double *t, int *step);
/* Spin-down luminosity of every run through the sphere r = radius[run]:
 * L = 2 pi r^2 \int (E x B)_r sin(theta) dtheta (orthonormal components, the
 * units of the fields; interpolated linearly in ln r).  No counterpart in the
 * reference; synchronous on `stream`, out[run] in double. */
int coffee_sph_batch_luminosity(coffee_batch *b, const double *radius, double *out, void *stream);

/* Fill n scalars at device pointer p with `value` (asynchronous).  Replaces
 * multi_array::assign_dev and its assign_single_value kernel
 * (src/cuda/data/multi_array.cu:12-19,119-128). */
int coffee_fill(void *p, unsigned long long n, int scalar_bytes, double value, void *stream);

/* Guard fill of `narrays` arrays of one grid without a solver context: the
 * same operation as coffee_fill_guards, for sim_environment::send_guard_cells
 * / send_guard_cell_array callers (src/sim_env_impl.hpp:165-195). */
int coffee_fill_guards_arrays(const coffee_grid *grid, const int periodic[3], void *const *arrays,
                              int narrays, int scalar_bytes, void *stream);

/* Streaming copy of `bytes` (a multiple of 16, 16-byte aligned pointers)
 * from src to dst, 16 bytes per lane, `unroll` (1, 2, 4, 8) loads in flight
 * per lane, `blocks` workgroups of 256 lanes, each copying one contiguous
 * chunk with nontemporal loads and stores (asynchronous).  No reference
 * counterpart: it measures the HBM bandwidth this device attains on a plain
 * read + write stream, the ceiling the stencil kernels are compared with. */
int coffee_stream_copy(const void *src, void *dst, unsigned long long bytes, int unroll,
                       int blocks, void *stream);

/* Last error message of this context (or of the last failed create when ctx
 * is NULL). */
const char *coffee_last_error(const coffee_ctx *ctx);

/* Library ABI version (COFFEE_ABI_VERSION). */
int coffee_abi_version(void);

#ifdef __cplusplus
}
#endif

#endif /* COFFEE_HIP_H */
