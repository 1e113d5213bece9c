/*
 * ygpu.h -- C-ABI of the MI355X SSA/DIVA velocity solver (libygpu.so).
 *
 * Drop-in boundary for Yelmo's velocity-solver call chain.  Two levels:
 *
 *   inner  ygpu_csr_solve      replaces linear_solver_matrix_solve
 *                              (src/physics/solver_linear.F90:151-190, the LIS
 *                              call at :194-400): a CSR system as the
 *                              reference assembles it, solved on the GPU.
 *   outer  ygpu_velocity_ssa   replaces calc_velocity_ssa
 *                              (src/physics/velocity_ssa.f90:60-357): the whole
 *                              Picard loop resident on the device.
 *          ygpu_velocity_diva  replaces calc_velocity_diva
 *                              (src/physics/velocity_diva.f90:59-429).
 *          ygpu_velocity_ssa_batch / ygpu_velocity_diva_batch
 *                              many independent domains in one call.
 *
 * Conventions follow the reference's C API (docs/c-api.md:23-38): plain
 * pointers, Fortran column-major (nx,ny[,nz]) arrays with i fastest, 3D
 * arrays with k slowest and k = 0 at the base.  The caller owns every host
 * array; the library copies in and out and keeps no host pointer after a call
 * returns.  Calls never abort: they return YGPU_OK, a negative error code
 * (text from ygpu_last_error), and report solver outcomes through LIS-style
 * status values (0 converged, 2 breakdown, 4 iteration limit).
 * One handle per host thread; handles share nothing.
 */
#ifndef YGPU_H
#define YGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* return codes */
#define YGPU_OK 0
#define YGPU_ERR_ARG (-1)         /* invalid argument / option string          */
#define YGPU_ERR_DEVICE (-2)      /* HIP runtime failure                       */
#define YGPU_ERR_UNSUPPORTED (-3) /* matrix pattern / option not supported     */
#define YGPU_ERR_NONFINITE (-6)   /* non-finite velocity (the reference's
                                     error stop, velocity_general.f90:2114-2152) */

/* boundaries strings (src/yelmo_tools.f90:248-271) */
#define YGPU_BND_ZEROS 0
#define YGPU_BND_INFINITE 1
#define YGPU_BND_MASK 2
#define YGPU_BND_MISMIP3D 3
#define YGPU_BND_TROUGH 4
#define YGPU_BND_PERIODIC 5
#define YGPU_BND_PERIODIC_X 6
#define YGPU_BND_PERIODIC_Y 7

/* arithmetic of the outer solve */
#define YGPU_PREC_FP64 0  /* FP64 fields, matrix and Krylov (plain FP64 path)          */
#define YGPU_PREC_REF 1   /* FP32 fields, FP64 matrix and Krylov (the reference, wp=sp) */
#define YGPU_PREC_MIXED 2 /* FP32 fields (FP64 with -ygpu_fields fp64), FP64 matrix
                             records and residual; FP32 SpMV/ILU inside FP64 refinement */

typedef struct ygpu_handle ygpu_handle;

/* The optional neff_hook of calc_velocity_diva (velocity_diva.f90:99-107):
 * called on the host after every Picard iteration, once the iteration's
 * basal velocity (ux_b, uy_b: (nx,ny) arrays of real_bytes each, the
 * caller's own arrays) is known; it updates c_bed in place (the same real
 * type).  ctx is the hook_ctx the caller registered.                        */
typedef void (*ygpu_neff_hook)(void* ctx, int32_t nx, int32_t ny, int32_t real_bytes, void* c_bed,
                               const void* ux_b, const void* uy_b);

/* Mirrors ssa_param_class (velocity_ssa.f90:22-52) and diva_param_class
 * (velocity_diva.f90:22-51); strings become the enums above.             */
typedef struct ygpu_ssa_params {
  int32_t ssa_solver;     /* 0 "residual" (solver_ssa_ac.f90), 1 "energy"
                             (solver_ssa_ac_energy.f90, the reference default)   */
  int32_t boundaries;     /* YGPU_BND_*                                          */
  int32_t visc_method;    /* 0, 1 (quadrature), 2 (aa-nodes)                     */
  int32_t beta_method;    /* -1 .. 5                                             */
  int32_t beta_gl_scale;  /* 0 .. 3 (2: zstar scaling, needs z_bed, z_sl)      */
  int32_t beta_gl_stag;   /* -1 .. 4 (3, 4: subgrid, need f_grnd_acx/acy)       */
  int32_t ssa_vel_lim_method; /* 0 "clip", 1 "drag"                              */
  int32_t ssa_iter_max;
  int32_t ssa_write_log;
  int32_t precision;      /* YGPU_PREC_*                                         */
  double visc_const, beta_const, beta_q, beta_u0, beta_gl_f, H_grnd_lim, beta_min, eps_0,
      ssa_vel_lim_tau, ssa_vel_max, ssa_iter_rel, ssa_iter_conv, rho_ice, rho_sw, g;
  int32_t no_slip;        /* DIVA: "diva-noslip"                                 */
  int32_t block;          /* threads per member block (0: pick automatically)   */
  char ssa_lis_opt[256];  /* LIS option string, e.g. the energy default
                             "-i cg -p jacobi -maxiter 200 -tol 1.0e-2 -initx_zeros false";
                             -i cg | bicgstab; minres and bicgsafe are accepted
                             and solved with bicgstab to the same tolerance,
                             and ygpu_last_note says so;
                             extensions: -ir_inner_tol T, -ir_max N              */
} ygpu_ssa_params;

/* One domain of a batch: the arguments of calc_velocity_ssa (and the extra
 * ones of calc_velocity_diva).  real_bytes = 4 for float arrays (Yelmo's
 * wp = sp), 8 for double arrays.  Unused DIVA pointers may be NULL for SSA. */
typedef struct ygpu_ssa_problem {
  int32_t nx, ny, nz, real_bytes;
  double dx, dy, n_glen;
  /* INOUT */
  void *ux_b, *uy_b, *taub_acx, *taub_acy, *visc_eff, *visc_eff_int;
  int32_t *ssa_mask_acx, *ssa_mask_acy;
  void *ssa_err_acx, *ssa_err_acy;
  void *beta, *beta_acx, *beta_acy;
  /* IN */
  const void *c_bed, *f_slide, *taud_acx, *taud_acy, *taul_int_acx, *taul_int_acy, *H_ice, *f_ice,
      *H_grnd, *f_grnd, *f_grnd_acx, *f_grnd_acy, *ATT, *zeta_aa, *z_sl, *z_bed, *z_srf;
  /* DIVA INOUT (NULL for SSA) */
  void *ux, *uy, *ux_bar, *uy_bar, *ux_i, *uy_i, *beta_eff, *duxdz, *duydz;
  /* OUT */
  int32_t ssa_iter_now, ssa_lin_iter, ssa_lin_fail, ssa_lim_n;
  int32_t status;         /* YGPU_OK or YGPU_ERR_NONFINITE for this domain       */
  int32_t pad;
  double picard_resid;    /* last Picard residual                                 */
  double solve_seconds;   /* device time of the whole batch call                   */
  /* DIVA only (NULL otherwise): neff_hook and its context.  With a hook,
   * c_bed is INOUT (the hook's last update is left in it), Picard
   * iterations run one launch sequence at a time, and an iteration counts
   * as converged only if the hook changed c_bed by at most ssa_iter_conv in
   * relative L1 (velocity_diva.f90:396-400).                                  */
  ygpu_neff_hook neff_hook;
  void* hook_ctx;
} ygpu_ssa_problem;

/* ---- handle ------------------------------------------------------------ */
int ygpu_create(int32_t device, ygpu_handle** out);
int ygpu_destroy(ygpu_handle* h);
const char* ygpu_last_error(const ygpu_handle* h);
/* What the last call that read an option string did differently from what
 * it was asked, or "" (e.g. "-i minres solved with BiCGSTAB").  Calls that
 * take options (one-shot solves, batch/session/server creation, the CSR
 * entry points, the autotuner) reset it.                                     */
const char* ygpu_last_note(const ygpu_handle* h);
/* fills params with Yelmo's &ydyn defaults (input/yelmo_defaults.nml:140-173)
 * and input/yelmo_phys_const.nml:6-10 (no handle needed)                     */
void ygpu_default_params(ygpu_ssa_params* par);

/* ---- inner boundary: replaces linear_solver_matrix_solve
 * (solver_linear.F90:151).  n rows, CSR arrays 1-based as in lgs%a_ptr /
 * a_index / a_value; x_inout is x0 on entry (unless -initx_zeros true) and the
 * solution on return.  OUT: iters, status (0/2/4), rel_resid
 * (||b-Ax||/||r0||, LIS conv_cond 0), seconds (device solve time).
 * Supported patterns: the C-grid SSA stencil of linear_solver_init's ordering
 * (solver_linear.F90:134-143), non-periodic.  Option grammar: LIS's -i, -p,
 * -maxiter, -tol, -initx_zeros, -conv_cond; extensions -ygpu_ny NY (grid hint),
 * -ygpu_precision fp64|mixed.                                                */
/* Extensions of the option string read by the outer entry points too:
 *   -ygpu_factor bf16|bf16u|fp32  storage of the mixed path's ILU(0)
 *                            factors: BF16 with U rows scaled by their FP32
 *                            inverse diagonals; bf16u: the U part in BF16
 *                            and the L part in FP32; default fp32
 *   -ygpu_ilu_blocks N       ILU(0) over N line blocks (mixed path default:
 *                            one per 64 lines, or the autotuned schedule)
 *   -ygpu_ir_tol T           residual at which refinement stops (tol/10)
 *   -ygpu_fields fp32|fp64   arithmetic of the mixed mode's field kernels
 *                            (viscosity, friction, assembly, Picard update):
 *                            FP32 by default, as the reference's wp = sp;
 *                            the Picard norm accumulates in FP64 either way
 *   -ir_inner_tol T          relative tolerance of each FP32 inner solve
 *   -ygpu_graph N            Picard iterations per captured hipGraph (4; 0 =
 *                            launch each iteration and read the flags after
 *                            it; runs with kernel statistics on are eager)   */
int ygpu_csr_solve(ygpu_handle* h, int32_t n, int32_t nnz, const int32_t* a_ptr1,
                   const int32_t* a_idx1, const double* a_val, const double* b, double* x_inout,
                   const char* opts, int32_t* iters, int32_t* status, double* rel_resid,
                   double* seconds);

/* ---- ygpu_csr_solve without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * Arguments, results and errors as ygpu_csr_solve (seconds: host wall time
 * of the Krylov launch).  Every device buffer comes from one static device
 * array of YGPU_NOALLOC_SCRATCH_BYTES per GPU, held from library load until
 * the process ends; systems above YGPU_NOALLOC_MAX_CELLS cells (n/2), or
 * with more than 9 entries per row, are rejected with YGPU_ERR_ARG.  All
 * calls on one GPU share that array (each clears what it needs): calls that
 * overlap on one GPU, from several threads or handles, race.               */
#define YGPU_NOALLOC_MAX_CELLS 32768
#define YGPU_NOALLOC_SCRATCH_BYTES (32 * 1024 * 1024)
int ygpu_csr_solve_noalloc(ygpu_handle* h, int32_t n, int32_t nnz, const int32_t* a_ptr1,
                           const int32_t* a_idx1, const double* a_val, const double* b,
                           double* x_inout, const char* opts, int32_t* iters, int32_t* status,
                           double* rel_resid, double* seconds);

/* y = A x for a CSR matrix of the same pattern (LIS's lis_matvec at the
 * boundary of solver_linear.F90; 1-based arrays as in lgs), on the device.  */
int ygpu_csr_spmv(ygpu_handle* h, int32_t n, int32_t nnz, const int32_t* a_ptr1,
                  const int32_t* a_idx1, const double* a_val, const double* x, double* y,
                  const char* opts);

/* ---- outer boundary: calc_velocity_ssa (velocity_ssa.f90:60-63), argument
 * order as in the reference, wp = sp arrays.                                  */
int ygpu_velocity_ssa(ygpu_handle* h, const ygpu_ssa_params* par, int32_t nx, int32_t ny,
                      int32_t nz, float* ux_b, float* uy_b, float* taub_acx, float* taub_acy,
                      float* visc_eff, float* visc_eff_int, int32_t* ssa_mask_acx,
                      int32_t* ssa_mask_acy, float* ssa_err_acx, float* ssa_err_acy,
                      int32_t* ssa_iter_now, int32_t* ssa_lin_iter, int32_t* ssa_lin_fail,
                      int32_t* ssa_lim_n, float* beta, float* beta_acx, float* beta_acy,
                      const float* c_bed, const float* f_slide, const float* taud_acx,
                      const float* taud_acy, const float* taul_int_acx,
                      const float* taul_int_acy, const float* H_ice, const float* f_ice,
                      const float* H_grnd, const float* f_grnd, const float* f_grnd_acx,
                      const float* f_grnd_acy, const float* ATT, const float* zeta_aa,
                      const float* z_sl, const float* z_bed, const float* z_srf, float dx,
                      float dy, float n_glen);

/* ---- many independent domains (ensembles, uneven batches) in one call.
 * Every domain runs the full Picard loop; per-domain outcomes are written into
 * probs[k] (ssa_iter_now ... status).  Returns YGPU_OK unless an argument or
 * the device failed.                                                          */
int ygpu_velocity_ssa_batch(ygpu_handle* h, const ygpu_ssa_params* par, int32_t nprob,
                            ygpu_ssa_problem* probs);

/* ---- outer boundary: calc_velocity_diva (velocity_diva.f90:59-64),
 * argument order as in the reference, wp = sp arrays.  c_bed is INOUT as in
 * the reference (the hook updates it); neff_hook may be NULL (the optional
 * argument absent).                                                         */
int ygpu_velocity_diva(ygpu_handle* h, const ygpu_ssa_params* par, int32_t nx, int32_t ny,
                       int32_t nz, float* ux, float* uy, float* ux_bar, float* uy_bar, float* ux_b,
                       float* uy_b, float* ux_i, float* uy_i, float* taub_acx, float* taub_acy,
                       float* beta, float* beta_acx, float* beta_acy, float* beta_eff,
                       float* visc_eff, float* visc_eff_int, float* duxdz, float* duydz,
                       int32_t* ssa_mask_acx, int32_t* ssa_mask_acy, float* ssa_err_acx,
                       float* ssa_err_acy, int32_t* ssa_iter_now, int32_t* ssa_lin_iter,
                       int32_t* ssa_lin_fail, int32_t* ssa_lim_n, float* c_bed,
                       const float* f_slide, const float* taud_acx, const float* taud_acy,
                       const float* taul_int_acx, const float* taul_int_acy, const float* H_ice,
                       const float* f_ice, const float* H_grnd, const float* f_grnd,
                       const float* f_grnd_acx, const float* f_grnd_acy, const float* ATT,
                       const float* zeta_aa, const float* z_sl, const float* z_bed,
                       const float* z_srf, float dx, float dy, float n_glen,
                       ygpu_neff_hook neff_hook, void* hook_ctx);

/* ---- per-solve inputs of the velocity solve (what calc_ydyn computes
 * before calling it, yelmo_dynamics.f90:105-122,464, with the topography
 * step's surface slopes, yelmo_topography.f90:1373-1374,1472), on the
 * device, wp = sp arrays (nx, ny):
 *   mask_frnt        calc_ice_front (topography.f90:724-818)
 *   dzsdx, dzsdy     calc_gradient_acx/acy of z_srf (yelmo_tools.f90:341-478;
 *                    zero_outside false, slope_bg_x/y, clipped at grad_lim)
 *   taud_acx/acy     calc_driving_stress (velocity_general.f90:1118-1205)
 *   taul_int_acx/acy calc_lateral_bc_stress_2D (velocity_general.f90:1453-1609)
 *   ssa_mask_acx/acy set_ssa_masks (solver_ssa_ac.f90:851-996)
 * par supplies boundaries, rho_ice, rho_sw and g; lateral_bc is ssa_lat_bc
 * (0 "none", 1 "floating", 2 "marine", 3 "all").  Any output may be NULL. */
int ygpu_dyn_inputs(ygpu_handle* h, const ygpu_ssa_params* par, int32_t nx, int32_t ny,
                    float dx, float dy, float taud_lim, float grad_lim, float slope_bg_x,
                    float slope_bg_y, int32_t lateral_bc, const float* H_ice, const float* f_ice,
                    const float* f_grnd, const float* z_bed, const float* z_sl,
                    const float* z_srf, int32_t* mask_frnt, float* dzsdx, float* dzsdy,
                    float* taud_acx, float* taud_acy, float* taul_int_acx, float* taul_int_acy,
                    int32_t* ssa_mask_acx, int32_t* ssa_mask_acy);

/* ---- calc_velocity_diva (velocity_diva.f90:59-429), batched form.         */
int ygpu_velocity_diva_batch(ygpu_handle* h, const ygpu_ssa_params* par, int32_t nprob,
                             ygpu_ssa_problem* probs);

/* ---- resident batches: upload once, solve many times.  ygpu_batch_run
 * solves every domain from the device-resident state (reset != 0: from the
 * state the batch was created with, restored device-to-device), with no
 * host traffic besides one flag read per Picard iteration; seconds is the
 * device time of the call.  ygpu_batch_download copies the results into
 * probs (the same array layout as at creation).                             */
typedef struct ygpu_batch ygpu_batch;
int ygpu_batch_create(ygpu_handle* h, const ygpu_ssa_params* par, int32_t nprob,
                      const ygpu_ssa_problem* probs, int32_t diva, ygpu_batch** out,
                      int64_t* device_bytes);
int ygpu_batch_run(ygpu_batch* b, int32_t reset, double* seconds);
int ygpu_batch_download(ygpu_batch* b, ygpu_ssa_problem* probs);
int ygpu_batch_destroy(ygpu_batch* b);
/* y = A x over every domain of a resident batch with the stencil records of
 * the systems assembled in its last Picard iteration and x its last
 * solution, in one full-device launch (fp32 != 0: the mixed mode's FP32
 * records and vectors); reps timed launches after one untimed.  y (may be
 * NULL) receives the 2-per-cell result of every domain in batch order;
 * ms_per_launch the average launch time; bytes_per_launch the algorithmic
 * bytes (18 matrix values, one velocity pair read and written per cell).  */
int ygpu_batch_spmv(ygpu_batch* b, int32_t fp32, int32_t reps, double* y, double* ms_per_launch,
                    double* bytes_per_launch);

/* ---- resident sessions (warm starts for served solve sequences).
 * A session keeps one domain's fields, velocity, viscosity and workspace on
 * the device between solves.  ygpu_session_solve uploads the IN fields of
 * prob, keeps the device-resident INOUT state (velocity, viscosity, beta) as
 * the warm start, runs the Picard loop and downloads the results into prob.
 * Sessions of one handle run on that handle's stream in call order.          */
typedef struct ygpu_session ygpu_session;
int ygpu_session_create(ygpu_handle* h, const ygpu_ssa_params* par, const ygpu_ssa_problem* prob,
                        ygpu_session** out, int64_t* device_bytes);
int ygpu_session_solve(ygpu_session* s, ygpu_ssa_problem* prob);
int ygpu_session_destroy(ygpu_session* s);

/* ---- kernel statistics of a handle (for roofline accounting).  When
 * enabled, every launch of the persistent Krylov kernel is bracketed by HIP
 * events on the handle's stream; algorithmic bytes are counted from the
 * matrix size and the Krylov iterations each launch performed.              */
typedef struct ygpu_stats {
  int64_t krylov_launches;
  double krylov_ms;             /* summed event time of those launches          */
  double krylov_bytes;          /* algorithmic HBM bytes of those launches      */
  int64_t krylov_iters;         /* Krylov iterations summed over members        */
  double unknown_iters;         /* sum over members of N * iterations           */
  int64_t picard_iters;         /* Picard iterations summed over members        */
  int32_t krylov_group;         /* workgroups per member of the last launch (the
                                   largest group when members of uneven sizes
                                   have groups of their own sizes)              */
  int32_t krylov_block;         /* threads per workgroup of the last launch     */
  int64_t refine_sweeps;        /* mixed mode: FP64 residual passes summed over
                                   members (their bytes are in krylov_bytes)    */
} ygpu_stats;
int ygpu_stats_enable(ygpu_handle* h, int32_t on);
int ygpu_stats_get(ygpu_handle* h, ygpu_stats* out);  /* returns and resets     */

/* ---- served solve sequences with continuous batching (no counterpart in
 * the reference, which solves one domain per call; these are the submit /
 * poll entry points of a request loop).  A server holds nlanes size classes:
 * lane l takes domains up to lane_max_side[l] (increasing) in lane_slots[l]
 * member slots, with its own HIP stream and its share of one device arena of
 * at most hbm_budget_bytes (never more than its slots can fill; up to nz_max
 * levels).  ygpu_submit copies one solve's arrays and queues it under a
 * caller-chosen sequence id: the first solve of a sequence carries its INOUT
 * state; solves flagged YGPU_SEQ_CONTINUE read only the IN fields (and masks)
 * and start from the previous solve's device-resident velocity, viscosity
 * and friction; YGPU_SEQ_LAST releases the sequence's slot once that solve
 * finishes.  A sequence holds its slot until then.  ygpu_server_step admits
 * queued solves as slots and arena cells free up (the smallest lane that
 * takes the grid, else a larger one with room), starts one Picard iteration
 * of every admitted solve on each idle lane, and returns once some lane has
 * finished an iteration; finished solves are downloaded at once.  ygpu_poll
 * returns 1 and the next finished ticket (0 when none); ygpu_fetch copies that
 * solve's results into out (same grid; solve_seconds = time from admission to
 * completion).  Host arrays passed to ygpu_submit may be reused as soon as it
 * returns.  Calls on one server are not thread-safe.                        */
#define YGPU_SEQ_CONTINUE 1
#define YGPU_SEQ_LAST 2
typedef struct ygpu_server ygpu_server;
int ygpu_server_create(ygpu_handle* h, const ygpu_ssa_params* par, int32_t diva, int32_t nlanes,
                       const int32_t* lane_max_side, const int32_t* lane_slots, int32_t nz_max,
                       int64_t hbm_budget_bytes, ygpu_server** out, int64_t* device_bytes);
int ygpu_submit(ygpu_server* s, const ygpu_ssa_problem* prob, int64_t seq, int32_t flags,
                int64_t* ticket);
int ygpu_server_step(ygpu_server* s, int32_t* n_running, int32_t* n_queued);
int ygpu_poll(ygpu_server* s, int64_t* ticket);
int ygpu_fetch(ygpu_server* s, int64_t ticket, ygpu_ssa_problem* out);
int ygpu_server_destroy(ygpu_server* s);

/* ---- schedule autotuner of the mixed-precision path (no counterpart in the
 * reference).  The persistent Krylov kernel's schedule for a domain is its
 * number of ILU(0) line blocks (one wave per block; at least ceil(nx/64))
 * and the threads per member block (512 or 1024).  ygpu_autotune times the
 * first Picard iteration's linear solve of prob (best of reps) under each
 * candidate schedule and records the fastest converged one in the handle
 * for prob's grid size (nx, ny); batches, sessions and servers created on
 * the handle afterwards use it for every member of that size when the
 * option string leaves the block count open (-p ilu, no -ygpu_ilu_blocks).
 * ygpu_set_schedule records a schedule directly (ilu_blocks <= 0 removes
 * it; block 0 keeps the size rule).                                         */
int ygpu_autotune(ygpu_handle* h, const ygpu_ssa_params* par, const ygpu_ssa_problem* prob,
                  int32_t diva, int32_t reps, int32_t* ilu_blocks, int32_t* block,
                  double* solve_ms);
int ygpu_set_schedule(ygpu_handle* h, int32_t nx, int32_t ny, int32_t ilu_blocks, int32_t block);

/* device bytes a domain of this size needs (admission control)               */
int64_t ygpu_footprint_bytes(const ygpu_ssa_params* par, int32_t nx, int32_t ny, int32_t nz);

#ifdef __cplusplus
}
#endif
#endif /* YGPU_H */
