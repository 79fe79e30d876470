/*
 * eafp.h -- C-ABI of libeafp.so, the MI355X p' (pressure-correction) solver
 * for EAF3D.
 *
 * The drop-in entry point eafp_cg_solve() replaces the body of the
 * reference's cg_3d_mpi (src/mod_solver_3d.f90:485-620), called from
 * solve_pressure_correction (src/mod_pressure_3d.f90:316-318).  The Fortran
 * shim mod_eafp_solver (hornofusion-full3d_amd/fortran/mod_eafp_solver.f90)
 * keeps cg_3d_mpi's subroutine interface and binds to it with ISO_C_BINDING.
 *
 * Host arrays are the reference's storage: real(8), Fortran order (r
 * fastest, then theta, then z), bounds (-1:nr+2, -1:ntheta+2, -1:nz+2)
 * (src/mod_workspace.f90:124-125), i.e. (nr+4)*(ntheta+4)*(nz+4) doubles.
 *
 * Conventions
 *  - Return value: EAFP_OK (0) on success, a negative EAFP_ERR_* for invalid
 *    input, a positive hipError_t value for a HIP failure.  The Fortran
 *    shim turns any nonzero status into residual = NaN, the failure signal
 *    the reference's caller already consumes (src/mod_pressure_3d.f90:322).
 *  - Every call is synchronous on return and runs on the library's own HIP
 *    stream.  Handles are not thread-safe; one handle, one host thread.
 *  - One GPU (the current HIP device); no MPI inside the solve.
 */
#ifndef EAFP_H
#define EAFP_H

#ifdef __cplusplus
extern "C" {
#endif

#define EAFP_OK 0
#define EAFP_ERR_ARG (-1)          /* bad pointer, size or option */
#define EAFP_ERR_UNSUPPORTED (-2)  /* is_parallel, halo != 2, lbound != -1, ... */
#define EAFP_ERR_COMPONENTS (-3)   /* more singular components than the projection holds */
#define EAFP_ERR_NO_DEVICE (-4)    /* no HIP device visible */
#define EAFP_ERR_STATE (-5)        /* solve before upload, index out of range */
#define EAFP_ERR_SIZE (-6)         /* mesh beyond an allocation-free entry point's scratch */

/* Mesh descriptor: the fields of mesh_t that cg_3d_mpi reads
 * (src/mod_types_3d.f90:143-193; bounds() src/mod_solver_3d.f90:573-585). */
typedef struct eafp_desc {
    int nr, ntheta, nz;   /* owned cells (serial: the whole mesh) */
    int halo;             /* halo depth of the host arrays: must be 2 */
    int lbound;           /* Fortran lower bound of the host arrays: must be -1 */
    int is_parallel;      /* mesh_t%is_parallel: must be 0 (one GPU) */
    int periodic_theta;   /* topo%periodic_theta: must be 1 (EAF cylinder) */
} eafp_desc;

/* Outcome of one solve. */
typedef struct eafp_stats {
    double residual;      /* FP64 CG: relative recurrence residual, as cg_3d_mpi's
                             `residual`; mixed: true relative residual ||b-Ax||/||b|| */
    int n_iter;           /* FP64 CG: iterations, as cg_3d_mpi's `n_iter`;
                             mixed: total inner FP32 iterations */
    int n_refine;         /* mixed: FP64 refinement steps (0 for FP64 CG) */
    int status;           /* EAFP_STOP_* */
    int pad_;
    double true_rel_l2;   /* ||b - A x||_2 / ||b||_2 with the FP64 operator */
    double true_rel_l1;   /* sum|b - A x| / sum|b| (the L1 form of res_cont) */
    double seconds;       /* wall time of the device work of this solve */
} eafp_stats;

#define EAFP_STOP_RUNNING 0
#define EAFP_STOP_CONVERGED 1    /* residual < tol */
#define EAFP_STOP_BREAKDOWN 2    /* |p.q| < 1e-30 (src/mod_solver_3d.f90:535) */
#define EAFP_STOP_MAX_ITER 3     /* max_iter reached */
#define EAFP_STOP_ZERO_RHS 4     /* ||Su|| < 1e-30 (src/mod_solver_3d.f90:516-521) */

/* Coefficient storage of the mixed-precision inner solver. */
#define EAFP_COEF_FP32 0
#define EAFP_COEF_BF16 1
#define EAFP_COEF_FP16 2         /* FP16 with a per-cell FP32 scale */

typedef struct eafp_mixed_opts {
    double tol;           /* target: true relative residual ||b-Ax||/||b|| (e.g. 1e-8) */
    int interval;         /* inner FP32 iterations between FP64 refinement steps
                             (rounded up to even; 0: library default) */
    int max_steps;        /* cap on FP64 refinement steps */
    int coef_format;      /* EAFP_COEF_* */
    int restart;          /* 0: the inner CG keeps its search direction across
                             refinement steps (residual replacement); 1: it restarts
                             at every step (classical iterative refinement) */
    int precond;          /* EAFP_PRECOND_*: preconditioner of the inner CG */
} eafp_mixed_opts;

/* Inner preconditioners of the mixed solver.  JACOBI: the diagonal of the
 * rounded operator (what cg_3d_mpi uses).  ZLINE: its tridiagonal part along
 * every z-line, M = L D L^T (l in FP16, 1/d in BF16: 4 bytes per cell, as
 * Jacobi's 1/D), applied by two scans per line; fewer inner iterations on
 * the furnace's p' systems.  ZLINE needs every nz at most 1024
 * (EAFP_ERR_UNSUPPORTED otherwise). */
#define EAFP_PRECOND_JACOBI 0
#define EAFP_PRECOND_ZLINE 1

/* ---------------------------------------------------------------------------
 * Drop-in: cg_3d_mpi(aW,aE,aS,aN,aB,aT,aP,Su,phi,m,max_iter,tol,residual,n_iter)
 * (src/mod_solver_3d.f90:485-496).  Jacobi-PCG in FP64 on the GPU with the
 * reference's semantics: x0 = phi, recurrence-residual stop test, |p.q| <
 * 1e-30 exit, ||Su|| < 1e-30 early return, phi updated on the interior and
 * its theta halos (j = -1, 0, ntheta+1, ntheta+2) filled by periodic copy;
 * r/z halos of phi are left as they were.  Device buffers are cached per
 * mesh shape across calls (the role of ensure_workspace,
 * src/mod_workspace.f90:120-144).
 * ------------------------------------------------------------------------- */
int eafp_cg_solve(const eafp_desc* desc,
                  const double* aW, const double* aE, const double* aS, const double* aN,
                  const double* aB, const double* aT, const double* aP, const double* Su,
                  double* phi, int max_iter, double tol, double* residual, int* n_iter);

/* Releases the device buffers cached by eafp_cg_solve, eafp_pp_assemble,
 * eafp_pp_correct, eafp_pp_step and eafp_tdma_solve(_batch).  No reference counterpart
 * (the reference's workspace lives until exit). */
int eafp_release(void);

/* ---------------------------------------------------------------------------
 * The p' system assembled on the GPU from the phase fields: the work that
 * solve_pressure_correction (src/mod_pressure_3d.f90:88-306) does before
 * its cg_3d_mpi call, for one serial mesh.  It covers:
 *   - the active phases (:127-169);
 *   - the cell pressure gradients (cell_gradients :612-651);
 *   - the Rhie-Chow face terms of the liquid and, with gas_in_poisson and
 *     solve_multiphase, of the gas (add_phase_contribution :469-610);
 *   - the gas compressibility source (:226-246);
 *   - the compliance and acoustic diagonal, or the trivial rows (:249-281);
 *   - the pressure boundary rows (apply_pressure_bc,
 *     src/mod_boundary_3d.f90:180-277);
 *   - the continuity residual sum|Su| / flux_ref (:290-306).
 * Every 3-D array is padded Fortran order with bounds (-1:n+2); r, theta, z
 * are the cell-centre coordinates with bounds (-1:n+2) (the mesh_t fields
 * m%r, m%theta, m%z); the face-area arrays index faces as the reference
 * does (Ar(i) the face i+1/2, Az(k) the face k+1/2).  cell_type follows
 * mesh_t (0 outside the fluid and in every r/z halo); outlet is the top
 * layer's electrode mask, (nr+4)*(ntheta+4) in the same order
 * (is_electrode(:,:,nz,:) of any electrode).  The four workspace masks are
 * optional: NULL means "not valid" (ws_liq_cont_valid, ws_pv_valid,
 * ws_gas_perc_valid, ws_liq_room_valid false).  Integer masks are 0/1.
 * ------------------------------------------------------------------------- */
typedef struct {
    const double *r, *theta, *z;                /* m%r, m%theta, m%z: (-1:n+2) */
    const double *vol, *Ar, *Ath, *Az;          /* m%vol, m%Ar, m%Ath, m%Az */
    const int* cell_type;                       /* m%cell_type */
    const int* outlet;                          /* any(m%is_electrode(:,:,nz,:)) */
    const double *alpha_l, *ur_l, *uth_l, *uz_l, *aPr_l, *aPth_l, *aPz_l, *rho_l, *T_l;        /* liq */
    const double *alpha_g, *ur_g, *uth_g, *uz_g, *aPr_g, *aPth_g, *aPz_g, *rho_g, *T_old_g;    /* gas, gas_T_old */
    const double* p;                            /* sh%p */
    const int *liq_cont, *pv_active, *gas_reach;   /* ws_liq_cont, ws_pv_active, ws_gas_reach */
    const double* liq_room;                     /* ws_liq_room */
} eafp_pp_fields;

typedef struct {
    double dt, rho_gas, T_ambient, beta_expansion;               /* cfg%dt, cfg%rho_gas, ... */
    int gas_in_poisson, solve_multiphase, gas_compressibility;   /* cfg flags (0/1) */
    int pad_;
} eafp_pp_config;

/* The liquid's +face terms that the assembly keeps for the post-solve
 * fluxes: F* and a_nb of every open E, N and T face of the liquid (Fs_r,
 * Fs_th, Fs_z, ac_r, ac_th, ac_z, locals of solve_pressure_correction,
 * src/mod_pressure_3d.f90:117-118, set by add_phase_contribution :525-527,
 * :562-564, :601-603) and their link flags (ws_Fc_lk_r/th/z, 0/1).  Padded
 * arrays, zero where no open liquid face. */
typedef struct {
    double *Fs_r, *Fs_th, *Fs_z, *ac_r, *ac_th, *ac_z;
    int *lk_r, *lk_th, *lk_z;
} eafp_pp_exports;

/* Writes the assembled aW..aT, aP, Su (padded, halos zero) and the
 * continuity residual and its normaliser flux_ref (either may be NULL).
 * exports: NULL, or where to write the liquid's face terms (every pointer
 * set).  arrays_on_device = 1: every array pointer, inputs and outputs, is a
 * device pointer (nothing crosses PCIe).  The coefficients and exports are
 * bit-identical to a sequential FP64 evaluation of the reference's
 * expressions in their written order without contraction; the two global
 * sums of the residual are reduced in a different order.  Serial meshes
 * only (EAFP_ERR_UNSUPPORTED when desc->is_parallel). */
int eafp_pp_assemble(const eafp_desc* desc, const eafp_pp_fields* fields, const eafp_pp_config* cfg,
                     double* aW, double* aE, double* aS, double* aN, double* aB, double* aT,
                     double* aP, double* Su, double* residual, double* flux_ref,
                     eafp_pp_exports* exports, int arrays_on_device);

/* Largest padded mesh, (nr+4)*(ntheta+4)*(nz+4) cells, that
 * eafp_pp_assemble_noalloc takes (60x120x84 is 698,368). */
#define EAFP_NOALLOC_MAX_CELLS 1048576

/* eafp_pp_assemble with every array in device memory, for callers that
 * cannot allocate device memory: code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc.  No counterpart in the
 * reference.  The same kernels and results, bit for bit (the residual's
 * block sums are reduced on the device in the same order), but no
 * hipMalloc, hipFree or other allocating call: the kernels' scratch (the
 * cell gradients, the residual's block sums) is a static device array of
 * the library, 32.5 MiB held by each GPU the library runs on until the
 * process ends.  Meshes beyond EAFP_NOALLOC_MAX_CELLS padded cells:
 * EAFP_ERR_SIZE.  Runs on `stream` (a hipStream_t; NULL: the default
 * stream) and returns synchronised; residual and flux_ref are host
 * pointers (may be NULL).  Calls on one GPU share its scratch array, so
 * calls that overlap on one GPU (from other threads or streams) race. */
int eafp_pp_assemble_noalloc(const eafp_desc* desc, const eafp_pp_fields* fields, const eafp_pp_config* cfg,
                             double* aW, double* aE, double* aS, double* aN, double* aB, double* aT,
                             double* aP, double* Su, double* residual, double* flux_ref,
                             eafp_pp_exports* exports, void* stream);

/* State the post-solve consumers write (padded arrays). */
typedef struct {
    double *Fc_r, *Fc_th, *Fc_z;       /* out: ws_Fc_r, ws_Fc_th, ws_Fc_z */
    int *lk_r, *lk_th, *lk_z;          /* out: ws_Fc_lk_r/th/z after the halo copy */
    double *ur_l, *uth_l, *uz_l;       /* in/out: liq%ur, liq%uth, liq%uz */
    double *ur_g, *uth_g, *uz_g;       /* in/out: gas%ur, gas%uth, gas%uz */
    double* p;                         /* in/out: sh%p */
} eafp_pp_update;

/* What solve_pressure_correction does with p' after its cg_3d_mpi call
 * (src/mod_pressure_3d.f90:324-397), for one serial mesh:
 *   - the conservative liquid fluxes Fc = Fs + ac (pp_P - pp_nb) of the
 *     linked E, N and T faces (:333-351), zero elsewhere; the theta halos of
 *     Fc and of the link flags by periodic copy;
 *   - correct_velocities (:689-746) of the liquid and, with gas_in_poisson
 *     and solve_multiphase, of the gas: u -= alpha vol / aP_u * (centred
 *     difference of pp), skipped where |aP_u| <= 1e-30, the r part on
 *     i = 1, nr and the z part on k = 1, nz;
 *   - p = min(max(p + alpha_p pp, -2e6), 2e6) over the whole array (:358-360);
 *   - sealed cells (fluid, neither phase active) take the mean p of their
 *     neighbours with an active phase, 0 without one (:373-397); then the
 *     theta halo copy of p.
 * fields: the assembly's inputs, of which r, theta, z, vol, cell_type,
 * alpha_l/alpha_g, the six aP_u arrays and the masks are read (the
 * velocities and p updated are io's).  pp: the solved correction with its
 * theta halos filled (cg_3d_mpi and eafp_cg_solve leave them so).
 * exports: as eafp_pp_assemble wrote them for the same fields.  Every
 * output is bit-identical to a sequential FP64 evaluation (nothing is
 * reduced).  arrays_on_device as for eafp_pp_assemble. */
int eafp_pp_correct(const eafp_desc* desc, const eafp_pp_fields* fields, const eafp_pp_config* cfg,
                    double alpha_p, const double* pp, const eafp_pp_exports* exports, const eafp_pp_update* io,
                    int arrays_on_device);

/* ---------------------------------------------------------------------------
 * One whole p' step of solve_pressure_correction (src/mod_pressure_3d.f90:
 * 127-397) in one call, every intermediate kept in HBM: sh%pp = 0 (:174),
 * the assembly (as eafp_pp_assemble, with the face exports), the solve of
 * the assembled system from pp = 0 (:316-318), then the post-solve steps
 * (as eafp_pp_correct) with alpha_p.  As in the reference a NaN solver
 * residual replaces the continuity residual (:322) and the corrections are
 * applied regardless.
 *   solver EAFP_STEP_FP64: cg_3d_mpi's Jacobi-PCG (max_iter, tol, the
 *     recurrence residual); every output equals what eafp_pp_assemble,
 *     eafp_cg_solve (phi = 0) and eafp_pp_correct give called in turn.
 *   solver EAFP_STEP_MIXED: the mixed-precision solver with `mixed`
 *     (tol on the true relative residual); max_iter and tol unused.
 * pp receives p' (padded; theta halos filled, r/z halos zero); io is updated
 * as by eafp_pp_correct.  residual (the continuity residual of the step)
 * and stats (the solve's outcome) may be NULL.  The exports and the p'
 * system stay inside the library.  arrays_on_device: every array of
 * fields, io and pp is device memory (1) or host memory (0; uploaded once,
 * the results downloaded once).
 * No single reference counterpart: it fuses the three steps of one
 * subroutine, for callers whose fields already live on the GPU.
 * ------------------------------------------------------------------------- */
#define EAFP_STEP_FP64 0
#define EAFP_STEP_MIXED 1

typedef struct eafp_pp_step_opts {
    int solver;               /* EAFP_STEP_* */
    int max_iter;             /* FP64: cfg%max_inner_pres */
    double tol;               /* FP64: SOR_TOL_PRESSURE */
    double alpha_p;           /* cfg%alpha_p */
    eafp_mixed_opts mixed;    /* EAFP_STEP_MIXED */
} eafp_pp_step_opts;

int eafp_pp_step(const eafp_desc* desc, const eafp_pp_fields* fields, const eafp_pp_config* cfg,
                 const eafp_pp_step_opts* opts, double* pp, const eafp_pp_update* io, double* residual,
                 eafp_stats* stats, int arrays_on_device);

/* ---------------------------------------------------------------------------
 * Device-resident systems.  A batch holds nsys systems of any shapes; nsys
 * = 1 is the single-system case.  Upload once, solve many times: every
 * solve restarts from the uploaded phi (the reference zeroes sh%pp before
 * each call, src/mod_pressure_3d.f90:174).  These entry points have no
 * one-to-one reference counterpart: they split cg_3d_mpi's implicit
 * copy-in / solve / copy-out so operands can stay resident in HBM.
 * ------------------------------------------------------------------------- */
typedef struct eafp_batch eafp_batch;

int eafp_batch_create(int nsys, const eafp_desc* descs, eafp_batch** out);

/* Uploads system s.  Arrays are padded Fortran layout; arrays_on_device = 0
 * for host pointers, 1 for device pointers (packed where they lie).  phi0
 * may be NULL (zero).  The singular components the mixed solver projects
 * are labelled on the GPU by the first mixed solve (or
 * eafp_batch_components call) after the upload; FP64 solves never label. */
int eafp_batch_upload(eafp_batch* b, int s,
                      const double* aW, const double* aE, const double* aS, const double* aN,
                      const double* aB, const double* aT, const double* aP, const double* Su,
                      const double* phi0, int arrays_on_device);

/* The reference algorithm (FP64 Jacobi-PCG, as eafp_cg_solve) on every
 * system of the batch in one launch sequence with per-system convergence
 * masks.  stats: nsys entries (may be NULL). */
int eafp_batch_solve_fp64(eafp_batch* b, int max_iter, double tol, eafp_stats* stats);

/* Mixed precision: FP32 inner PCG (Jacobi or z-line preconditioner,
 * opts->precond) on low-precision symmetric face coefficients, wrapped in
 * FP64 iterative refinement: every `interval`
 * inner iterations the FP64 correction is folded into x, the true residual
 * b - A x is recomputed with the caller's FP64 operator and replaces the
 * inner residual, projected to zero mean on every singular Neumann
 * component (FP64); every dot product is accumulated in FP64.  The result
 * is gauge-fixed: zero mean per singular component.  stats: nsys entries
 * (may be NULL). */
int eafp_batch_solve_mixed(eafp_batch* b, const eafp_mixed_opts* opts, eafp_stats* stats);

/* Copies system s's phi (padded Fortran layout: interior from the last
 * solve, theta halos periodic copies, r/z halos as uploaded) to phi;
 * to_device = 1 if phi is a device pointer. */
int eafp_batch_download(eafp_batch* b, int s, double* phi, int to_device);

/* Launch shape in use: blocks per launch and cells per tile (for reports). */
int eafp_batch_info(const eafp_batch* b, int* grid, int* tile_cells, long* ncells);

/* Launch shape of every kernel of the batch: a tile of 256*cells_per_thread
 * cells (cells_per_thread in {1, 2, 4, 8}) and `blocks` blocks per launch
 * (0: one per tile, at most 8 per CU).  No reference counterpart. */
int eafp_batch_set_tiling(eafp_batch* b, int cells_per_thread, int blocks);

/* The mixed solver's stencil as a column march: r_splits > 0 cuts every
 * system into columns of theta-rows (about 1024 cells per r-plane) marched
 * through the r-planes, each column in r_splits segments (neighbouring
 * segments march in opposite directions, so the planes at a boundary are
 * mostly read from HBM once); 0 returns to the tile kernels.
 * EAFP_ERR_UNSUPPORTED when a system's nz exceeds 256 or a system has fewer
 * tiles than segments.  Kept across eafp_batch_set_tiling while the new
 * tiles allow it.  No reference counterpart. */
int eafp_batch_set_columns(eafp_batch* b, int r_splits);

/* Column setting in use (0: tile kernels). */
int eafp_batch_get_columns(const eafp_batch* b, int* r_splits);

/* Measurement: average duration (ms) of one launch of each of the two
 * iteration kernels of a solver (solver 0: FP64 CG, k64_spmv/k64_update;
 * solver 1: mixed, k_mspmv/k_mupdate with coef_format; solver 2: mixed
 * with the z-line inner preconditioner, k_mspmv/k_mupdate_line) over `reps`
 * iterations on the uploaded systems, timed with HIP events on the
 * library's stream.  Leaves the batch's results undefined (solve again).
 * No reference counterpart. */
int eafp_batch_time_kernels(eafp_batch* b, int solver, int coef_format, int reps, float* ms_spmv,
                            float* ms_update);

/* Autotuner: times every tile shape (cells per thread x blocks per CU) for
 * the solver's iteration kernels (solver as for eafp_batch_time_kernels)
 * and keeps the fastest; for the mixed solver it then times the
 * column-marching stencil at 1, 2, 4, 6, 8, 10, 12 and 16 r-segments and
 * keeps it if it is faster (eafp_batch_get_columns tells).
 * Returns the tile shape and the iteration time (ms).  No reference
 * counterpart. */
int eafp_batch_autotune(eafp_batch* b, int solver, int coef_format, int reps, int* cells_per_thread,
                        int* blocks, float* ms_iter);

/* The singular components of system s: connected components of the link
 * graph (a link where either side's coefficient is nonzero, theta periodic)
 * whose every row is a pure-Neumann row (aP = sum of the interior links
 * within 1e-10 relative, no wall link), i.e. the blocks whose p' is fixed
 * only up to a constant.  Labels the batch's systems on the GPU if an
 * upload has not been labelled yet.  *nsingular: how many; cells[q]: the
 * cell count of each, ordered by first cell (cell index
 * (i*ntheta + j)*nz + k), when there are at most 4 (0 past the last, and
 * all 0 beyond 4).  The mixed solver projects at most 4 per system and
 * refuses more (EAFP_ERR_COMPONENTS).  No reference
 * counterpart (the reference regularises with PP_COMPLIANCE instead,
 * src/mod_pressure_3d.f90:261-265). */
int eafp_batch_components(eafp_batch* b, int s, int* nsingular, int cells[4]);

int eafp_batch_destroy(eafp_batch* b);

/* Process-wide activity counters, for diagnostics and tests: systems packed
 * into HBM (every eafp_cg_solve, eafp_batch_upload and eafp_pp_step call),
 * systems whose singular components were labelled, solver graphs captured
 * and instantiated, and solver graph launches.  No reference counterpart. */
typedef struct eafp_counters {
    long long uploads;
    long long labelled;
    long long graph_captures;
    long long graph_launches;
} eafp_counters;

int eafp_get_counters(eafp_counters* out);

/* ---------------------------------------------------------------------------
 * Line-by-line TDMA smoother: tdma_3d_mpi (src/mod_solver_3d.f90:252-366)
 * on a serial mesh, with its Thomas solve `tdma` (:22-46).  Each sweep:
 * the theta halo copy, the r lines (k outer, j inner), the copy, the theta
 * lines (k outer, i inner), the copy, the z lines (j outer, i inner); each
 * line takes the other directions' neighbours from phi as it stands and its
 * two end neighbours from the halos; no copy follows the last z pass.  The
 * GPU keeps this order exactly: lines on one anti-diagonal of the line grid
 * never neighbour each other, so they are solved together, one diagonal
 * after another.  Every value of phi (halos included) is bit-identical to a
 * sequential FP64 evaluation of the reference's expressions without
 * contraction.  Serial meshes only (EAFP_ERR_UNSUPPORTED otherwise);
 * n_sweep = 0 leaves phi as it is.
 * ------------------------------------------------------------------------- */
int eafp_tdma_solve(const eafp_desc* desc, const double* aW, const double* aE, const double* aS, const double* aN,
                    const double* aB, const double* aT, const double* aP, const double* Su, double* phi,
                    int n_sweep, int arrays_on_device);

/* One system of a batched TDMA call (padded Fortran arrays). */
typedef struct {
    const double *aW, *aE, *aS, *aN, *aB, *aT, *aP, *Su;
    double* phi;
} eafp_tdma_system;

/* n_sweep sweeps of eafp_tdma_solve on each of nsys independent systems of
 * any shapes (the momentum components, energy, species, ... of one time
 * step), solved concurrently.  No reference counterpart (the reference
 * calls tdma_3d_mpi once per equation). */
int eafp_tdma_solve_batch(int nsys, const eafp_desc* descs, const eafp_tdma_system* systems, int n_sweep,
                          int arrays_on_device);

/* Library version string. */
const char* eafp_version(void);

#ifdef __cplusplus
}
#endif

#endif /* EAFP_H */
