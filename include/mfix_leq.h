/*
 * mfix_leq.h - C ABI of the MI355X (gfx950) BiCGSTAB linear-equation solver
 * that replaces MFIX's LEQ_BICGS hot path.
 *
 * Reference interfaces replaced (paths relative to the MFIX 2016-1 tree):
 *   LEQ_BICGS(VNAME,VNO,VAR,A_M,B_M,CMETHOD,TOL,PC,ITMAX,IER)
 *       model/leq_bicgs.f:23-90 (PC dispatch) and :107-509 (LEQ_BICGS0)
 *   SOLVE_LIN_EQ(VNAME,VNO,VAR,A_M,B_M,M,ITMAX,METHOD,SWEEP,TOL1,PC,IER)
 *       model/solve_lin_eq.f:17-195 (METHOD=2 branch, :151-152)
 *   LEQ_MATVEC   model/leqsol_mod.f:103-265
 *   LEQ_MSOLVE   model/leqsol_mod.f:286-598 (LINE; sweeps :767-1086, DGTSV)
 *   LEQ_MSOLVE0  model/leqsol_mod.f:616-663 (NONE)
 *   LEQ_MSOLVE1  model/leqsol_mod.f:681-749 (DIAG)
 *   dot_product_par / dot_product_par2  model/leqsol_mod.f:1094-1295
 *   grid descriptor (module globals c0,c1,c2, ranges, CELL_CLASS,
 *   INCREMENT_FOR_MP)  model/dmp_modules/gridmap_mod.f:259-556,
 *   model/set_increments.f:72-178
 *
 * Storage contract (identical to the reference's serial layout):
 *   ni = IMAX+2, nj = JMAX+2, nk = KMAX+2 (nk = 1 for a 2-D run, NO_K),
 *   DIMENSION_3 = ni*nj*nk, ijk (0-based) = (j-1) + (i-1)*nj + (k-1)*nj*ni
 *   with 1-based i,j,k: J is the stride-1 axis.
 *   A_M is DIMENSION_3 x 7 column-major: plane l (l = -3..3) starts at
 *   element (l+3)*DIMENSION_3; l = -3 bottom (K-), -2 south (J-),
 *   -1 west (I-), 0 diagonal, +1 east (I+), +2 north (J+), +3 top (K+).
 *
 * Return values: every function returns 0 on success or a negative
 * MFIX_LEQ_E* code.  The reference's own solver status (IER = 0/-1/-2) is
 * returned through the ier argument exactly as LEQ_BICGS sets it.
 */
#ifndef MFIX_LEQ_H
#define MFIX_LEQ_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- error codes of the C ABI (never confused with IER, which is <= 0) -- */
#define MFIX_LEQ_OK 0
#define MFIX_LEQ_EBADPC (-101)   /* PC not LINE/DIAG/NONE: reference calls mfix_exit (leq_bicgs.f:83-87) */
#define MFIX_LEQ_ENOGRID (-102)  /* mfix_leq_grid_init not called / invalid grid */
#define MFIX_LEQ_EHIP (-103)     /* a HIP runtime call failed */
#define MFIX_LEQ_EINVAL (-104)   /* invalid argument */
#define MFIX_LEQ_ENOMEM (-105)   /* device allocation failed / over budget */
#define MFIX_LEQ_EDGTSV (-106)   /* DGTSV INFO>0 in a JK/IJ sweep: reference calls mfix_exit (leqsol_mod.f:991-996,1073-1078) */
#define MFIX_LEQ_EBUSY (-107)    /* serving queue full */
#define MFIX_LEQ_EMETHOD (-108)  /* SOLVE_LIN_EQ METHOD other than 2 (BiCGSTAB) */
#define MFIX_LEQ_ETOOBIG (-109)  /* system larger than mfix_leq_bicgs_dev_noalloc's fixed workspace */

/* ---- grid descriptor: replaces the compar/geometry/indices module state -- */
typedef struct mfix_leq_grid {
    int imax, jmax, kmax;        /* interior cells per axis (IMAX, JMAX, KMAX) */
    int no_k;                    /* 1 = two-dimensional run (NO_K); kmax ignored */
    int cyclic_x, cyclic_y, cyclic_z; /* CYCLIC_X/Y/Z with NODES*=1 (set_increments.f:72-112) */
    int is_serial;               /* IS_SERIAL (init_namelist.f:823): 1 = dot over all DIMENSION_3 cells */
    int minimize_dotproducts;    /* MINIMIZE_DOTPRODUCTS (leqsol_mod.f:29) */
    int icheck_bicgs;            /* ICHECK_BICGS (init_namelist.f:802), >= 1 */
} mfix_leq_grid;

/* Per-solve statistics (the reference exposes only IER and ITER_TOT). */
typedef struct mfix_leq_stats {
    int iterations;              /* BiCGSTAB iterations executed (loop trips) */
    int iter_tot_increment;      /* what LEQ_BICGS0 adds to ITER_TOT(VNO) */
    int ier;                     /* IER as LEQ_BICGS returns it */
    int exit_reason;             /* 0 converged, 1 S-norm exit, 2 rho=0, 3 itmax, 4 zero row */
    double rnorm0;               /* ||b - A x0|| of the scaled system */
    double rnorm;                /* final recursive residual norm */
    double true_rnorm;           /* ||b - A x|| recomputed in FP64 (mixed/refined paths; else -1) */
    int refinements;             /* FP64 refinement cycles (mixed precision), 0 for FP64 solves */
    int inner_iterations;        /* FP32 inner iterations summed over cycles (mixed) */
    double device_ms;            /* device time of the solve measured with HIP events */
    int restarts;                /* mixed precision: recurrence restarts after a drifting update */
} mfix_leq_stats;

/* ---- lifetime ----------------------------------------------------------- */
int mfix_leq_init(int device);                 /* select device, create stream */
int mfix_leq_finalize(void);                   /* free every device buffer */
int mfix_leq_grid_init(const mfix_leq_grid *g);/* replaces GRIDMAP_INIT+SET_INCREMENTS state */
int64_t mfix_leq_dimension_3(void);            /* DIMENSION_3 of the current grid, or <0 */
int mfix_leq_iter_tot(int vno);                /* ITER_TOT(VNO), VNO in 1..10 (leqsol_mod.f:17) */
void mfix_leq_reset_iter_tot(void);
const char *mfix_leq_error_string(int code);

/* ---- drop-in solver entry (host arrays, blocking) ------------------------ */
/* LEQ_BICGS model/leq_bicgs.f:23-90.  VAR/A_M/B_M are host arrays in the
 * storage contract above; on return VAR holds the solution and A_M/B_M the
 * row-scaled system, as the reference leaves them (leq_bicgs.f:194-233). */
int mfix_leq_bicgs(const char *vname, int vname_len, int vno, double *var,
                   double *a_m, double *b_m, const char *cmethod,
                   int cmethod_len, double tol, const char *pc, int itmax,
                   int *ier);

/* SOLVE_LIN_EQ model/solve_lin_eq.f:17-195 for METHOD=2: A_M is
 * DIMENSION_3 x 7 x (DIMENSION_M+1) and B_M DIMENSION_3 x (DIMENSION_M+1),
 * phase M (0-based as in the reference) is sliced out (solve_lin_eq.f:151).
 * Other METHODs return MFIX_LEQ_EMETHOD. */
int mfix_solve_lin_eq(const char *vname, int vname_len, int vno, double *var,
                      double *a_m, double *b_m, int m, int itmax, int method,
                      const char *sweep, double tol1, const char *pc, int *ier);

/* Options of the device-resident entry points. */
typedef struct mfix_leq_opts {
    int write_back_scaled;       /* host entry: copy scaled A_M/B_M back (1 = reference behaviour) */
    const double *noise;         /* optional U[0,1) array (DIMENSION_3) for Rtilde (leq_bicgs.f:298-305); device ptr for _dev calls, host ptr for host calls; NULL = internal generator */
    uint64_t noise_seed;         /* seed of the internal counter-based U[0,1) generator */
    double inner_tol;            /* mixed precision: each FP32 cycle reduces the recursive residual by this
                                    factor before an FP64 reliable update (0 = default 0.1) */
    int max_refine;              /* mixed precision: end the solve after this many reliable updates
                                    (IER -1/-2 by the true residual); 0 = no limit */
} mfix_leq_opts;

/* Device-resident FP64 LEQ_BICGS: var/a_m/b_m are device pointers (storage
 * contract above), stream may be 0.  Same semantics as mfix_leq_bicgs. */
int mfix_leq_bicgs_dev(const char *cmethod, double tol, const char *pc,
                       int itmax, double *d_var, double *d_a_m, double *d_b_m,
                       const mfix_leq_opts *opts, mfix_leq_stats *stats,
                       void *stream);

/* mfix_leq_bicgs_dev without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under
 * stream capture, tight serving loops, harnesses that forbid it).  The
 * workspace, system table, block maps and poll words live in a fixed
 * device array of MFIX_LEQ_NOALLOC_BYTES per GPU (held until the process
 * ends; its address is taken when the library initialises on that GPU).
 * Same arguments, results and bits as mfix_leq_bicgs_dev; no graph replays
 * and no timing (stats->device_ms = 0).  Needs mfix_leq_init /
 * mfix_leq_grid_init first (they allocate); returns MFIX_LEQ_ETOOBIG for a
 * system whose workspace does not fit (FP64 DIAG/NONE up to about 1.3 M
 * cells, LINE about 0.7 M).  Every call on a GPU uses that GPU's one array:
 * calls that overlap on one GPU race and must be serialised by the caller. */
#define MFIX_LEQ_NOALLOC_BYTES ((size_t)64 << 20)
int mfix_leq_bicgs_dev_noalloc(const char *cmethod, double tol, const char *pc,
                               int itmax, double *d_var, double *d_a_m,
                               double *d_b_m, const mfix_leq_opts *opts,
                               mfix_leq_stats *stats, void *stream);

/* Mixed precision: FP32 inner BiCGSTAB (matvec, preconditioner, Krylov
 * vectors) inside FP64 iterative refinement; FP64 true residual, solution
 * accumulation and scalar reductions.  Converged when
 * ||b - A x|| <= tol * ||b - A x0|| in FP64 on the row-scaled system.
 * No counterpart in the reference. */
int mfix_leq_bicgs_mixed_dev(const char *cmethod, double tol, const char *pc,
                             int itmax, double *d_var, double *d_a_m,
                             double *d_b_m, const mfix_leq_opts *opts,
                             mfix_leq_stats *stats, void *stream);

struct mfix_leq_resid; /* defined with the residual entries below */

/* CALC_RESID, UNDER_RELAX_S and LEQ_BICGS of one equation, the residual and
 * the under-relaxation done by the solver's set-up pass (the pass that
 * scales the rows and forms R = B - A*VAR) instead of two passes of their
 * own over A_M, B_M and VAR: MFIX's sequence CALC_RESID_S, UNDER_RELAX_S,
 * SOLVE_LIN_EQ (model/solve_scalar_eq.f:129-152,
 * model/solve_energy_eq.f:253-279, model/solve_species_eq.f:163-175) or
 * CALC_RESID_PP, SOLVE_LIN_EQ (model/solve_pp_g.f:104-119).  resid_kind: 0 =
 * CALC_RESID_C, 1 = CALC_RESID_S (resid_tol, d_ep_s, dil_ep_s as in
 * mfix_calc_resid_s_dev), 2 = CALC_RESID_PP (norm), -1 = none; the residual
 * is of the equation as passed, before the under-relaxation.  ur: UR_FAC of
 * UNDER_RELAX_S, 0 = none.  mixed: 0 = mfix_leq_bicgs_dev's FP64 solve, 1 =
 * mfix_leq_bicgs_mixed_dev's.  Results: *resid as the mfix_calc_resid_*_dev
 * call gives it (NUM and DEN to summation-order rounding, the rest
 * exactly); VAR, A_M, B_M, the return code and *stats exactly as
 * mfix_under_relax_s_dev followed by the solve.  No counterpart in the
 * reference, which runs the routines in turn; device pointers. */
int mfix_leq_resid_relax_bicgs_dev(int resid_kind, const int *d_flag, double resid_tol,
                                   const double *d_ep_s, double dil_ep_s, double norm,
                                   int debug_resid, struct mfix_leq_resid *resid, double ur,
                                   const char *cmethod, double tol, const char *pc, int itmax,
                                   int mixed, double *d_var, double *d_a_m, double *d_b_m,
                                   const mfix_leq_opts *opts, mfix_leq_stats *stats,
                                   void *stream);

/* ---- kernels of the path, device pointers (for parity tests) ------------ */
/* LEQ_MATVEC leqsol_mod.f:103-265: y = A x over every cell. */
int mfix_leq_matvec_dev(const double *d_a_m, const double *d_x, double *d_y,
                        void *stream);
/* LEQ_MSOLVE / MSOLVE1 / MSOLVE0: z = M^-1 r for pc LINE (with cmethod)/DIAG/NONE. */
int mfix_leq_msolve_dev(const char *pc, const char *cmethod, const double *d_a_m,
                        const double *d_r, double *d_z, void *stream);
/* dot_product_par leqsol_mod.f:1094-1201 (grid's is_serial decides the domain). */
int mfix_leq_dot_dev(const double *d_x, const double *d_y, double *result,
                     void *stream);
/* Row max-norm unit scaling, leq_bicgs.f:194-233; *ier = -2 on a zero row. */
int mfix_leq_scale_dev(double *d_a_m, double *d_b_m, int *ier, void *stream);
/* FP32 stencil row of the mixed-precision inner solver (no reference
 * counterpart): d_a_m holds 7 FP32 planes in the difference form, planes
 * 0-2 and 4-6 the off-diagonal couplings as in A_M and plane 3 the row
 * excess A0 + sum of the six couplings; y = sum_l a_l (x_nb - x_c) + excess
 * x_c, products in FP32, sums in FP64, one rounding to FP32. */
int mfix_leq_matvec32_dev(const float *d_a_m, const float *d_x, float *d_y,
                          void *stream);
/* LINE preconditioner of the mixed-precision inner solver, the FP32
 * counterpart of LEQ_MSOLVE (model/leqsol_mod.f:286-598): z = r, then the
 * red-black J-line passes of `cmethod` (R/S characters; others run one line
 * per thread).  d_a_m holds 8 FP32 planes of the current grid: the six
 * couplings at planes 0-2, 4-6 as in A_M and A0 at plane 7 (plane 3, the
 * row excess, is not read).  Each line of a red-black pass is split over
 * several lanes (partitioned elimination + cyclic reduction, no pivoting),
 * so the result agrees with the FP64 line solves to FP32 rounding, not bit
 * for bit. */
int mfix_leq_msolve32_dev(const char *cmethod, const float *d_a_m, const float *d_r, float *d_z,
                          void *stream);

/* ---- synthetic systems (host only, deterministic, no GPU needed) -------- */
/* kind 0: symmetric pressure-correction, faces 10^U[-1,1], A0=-sum-1e-3U, B=U[-1,1]
 * kind 1: nonsymmetric first-order-upwind momentum, D=0.5*4^U, divergence-free
 *         8-mode flux with max cell Peclet 10, transient rho V/dt at CFL 0.5
 * kind 2: kind 0 faces times face-averaged eps_g (2000 spheres at 0.4 in 1.0),
 *         B=U[-1e-3,1e-3]
 * a_m: 7*N3 doubles, b_m: N3 doubles (storage contract above).  Ghost rows:
 * A0=-1, off-diagonals 0, B=0. */
int mfix_leq_gen_system(int kind, int imax, int jmax, int kmax, uint64_t seed,
                        int nspheres, double *a_m, double *b_m);
/* U[0,1) noise array from the same counter-based generator. */
int mfix_leq_gen_uniform(int64_t n, uint64_t seed, uint64_t stream, double *out);

/* ---- batched solves (many independent systems, one launch sequence) ---- */
typedef struct mfix_leq_batch_item {
    int imax, jmax, kmax;        /* grid of this system (serial layout, 3-D) */
    char pc[4];                  /* "DIAG", "LINE" or "NONE" */
    char cmethod[4];             /* LINE sweep string, e.g. "RSRS" */
    double tol;
    int itmax;
    int mixed;                   /* 1 = FP32 inner + FP64 refinement */
    double *d_var, *d_a_m, *d_b_m; /* device arrays of this system */
    uint64_t noise_seed;
} mfix_leq_batch_item;

/* Solve every item; systems advance together, one fused launch per phase
 * covering all systems (captured in a HIP graph).  stats may be NULL. */
int mfix_leq_batch_solve(int nsys, const mfix_leq_batch_item *items,
                         mfix_leq_stats *stats, void *stream);

/* ---- SIMPLE-sweep batching ----------------------------------------------- */
/* SOLVE_LIN_EQ(VNAME,VNO,VAR,A_M,B_M,M,ITMAX,METHOD,SWEEP,TOL1,PC,IER)
 * (model/solve_lin_eq.f:17-195), deferred: the call is checked now (METHOD
 * other than 2 -> MFIX_LEQ_EMETHOD, a PC other than LINE/DIAG/NONE ->
 * MFIX_LEQ_EBADPC, as the reference stops there) and queued; it is solved at
 * the next mfix_leq_sweep_flush together with every other queued call, as one
 * batched solve, each on the grid current at its call (one launch per phase
 * for all systems).
 * This is for the SOLVE_LIN_EQ calls of one DO_ITERATION stage whose systems
 * do not depend on each other's solutions: SOLVE_VEL_STAR solves U, V and W
 * of every phase into temporaries from the same old fields
 * (model/solve_vel_star.f:86-99, :315-329), the species of SOLVE_SPECIES_EQ
 * and the scalars of SOLVE_SCALAR_EQ likewise.  d_var, d_a_m (A_M(:,:,M))
 * and d_b_m (B_M(:,M)) are device arrays that must stay untouched until the
 * flush, one set per queued call (the reference reuses one A_M work array
 * between equations; a batched stage keeps one per equation).  mixed = 1
 * solves that system in mixed precision.  *ier is written and ITER_TOT(VNO)
 * updated at the flush, as the call itself would have done.  At most 1024
 * calls are queued (MFIX_LEQ_EBUSY beyond).  No counterpart in the
 * reference (batching). */
int mfix_solve_lin_eq_deferred(const char *vname, int vname_len, int vno,
                               double *d_var, double *d_a_m, double *d_b_m,
                               int itmax, int method, const char *sweep,
                               double tol1, const char *pc, int mixed,
                               int *ier);
/* Solve every queued call in one batched solve and empty the queue (also
 * when it fails).  stats: NULL or room for mfix_leq_sweep_pending() entries,
 * in call order.  Returns 0, MFIX_LEQ_EDGTSV if a JK/IJ line sweep's DGTSV
 * failed in one of them (IERs are still written), or another error. */
int mfix_leq_sweep_flush(mfix_leq_stats *stats, void *stream);
int mfix_leq_sweep_pending(void);   /* calls queued since the last flush */
void mfix_leq_sweep_discard(void);  /* drop the queued calls unsolved */

/* Autotuner: pick the block shape of the fused kernels (timed on a DIAG
 * solve) and the lines per block of the red-black line kernel (timed on a
 * LINE solve) for one grid size; results are cached per (n3 class,
 * precision).  Returns the chosen threads-per-block and lines-per-block. */
int mfix_leq_autotune(int imax, int jmax, int kmax, int mixed, int *block,
                      int *lines_per_block);
/* Pin the launch shape the autotuner would record for this grid size and
 * precision: threads per block (256, 512 or 1024) and cells per thread
 * (1, 2 or 4; the 16-byte grouped kernels run when it is 16/element size).
 * No counterpart in the reference. */
int mfix_leq_set_shape(int imax, int jmax, int kmax, int mixed, int block_threads,
                       int cells_per_thread);
/* Pin the lines per block (64, 32 or 16; 0 = the default for the size) of
 * the red-black line kernel that LINE solves of this grid size and
 * precision use: each block solves that many J-lines of one colour, one
 * line per lane (the LEQ_MSOLVE/LEQ_IKSWEEP line solves,
 * model/leqsol_mod.f:851-921).  The results do not depend on it.  No
 * counterpart in the reference. */
int mfix_leq_set_line_block(int imax, int jmax, int kmax, int mixed, int lines_per_block);
/* Reliable-update trigger of the mixed-precision solves that do not set
 * opts->inner_tol (batched and served ones included): an FP64 reliable
 * update (true residual, solution accumulation) when the FP32 recursive
 * residual fell by `delta` since the last one; 0 restores the default,
 * 1e-4 below 2^23 cells and 0.1 from there on.  No counterpart in the
 * reference. */
int mfix_leq_set_mixed_delta(double delta);

/* ---- residual norms of an assembled equation ---------------------------- */
/* CALC_RESID_C / CALC_RESID_S / CALC_RESID_PP (model/calc_resid.f:16-178,
 * :195-371, :391-534), serial semantics.  var/a_m/b_m: the phase's VAR,
 * A_M(:,:,M), B_M(:,M) in the storage contract above; flag: MFIX's FLAG
 * array (INTEGER, DIMENSION_3; 1 = fluid cell, >= 100 walls).  Cells of the
 * first interior layers (IS_ON_myPE_wobnd) that are fluid are summed; for C
 * a wall neighbour stands for the cell itself (EAST_OF etc.,
 * model/set_index1a.f:85-222).  debug_resid = DEBUG_RESID
 * (model/run_mod.f:244): with 0, C and S return num/den/ncells only (resid
 * and ijk_resid 0, max_resid -1; the reference leaves RESID and IJK_RESID
 * unset) and PP returns RESID with the unnormalised maximum.  ijk_resid is
 * 1-based like the reference's IJK. */
typedef struct mfix_leq_resid {
    double num, den;             /* NUM, DEN */
    double resid;                /* RESID */
    double max_resid;            /* MAX_RESID */
    int ijk_resid;               /* IJK_RESID */
    int ncells;                  /* fluid cells counted (NCELLS) */
} mfix_leq_resid;

/* CALC_RESID_C(VAR, A_M, B_M, M, NUM, DEN, RESID, MAX_RESID, IJK_RESID),
 * model/calc_resid.f:16: host arrays (copied to the device) and device
 * pointers (stream may be 0). */
int mfix_calc_resid_c(const double *var, const double *a_m, const double *b_m,
                      const int *flag, int debug_resid, mfix_leq_resid *out);
int mfix_calc_resid_c_dev(const double *d_var, const double *d_a_m,
                          const double *d_b_m, const int *d_flag,
                          int debug_resid, mfix_leq_resid *out, void *stream);
/* CALC_RESID_S(..., TOL), model/calc_resid.f:195: cells with |VAR| <= tol
 * are skipped, and for a solids phase (M > 0) those with EP_s(:,M) <=
 * dil_ep_s (DIL_EP_s = 1e-4, model/toleranc_mod.f:24); ep_s = NULL for M = 0. */
int mfix_calc_resid_s(const double *var, const double *a_m, const double *b_m,
                      const int *flag, double tol, const double *ep_s,
                      double dil_ep_s, int debug_resid, mfix_leq_resid *out);
int mfix_calc_resid_s_dev(const double *d_var, const double *d_a_m,
                          const double *d_b_m, const int *d_flag, double tol,
                          const double *d_ep_s, double dil_ep_s,
                          int debug_resid, mfix_leq_resid *out, void *stream);
/* CALC_RESID_PP(B_M, NORM, NUM, DEN, RESID, MAX_RESID, IJK_RESID),
 * model/calc_resid.f:391: |B_M| of the pressure-correction equation. */
int mfix_calc_resid_pp(const double *b_m, const int *flag, double norm,
                       int debug_resid, mfix_leq_resid *out);
int mfix_calc_resid_pp_dev(const double *d_b_m, const int *d_flag, double norm,
                           int debug_resid, mfix_leq_resid *out, void *stream);

/* ---- the gas pressure-correction equation, assembled on the device -------- */
/* Fields of INIT_AB_M + CONV_PP_G + SOURCE_PP_G for phase 0: DIMENSION_3
 * device arrays in the storage contract above. */
typedef struct mfix_pp_g_fields {
    const double *rop_ge, *rop_gn, *rop_gt;   /* ROP_gE/N/T face bulk densities, model/mflux_mod.f:36-40 */
    const double *ayz, *axz, *axy;            /* face areas, model/geometry_mod.f:206-210 */
    const double *vol;                        /* cell volumes, model/geometry_mod.f:212 */
    const double *rop_g, *rop_go;             /* ROP_g and its old-time value, model/fldvar_mod.f:38-41 */
    const double *u_g, *v_g, *w_g;            /* gas velocities, model/fldvar_mod.f:87-111 */
    const double *sum_r_g;                    /* SUM_R_g, model/rxns_mod.f:28 */
    const double *d_e, *d_n, *d_t;            /* D_E(:,0), D_N(:,0), D_T(:,0), model/pgcor_mod.f:12-16 */
    const int *flag;                          /* FLAG (1 fluid, 10/11 pressure flow, >= 100 walls) */
    double odt;                               /* ODT = 1/DT, model/run_mod.f:54 */
    int ijk_p_g;                              /* IJK_P_g (1-based reference pressure cell), 0 = UNDEFINED_I */
} mfix_pp_g_fields;

/* INIT_AB_M(A_M,B_M,IJKMAX2,0) + CONV_PP_G(A_M,B_M) + SOURCE_PP_G(A_M,B_M,
 * B_MMAX) (model/init_ab_m.f:20, model/conv_pp_g.f:33, model/source_pp_g.f:18)
 * for the gas phase with MMAX = 0, RO_G0 defined (incompressible: no
 * COMPRESSIBLE_PP_G), no cut cells, no shear: writes A_M(:,:,0), B_M(:,0)
 * and, on fluid cells, B_MMAX(:,0).  *n_singular (may be NULL) receives the
 * number of fluid cells left with A0 = 0 and B != 0, where the reference
 * writes an error message and carries on. */
int mfix_assemble_pp_g_dev(const mfix_pp_g_fields *fields, double *d_a_m,
                           double *d_b_m, double *d_b_mmax, int *n_singular,
                           void *stream);

/* SOLVE_PP_G(NORMG, RESG, IER) (model/solve_pp_g.f:25) for the gas phase:
 * PP_G = 0, the assembly above, RESG = CALC_RESID_PP(B_M, NORMGloc) with
 * NORMGloc = NORMG, or RESID of B_MMAX / 10 when NORMG = 0, then
 * SOLVE_LIN_EQ('Pp_g', 1, PP_G, A_M, B_M, 0, ...) (METHOD 2 only) as
 * LEQ_BICGS; mixed = 1 runs it in mixed precision.  d_pp_g receives PP_G,
 * d_a_m/d_b_m the (row-scaled) system, ITER_TOT(1) is updated. */
int mfix_solve_pp_g_dev(const mfix_pp_g_fields *fields, double normg,
                        double *d_pp_g, double *d_a_m, double *d_b_m,
                        double *d_b_mmax, int itmax, int method,
                        const char *sweep, double tol, const char *pc,
                        int mixed, double *resg, int *ier,
                        mfix_leq_stats *stats, void *stream);

/* ---- scalar transport equations, first-order upwinding, on the device ---- */
/* Fields of CONV_DIF_PHI (model/conv_dif_phi.f:17) for one phase: DIMENSION_3
 * device arrays in the storage contract above, and the 1-D geometry arrays
 * of model/geometry_mod.f with Fortran index 1 first (I = 1..IMAX2 at
 * [0..IMAX2-1], likewise J, and K = 1..KMAX2, one value when NO_K). */
typedef struct mfix_phi_fields {
    const double *dif;                        /* DIF, the diffusion coefficient */
    const double *u, *v, *w;                  /* UF, VF, WF: the upwind side of each face */
    const double *flux_e, *flux_n, *flux_t;   /* FLUX_E/N/T face mass fluxes */
    const double *ayz, *axz, *axy;            /* face areas, model/geometry_mod.f:206-210 */
    const double *odx_e, *ox, *fx_e, *fx_e_bar;   /* along I */
    const double *ody_n, *fy_n, *fy_n_bar;        /* along J */
    const double *odz_t, *fz_t, *fz_t_bar;        /* along K */
    const int *flag;                          /* FLAG */
} mfix_phi_fields;

/* INIT_AB_M(A_M,B_M,IJKMAX2,M) + CONV_DIF_PHI(PHI,DIF,DISC=0,UF,VF,WF,FLUX_E,
 * FLUX_N,FLUX_T,M,A_M,B_M) with DEF_COR off (model/init_ab_m.f:20,
 * model/conv_dif_phi.f:17-80, :214-339): writes phase M's A_M(:,:,M) slice
 * (d_a_m, 7 planes) and B_M(:,M) (d_b_m, zero).  No cut cells, no internal
 * surfaces.  Returns once the kernel is queued on the stream. */
int mfix_conv_dif_phi_dev(const mfix_phi_fields *fields, double *d_a_m,
                          double *d_b_m, void *stream);

/* SOURCE_PHI(S_P,S_C,EP,PHI,M,A_M,B_M) (model/source_phi.f:25-139) in place
 * on phase M's slice, fluid rows only; no internal surfaces, and a cyclic
 * axis's ghost layers carry FLAG 106/107 as SET_FLAGS gives them (IS_AT_*
 * then never changes the result).  DIL_EP_s is model/toleranc_mod.f:24's
 * 1e-4 in MFIX.  Queued on the stream. */
int mfix_source_phi_dev(const double *d_s_p, const double *d_s_c,
                        const double *d_ep, const double *d_phi,
                        const int *d_flag, double dil_ep_s, double *d_a_m,
                        double *d_b_m, void *stream);

/* UNDER_RELAX_S(VAR,A_M,B_M,M,UR) (model/under_relax.f:20-81) in place on
 * phase M's slice.  UR = 0 (a division by zero in the reference) is
 * MFIX_LEQ_EINVAL.  Queued on the stream. */
int mfix_under_relax_s_dev(const double *d_var, const int *d_flag, double ur,
                           double *d_a_m, double *d_b_m, void *stream);

/* ---- serving: continuous batching with HBM-budget admission ------------- */
typedef struct mfix_leq_server mfix_leq_server;
typedef struct mfix_leq_request {
    int imax, jmax, kmax;
    char pc[4];
    char cmethod[4];
    double tol;
    int itmax;
    int mixed;
    const double *a_m, *b_m;     /* host arrays (copied at admission) */
    double *var;                 /* host in/out: initial guess, then solution */
    uint64_t noise_seed;
} mfix_leq_request;

int mfix_leq_server_create(size_t hbm_budget_bytes, int max_slots,
                           mfix_leq_server **out);
/* Non-blocking: queue a request; returns its id (>=0) or a negative code. */
int64_t mfix_leq_server_submit(mfix_leq_server *s, const mfix_leq_request *r);
/* Advance the device by one scheduling round (admit, iterate, retire). */
int mfix_leq_server_step(mfix_leq_server *s);
/* Query a request: returns 1 when done (stats filled, var written), 0 if
 * pending, or the negative code it failed with.  stats->device_ms is the
 * wall time from admission to retirement. */
int mfix_leq_server_poll(mfix_leq_server *s, int64_t id, mfix_leq_stats *stats);
int mfix_leq_server_pending(mfix_leq_server *s);
int mfix_leq_server_destroy(mfix_leq_server *s);

/* ---- device memory helpers (resident A_M/B_M/VAR for the _dev entries) -- */
int mfix_leq_dev_alloc(size_t bytes, void **out);
int mfix_leq_dev_free(void *p);
int mfix_leq_copy_to_dev(void *dst, const void *src, size_t bytes);
int mfix_leq_copy_from_dev(void *dst, const void *src, size_t bytes);
int mfix_leq_copy_dev(void *dst, const void *src, size_t bytes);
/* Plain streaming copy kernel of n doubles (HBM traffic calibration:
 * reads 8n and writes 8n bytes). */
int mfix_leq_copy_kernel_dev(double *dst, const double *src, int64_t n);
int mfix_leq_sync(void);

/* ---- parity / diagnostics hooks of the host entry ------------------------ */
/* Use this U[0,1) array (host, DIMENSION_3) for Rtilde in mfix_leq_bicgs
 * instead of the internal generator; NULL restores the generator (seed). */
int mfix_leq_set_noise(const double *host_noise, uint64_t seed);
/* 1 (default): mfix_leq_bicgs copies the row-scaled A_M/B_M back as the
 * reference leaves them; 0: only VAR comes back (no caller reads them). */
int mfix_leq_set_write_back(int on);
/* Statistics of the last mfix_leq_bicgs call. */
int mfix_leq_last_stats(mfix_leq_stats *st);
/* 1 (default): iterations of single and batched solves run as replays of a
 * captured HIP graph (chunks of 16); 0: direct launches.  Results are
 * bit-identical either way. */
int mfix_leq_set_graphs(int on);
/* 1 (default): a mixed-precision DIAG solve whose unit-scaled diagonal is
 * -1 in every row (1/A0 = -1 in FP32: each row's largest coefficient is its
 * diagonal) applies the preconditioner as a negation and does not read the
 * 1/A0 plane; 0: always read it.  Results are bit-identical either way.
 * No counterpart in the reference (a bandwidth switch of this library). */
int mfix_leq_set_unit_diag(int on);

/* ---- timing of the dominant kernel (HIP events of each dispatch) ------- */
/* Enable per-launch timing of kernel `which` (0 = first stencil pass of each
 * iteration, 1 = line preconditioner colour phase): the timed launches
 * record their own start and end in HIP events (hipExtLaunchKernelGGL), and
 * iterations run as direct launches instead of graph replays while it is
 * on.  Read back the launch count, the summed milliseconds and the
 * algorithmic bytes of the launches that worked.  Launches of a single solve issued after it was
 * done (they return at once) are counted apart: mfix_leq_prof_idle.
 * No counterpart in the reference (measurement of this library). */
int mfix_leq_prof_enable(int which, int on);
int mfix_leq_prof_read(int which, int64_t *launches, double *total_ms,
                       double *alg_bytes);
int mfix_leq_prof_idle(int which, int64_t *launches, double *total_ms);

#ifdef __cplusplus
}
#endif
#endif /* MFIX_LEQ_H */
