/*
 * mfcfd_hip.h -- C ABI of libmfcfd_hip.so, the MI355X implementation of the
 * q-LSKUM pseudo-time iteration of Pathlessbark8/3d-mfcfd.
 *
 * The reference's hot path is a set of argument-free Fortran module
 * procedures that read and write the module globals of data_structure_mod
 * (data_structure.f90:46-59) and the PARAMETERs of parameter_mod
 * (parameter.f90:8-33).  This ABI carries the same arrays in the same
 * layout -- column-major (m,N) arrays passed as plain pointers, 1-based
 * point numbers in every index array -- so a Fortran host binds it with
 * iso_c_binding (see fortran/mfcfd_dropin.f90) and keeps its own
 * allocatables.  Device copies live in HBM for the life of a context; host
 * arrays are refreshed only by mfcfd_download().
 *
 * Every function returns 0 on success or a negative MFCFD_E* code; the
 * message is available from mfcfd_last_error().  Nothing throws across the
 * ABI.  A context is not re-entrant: one host thread drives it, and calls
 * run in order on the context's HIP stream.
 */
#ifndef MFCFD_HIP_H
#define MFCFD_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define MFCFD_ABI_VERSION 6

/* error codes */
#define MFCFD_OK 0
#define MFCFD_EINVAL (-1)      /* bad argument or inconsistent arrays */
#define MFCFD_ESTATE (-2)      /* call out of order (e.g. no geometry bound) */
#define MFCFD_EHIP (-3)        /* HIP runtime error */
#define MFCFD_ENOMEM (-4)      /* device or host allocation failed */
#define MFCFD_EUNSUPPORTED (-5)

/* iteration forms */
#define MFCFD_EXPLICIT 0       /* U <- U - delt*R (state_update.f90:35,67) */
#define MFCFD_IMPLICIT 1       /* LU-SGS sweeps (implicit_forward_and_reverse_sweeps.f90
                                  with state_update.f90) */
/* the reference's alternative implicit pairing: the type-ordered sweeps that
 * also sweep outer points (implicit_forward_and_reverse_sweeps_backup.f90:
 * 16-141, with implicit_D_inv_outer.f90 and the *_outer sweep terms) with
 * state_update_implicit.f90:14-102 (outer Ubar + delUn, no wall U(4) = 0) */
#define MFCFD_IMPLICIT_BACKUP 2
#define MFCFD_FP64 0
/* MIXED: q, qm, dq, the q-derivative right-hand sides and the split fluxes
 * (flux residual and LU-SGS sweeps) in FP32; positions, least-squares
 * geometry and 3x3 solves, Cramer weights, the flux sums, D_inv, the state
 * update and the residual in FP64.  Fused flux order and sync-free sweeps
 * only, no second-order inner iterations. */
#define MFCFD_MIXED 1

/* array ids for upload/download; shapes are the reference's */
#define MFCFD_ARR_PRIM 0       /* prim(5,N) */
#define MFCFD_ARR_U 1          /* U(5,N) */
#define MFCFD_ARR_Q 2          /* q(5,N) */
#define MFCFD_ARR_QM 3         /* qm(2,5,N) */
#define MFCFD_ARR_DQ 4         /* dq(3,5,N) */
#define MFCFD_ARR_DELT 5       /* delt(N) */
#define MFCFD_ARR_FLUX_RES 6   /* flux_res(5,N) */
#define MFCFD_ARR_DELUP 7      /* delUp(5,N) */
#define MFCFD_ARR_DELUN 8      /* delUn(5,N) */
#define MFCFD_ARR_COUNT 9

typedef struct mfcfd_ctx mfcfd_ctx;

/* parameter.f90:8-33, runtime instead of compile time */
typedef struct mfcfd_params {
    double mach;             /* Mach */
    double aoa;              /* aoa, degrees */
    double cfl;              /* CFL */
    double vl_const;         /* VL_CONST (Venkatakrishnan constant) */
    double power;            /* power (least-squares weight exponent) */
    int inner_iterations;    /* inner_iterations (second-order q-derivatives) */
    int mode;                /* MFCFD_EXPLICIT, MFCFD_IMPLICIT or MFCFD_IMPLICIT_BACKUP */
    int precision;           /* MFCFD_FP64 or MFCFD_MIXED */
    int max_iters;           /* max_iters (used by the Fortran drop-in's final download) */
    int flux_order;          /* MFCFD_FLUX_FUSED (default) or MFCFD_FLUX_LISTS */
    int sweep_schedule;      /* MFCFD_SWEEP_ASYNC (default) or MFCFD_SWEEP_LEVELS */
} mfcfd_params;

/* flux residual evaluation order:
 *  FUSED  regular points walk conn once, one limited reconstruction per edge
 *         end, each list's least-squares solve applied edge by edge through
 *         its Cramer weights (same fluxes, different rounding);
 *  LISTS  every point walks each direction list and solves the accumulated
 *         3x3 system at the end, exactly as interior_flux_dGxpos.f90:61-123 */
#define MFCFD_FLUX_FUSED 0
#define MFCFD_FLUX_LISTS 1

/* schedule of the implicit LU-SGS sweeps; both visit each point's
 * dependencies (stencil neighbours earlier in the sweep's alias order,
 * implicit_forward_and_reverse_sweeps.f90:26-85) before the point itself:
 *  ASYNC   one persistent launch per sweep; workgroups take points in sweep
 *          order and wait on per-point completion flags;
 *  LEVELS  one launch per dependency level (host-built level sets). */
#define MFCFD_SWEEP_ASYNC 0
#define MFCFD_SWEEP_LEVELS 1

/* the point/connectivity arrays of data_structure.f90:11-28,49-56 as the
 * host holds them (all point numbers 1-based) */
typedef struct mfcfd_geometry {
    int n;                          /* max_points */
    const double *x, *y, *z;        /* (N) */
    const double *tan1, *tan2, *nor;/* (3,N) */
    const double *min_dist;         /* (N) */
    const int *status;              /* (N): 0 interior, 1 wall, 2 outer, 5/6 supersonic */
    const int *nbhs;                /* (N), each <= 27 */
    const int *conn;                /* (27,N) */
    const int *split_nbhs[6];       /* xpos,xneg,ypos,yneg,zpos,zneg _nbhs (N) */
    const int *split_conn[6];       /* xpos..zneg _conn (22,N), exactly as the host
                                       built them: lists longer than 22 continue
                                       into the next point's column */
    const int *alias;               /* (N) implicit_aliasing.f90 */
    const int *point_with_alias;    /* (N) */
} mfcfd_geometry;

typedef struct mfcfd_residual {
    double res_old, res_new, residue;   /* data_structure.f90:58 */
    double max_res;                     /* largest (U1 change)^2 */
    double sum_res_sqr;
    int max_res_point;                  /* 1-based */
    int iteration;                      /* t of the last fpi_solver call */
} mfcfd_residual;

/* ---- context ---------------------------------------------------------- */
int mfcfd_create(int device, mfcfd_ctx **out);
int mfcfd_destroy(mfcfd_ctx *ctx);
const char *mfcfd_last_error(void);
int mfcfd_abi_version(void);

/* parameter_mod (parameter.f90:8-33) */
int mfcfd_set_params(mfcfd_ctx *ctx, const mfcfd_params *p);

/* copies the point arrays to HBM once; replaces the host->module-global
 * handoff after read_input_point_data (point_preprocessor.f90:13-98),
 * generate_split_stencils (generate_connectivity.f90:12-34) and aliasing
 * (implicit_aliasing.f90:13-44).  Also builds the sweep schedule.
 * Returns MFCFD_EINVAL, naming the point, for an empty walked direction list
 * or a least-squares stencil (direction list or conn) whose 3x3 matrix is
 * singular -- coplanar or collinear neighbours, where the reference divides
 * by a zero determinant. */
int mfcfd_bind_geometry(mfcfd_ctx *ctx, const mfcfd_geometry *g);

/* the same binding with one argument per array, for hosts that pass arrays
 * by reference (a Fortran bind(C) interface passes point%x, point%tan1, ...
 * directly); split lists in xpos, xneg, ypos, yneg, zpos, zneg order */
int mfcfd_bind_geometry_arrays(mfcfd_ctx *ctx, int n, const double *x, const double *y,
                               const double *z, const double *tan1, const double *tan2,
                               const double *nor, const double *min_dist, const int *status,
                               const int *nbhs, const int *conn, const int *xpos_nbhs,
                               const int *xneg_nbhs, const int *ypos_nbhs, const int *yneg_nbhs,
                               const int *zpos_nbhs, const int *zneg_nbhs, const int *xpos_conn,
                               const int *xneg_conn, const int *ypos_conn, const int *yneg_conn,
                               const int *zpos_conn, const int *zneg_conn, const int *alias,
                               const int *point_with_alias);

/* host preprocessing drop-ins: generate_split_stencils
 * (generate_connectivity.f90:12-246) and aliasing (implicit_aliasing.f90:
 * 13-44) on the caller's arrays.  split_conn buffers must hold 22*N+32 ints
 * (lists longer than 22 spill as in the reference). */
int mfcfd_generate_split_stencils(int n, const double *x, const double *y, const double *z,
                                  const double *tan1, const double *tan2, const double *nor,
                                  const int *status, const int *nbhs, const int *conn,
                                  int *split_nbhs[6], int *split_conn[6]);
int mfcfd_aliasing(int n, const int *nbhs, const int *conn, int *alias, int *point_with_alias);

/* a batch of independent clouds iterated together (no counterpart in the
 * reference, which holds one cloud): the clouds are laid out one after the
 * other in device memory, every kernel launch covers all of them, and each
 * keeps its own Mach and angle of attack (the far-field state of
 * conserved_vector_Ubar.f90) and its own residual.  Host arrays of
 * upload/download then hold the clouds' points one cloud after the other.
 * The other parameters (CFL, mode, precision, ...) are shared. */
int mfcfd_bind_batch(mfcfd_ctx *ctx, int n_instances, const mfcfd_geometry *geoms,
                     const double *mach, const double *aoa);
/* number of clouds bound (1 for mfcfd_bind_geometry) and their point counts */
int mfcfd_batch_info(mfcfd_ctx *ctx, int *n_instances, int *counts);

/* state transfer by array id (MFCFD_ARR_*), reference layout.  An upload of
 * prim with a non-finite value, or a density or pressure not above zero,
 * returns MFCFD_EINVAL and leaves the device state unchanged (the reference
 * would carry the NaN into every stencil it touches). */
int mfcfd_upload(mfcfd_ctx *ctx, int array_id, const double *host);
int mfcfd_download(mfcfd_ctx *ctx, int array_id, double *host);

/* ---- the hot path: one call per reference entry point ----------------- */
int mfcfd_compute_conserved_vector(mfcfd_ctx *ctx); /* compute_conserved_vector.f90:13 */
int mfcfd_eval_q_variables(mfcfd_ctx *ctx);         /* q_variables.f90:10 */
int mfcfd_eval_q_derivatives(mfcfd_ctx *ctx);       /* q_derivatives.f90:10 */
int mfcfd_timestep_delt(mfcfd_ctx *ctx);            /* timestep_delt.f90:8 */
int mfcfd_eval_flux_residual(mfcfd_ctx *ctx);       /* flux_residual.f90:30 */
int mfcfd_implicit_forward_backward_sweeps(mfcfd_ctx *ctx); /* implicit_forward_and_reverse_sweeps.f90:16;
                                                       explicit mode: delUn = -delt*flux_res */
int mfcfd_state_update(mfcfd_ctx *ctx);             /* state_update.f90:15 */

/* fpi_solver.f90:20-45: the whole iteration plus the residue bookkeeping;
 * res_new/residue may be NULL */
int mfcfd_fpi_solver(mfcfd_ctx *ctx, int t, double *res_new, double *residue);

/* q_lskum.f90:28-32 without the printing: iterations t0..t0+n-1 with no host
 * round trip; per-iteration res_new and residue (either may be NULL) */
int mfcfd_run(mfcfd_ctx *ctx, int t0, int n_iters, double *res_new_hist, double *residue_hist);

/* mfcfd_run that allocates nothing (no counterpart in the reference): for
 * callers that cannot allocate once they are set up -- serving loops and
 * harnesses that forbid hipMalloc after binding.  mfcfd_run grows a device history buffer with hipMalloc when n_iters
 * exceeds every earlier call's; this variant keeps the per-iteration
 * residuals in a fixed device array of MFCFD_NOALLOC_MAX_ITERS doubles (one
 * per device, held until the process ends), returns MFCFD_EINVAL for longer
 * runs, and MFCFD_ESTATE when the implicit form's sweep plan has not been
 * built yet (mfcfd_sweep_info, mfcfd_fpi_solver or mfcfd_run builds it) or
 * kernel timing is on.  Results are those of mfcfd_run.  Calls on one device
 * share the array: concurrent calls from two contexts on one device race. */
#define MFCFD_NOALLOC_MAX_ITERS 4096
int mfcfd_run_noalloc(mfcfd_ctx *ctx, int t0, int n_iters, double *res_new_hist, double *residue_hist);

/* mfcfd_run with the iteration captured once as a HIP graph and replayed
 * n_iters times; res_new_hist as in mfcfd_run (the whole bound set), and
 * inst_res_hist[it*n_instances + b] the residual of cloud b alone
 * (sqrt(sum over its wall and interior points)/its point count).  Either
 * may be NULL. */
int mfcfd_run_graph(mfcfd_ctx *ctx, int t0, int n_iters, double *res_new_hist, double *inst_res_hist);

/* launch-shape autotuner (no counterpart in the reference).  Times, on the
 * bound cloud(s) and the current state, every compiled variant of the
 * gather kernels -- block size 64/128/256 x minimum waves per SIMD (the
 * register cap: 2/4/6 for the q-derivative kernel, 1..4 for the fused flux
 * kernels, 1..6 for their mixed-precision FP32 variants), then for the chosen shape an LDS reservation that caps the
 * resident waves per CU (fewer points in flight per XCD keep their
 * neighbours in that XCD's L2) -- and, in implicit mode, the sync-free
 * sweeps on 1, 1/2 or 1/4 of the resident grid; keeps the fastest.  Results do not depend on the
 * shapes.  It recomputes U, q, qm, dq, delt, flux_res and (implicit)
 * delUp/delUn from prim as one iteration's first steps would; prim is not
 * touched. */
typedef struct mfcfd_tuning {
    int flux_interior_block, flux_interior_waves;
    int flux_wall_block, flux_wall_waves;
    int qderiv_block, qderiv_waves;
    int sweep_div;
    int flux_interior_lds, flux_wall_lds, qderiv_lds;   /* dynamic LDS bytes per block (0: none) */
    double flux_interior_ms, flux_wall_ms, qderiv_ms, sweeps_ms;   /* chosen shapes */
    double flux_interior_default_ms, flux_wall_default_ms, qderiv_default_ms, sweeps_default_ms;
} mfcfd_tuning;
int mfcfd_autotune(mfcfd_ctx *ctx, int repeats, mfcfd_tuning *out);
/* apply shapes chosen earlier (times ignored) */
int mfcfd_set_tuning(mfcfd_ctx *ctx, const mfcfd_tuning *in);

int mfcfd_get_residual(mfcfd_ctx *ctx, mfcfd_residual *out);
int mfcfd_synchronize(mfcfd_ctx *ctx);

/* diagnostics: run n_iters iterations starting at t0 (advancing the state
 * exactly as mfcfd_run does) with HIP events recorded on the context's
 * stream between phases; out_ms[MFCFD_PHASE_*] receives the mean
 * milliseconds per iteration of each phase */
#define MFCFD_PHASE_CONSERVED_Q 0    /* compute_conserved_vector + eval_q_variables */
#define MFCFD_PHASE_QDERIV_DT 1      /* eval_q_derivatives + timestep_delt */
#define MFCFD_PHASE_FLUX_INTERIOR 2  /* eval_flux_residual, interior points */
#define MFCFD_PHASE_FLUX_BOUNDARY 3  /* eval_flux_residual, wall and outer points */
#define MFCFD_PHASE_SWEEPS 4         /* implicit_forward_backward_sweeps (or explicit form) */
#define MFCFD_PHASE_STATE_UPDATE 5   /* state_update + residual */
#define MFCFD_PHASE_COUNT 6
int mfcfd_time_phases(mfcfd_ctx *ctx, int t0, int n_iters, double *out_ms);

/* diagnostics: sweep schedule depth (levels of the forward and backward
 * Gauss-Seidel dependency chains) and swept point count */
int mfcfd_sweep_info(mfcfd_ctx *ctx, int *levels_fwd, int *levels_bwd, int *n_swept);

/* diagnostics: how many points take the fused flux kernel's interior class
 * (interior points whose lists partition conn), its general class (wall and
 * outer points, and interior points whose lists the 22-column spill
 * altered) and the list-by-list kernel (points whose conn and out-of-order
 * list entries together exceed 64) */
int mfcfd_point_classes(mfcfd_ctx *ctx, int *n_fused_interior, int *n_fused_wall, int *n_listwise);

/* diagnostics: the same classes per point, host order (N ints):
 * 0 fused interior, 1 fused general, 2 list-by-list */
int mfcfd_point_class_map(mfcfd_ctx *ctx, int *cls);

/* diagnostics: with timing enabled, every iteration records HIP events on
 * the context's stream around the launches of one kernel:
 *   MFCFD_TIME_FLUX_INTERIOR  the fused interior flux kernel (the explicit
 *                             form's dominant kernel);
 *   MFCFD_TIME_SWEEPS         each sync-free sweep kernel launch (forward
 *                             and backward; the implicit form's dominant one).
 * mfcfd_kernel_time returns the summed milliseconds and the number of timed
 * launches since mfcfd_kernel_timing was last called (0 disables timing) */
#define MFCFD_TIME_FLUX_INTERIOR 1
#define MFCFD_TIME_SWEEPS 2
int mfcfd_kernel_timing(mfcfd_ctx *ctx, int enable);
int mfcfd_kernel_time(mfcfd_ctx *ctx, double *total_ms, long long *launches);

#ifdef __cplusplus
}
#endif
#endif
