/*
 * mfcfd_oracle.h -- CPU restatement of the reference q-LSKUM iteration.
 *
 * TEST INFRASTRUCTURE ONLY.  This oracle is the checker for the HIP path:
 * only tests, the smoke check and the CPU-baseline leg of the benchmark load
 * it.  The product library never links or calls it.
 *
 * Layout follows data_structure.f90:7-44 exactly (Fortran column-major,
 * 1-based point numbers inside index arrays):
 *   prim/U/q/flux_res/delUp/delUn (5,N)  -> a[k*5 + r]
 *   qm (2,5,N)                           -> a[(k*5 + r)*2 + s]
 *   dq/temp (3,5,N)                      -> a[(k*5 + r)*3 + c]
 *   tan1/tan2/nor (3,N)                  -> a[k*3 + c]
 *   conn (27,N)                          -> a[k*27 + j]
 *   *_conn (22,N)                        -> a[k*22 + j]  (raw offset, j may
 *                                           exceed 21: entries spill into the
 *                                           next point's column as they do in
 *                                           the reference; buffers need
 *                                           22*N + ORC_SPLIT_SLACK entries)
 * with k the 0-based point row, r/s/c/j 0-based components.
 */
#ifndef MFCFD_ORACLE_H
#define MFCFD_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_CONN_COLS 27
#define ORC_SPLIT_COLS 22
#define ORC_SPLIT_SLACK 32

typedef struct orc_state {
    /* parameter.f90:8-33 (runtime here) */
    int n;                   /* max_points */
    int inner_iterations;
    double mach, aoa, cfl, vl_const, power;
    double pi, theta, rho_inf, u1_inf, u2_inf, u3_inf, pr_inf;

    /* geometry (data_structure.f90:11-28) */
    double *x, *y, *z, *tan1, *tan2, *nor, *min_dist;
    int *status, *nbhs, *conn;
    int *xpos_nbhs, *xneg_nbhs, *ypos_nbhs, *yneg_nbhs, *zpos_nbhs, *zneg_nbhs;
    int *xpos_conn, *xneg_conn, *ypos_conn, *yneg_conn, *zpos_conn, *zneg_conn;
    int *alias, *point_with_alias;

    /* point classes (data_structure.f90:49-56), 1-based point numbers */
    int interior_points, wall_points, outer_points;
    int supersonic_inlet_points, supersonic_outlet_points;
    int *interior_points_index, *wall_points_index, *outer_points_index;
    int *supersonic_inlet_points_index, *supersonic_outlet_points_index;

    /* flow state (data_structure.f90:32-41) */
    double *U, *delUp, *delUn, *prim, *flux_res, *q, *qm, *dq, *temp, *delt;

    /* residual scalars (data_structure.f90:58-59) */
    double res_old, res_new, max_res, residue;
    int max_res_point;
} orc_state;

/* derived freestream values (parameter.f90:19-33) from mach/aoa */
void orc_set_flow(orc_state *s, double mach, double aoa);

/* host preprocessing */
void orc_classify_points(orc_state *s);          /* point_preprocessor.f90:41-95 */
void orc_generate_split_stencils(orc_state *s);  /* generate_connectivity.f90:12-246 */
void orc_aliasing(orc_state *s);                 /* implicit_aliasing.f90:13-44 */

/* the iteration, one function per reference entry point */
void orc_compute_conserved_vector(orc_state *s); /* compute_conserved_vector.f90:13-34 */
void orc_eval_q_variables(orc_state *s);         /* q_variables.f90:10-40 */
void orc_eval_q_derivatives(orc_state *s);       /* q_derivatives.f90:10-204 */
void orc_timestep_delt(orc_state *s);            /* timestep_delt.f90:8-54 */
void orc_eval_flux_residual(orc_state *s);       /* flux_residual.f90:30-84 */
void orc_implicit_sweeps(orc_state *s);          /* implicit_forward_and_reverse_sweeps.f90:16-89 */
void orc_explicit_sweep(orc_state *s);           /* explicit form: delUn = -delt*flux_res */
void orc_state_update(orc_state *s);             /* state_update.f90:15-118 */
/* alternative implicit pairing (implicit_forward_and_reverse_sweeps_backup.f90
 * with state_update_implicit.f90) */
void orc_implicit_sweeps_backup(orc_state *s);   /* ..._backup.f90:16-141 */
void orc_state_update_implicit(orc_state *s);    /* state_update_implicit.f90:14-102 */
/* implicit_mode: 0 explicit, 1 canonical implicit, ORC_IMPLICIT_BACKUP the
 * alternative pairing */
#define ORC_IMPLICIT_BACKUP 2
void orc_fpi_solver(orc_state *s, int t, int implicit_mode); /* fpi_solver.f90:20-45 */
void orc_run(orc_state *s, int t0, int n_iters, int implicit_mode,
             double *res_new_hist, double *residue_hist);

/* single-call kernels exposed for unit tests */
void orc_flux_half(int axis, int sign, const double t1[3], const double t2[3],
                   const double n[3], const double prim[5], double pi, double G[5]);
void orc_flux_octant(int axis, int sign_a, int sign_n, const double t1[3],
                     const double t2[3], const double n[3], const double prim[5],
                     double pi, double G[5]);
void orc_venkat_limiter(const orc_state *s, const double qtilde[5], double phi[5], int k0);
void orc_qtilde_to_primitive(const double q[5], double prim[5]);
double orc_implicit_D_inv(const orc_state *s, int i0);
double orc_implicit_D_inv_outer(const orc_state *s, int i0); /* implicit_D_inv_outer.f90:17-363 */
void orc_conserved_vector_Ubar(const orc_state *s, int k0, double Ubar[5]);

#ifdef __cplusplus
}
#endif
#endif
