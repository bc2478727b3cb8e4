/*
 * mfcfd_oracle.c -- CPU restatement of the reference q-LSKUM pseudo-time
 * iteration (Pathlessbark8/3d-mfcfd), written from the reference's
 * algorithm, routine by routine, in the reference's evaluation order.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Only tests, the
 * smoke check and the CPU-baseline leg of the benchmark may load it.
 *
 * Floating-point contract: compile with -O2 -ffp-contract=off and no
 * -ffast-math, so that every Fortran expression below is evaluated with the
 * same association and rounding as the reference built by flang -O2 for the
 * default x86-64 target (no FMA).  exp/log/erf/sqrt come from the same glibc.
 * Fortran evaluates a+b+c as (a+b)+c and a*b*c as (a*b)*c; the C below is
 * written with the same left-to-right association.  "-a*b" in Fortran is
 * -(a*b), which rounds identically to (-a)*b.
 *
 * Parity pins (tests/test_oracle_*.py): bit-for-bit equality with the
 * reference compiled by flang (oracle/_ref, built by oracle/build_ref.sh) on
 * the golden clouds -- the canonical explicit and implicit forms, supersonic
 * inlet/outlet points in both forms, power = 2 with one inner q-derivative
 * iteration in both forms, and the alternative implicit pairing (outer
 * points swept, state_update_implicit.f90) -- and on the first iterations of
 * the 50k-point history; closed forms (KFVS and octant flux identities,
 * q <-> prim and U <-> prim round trips, WLS exactness on linear fields,
 * limiter special cases, freestream preservation, brute-force split stencils
 * and alias BFS).  Every function below is reached by at least one of the
 * bit-for-bit fixtures.
 */
#include "mfcfd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* limiters.f90:26,28 compare against the REAL(4) literal 10e-6 */
#define LIMITER_THRESHOLD ((double)10e-6f)

#define P5(a, k, r) ((a)[(size_t)(k) * 5 + (r)])
#define QM(s_, k, r, m) ((s_)->qm[((size_t)(k) * 5 + (r)) * 2 + (m)])
#define DQ(s_, k, r, c) ((s_)->dq[((size_t)(k) * 5 + (r)) * 3 + (c)])
#define TMP(s_, k, r, c) ((s_)->temp[((size_t)(k) * 5 + (r)) * 3 + (c)])
#define F3(a, k, c) ((a)[(size_t)(k) * 3 + (c)])
#define CONN(s_, k, j) ((s_)->conn[(size_t)(k) * 27 + (j)])

/* ------------------------------------------------------------------------ */
/* parameters                                                               */
/* ------------------------------------------------------------------------ */

void orc_set_flow(orc_state *s, double mach, double aoa)
{
    /* parameter.f90:13-33 */
    s->mach = mach;
    s->aoa = aoa;
    s->pi = 4.0 * atan(1.0);
    s->theta = aoa * s->pi / 180.0;
    s->rho_inf = 1.0;
    s->u1_inf = mach * cos(s->theta);
    s->u2_inf = mach * sin(s->theta);
    s->u3_inf = 0.0;
    s->pr_inf = 1.0 / 1.4;
}

/* ------------------------------------------------------------------------ */
/* host preprocessing                                                       */
/* ------------------------------------------------------------------------ */

void orc_classify_points(orc_state *s)
{
    /* point_preprocessor.f90:41-95; the index arrays must be sized by the
     * caller (counts are recomputed here and must agree). */
    int ci = 0, cw = 0, co = 0, csi = 0, cso = 0;
    for (int k = 0; k < s->n; ++k) {
        int st = s->status[k];
        if (st == 0) s->interior_points_index[ci++] = k + 1;
        else if (st == 1) s->wall_points_index[cw++] = k + 1;
        else if (st == 2) s->outer_points_index[co++] = k + 1;
        else if (st == 6) s->supersonic_outlet_points_index[cso++] = k + 1;
        else if (st == 5) s->supersonic_inlet_points_index[csi++] = k + 1;
    }
    s->interior_points = ci;
    s->wall_points = cw;
    s->outer_points = co;
    s->supersonic_inlet_points = csi;
    s->supersonic_outlet_points = cso;
}

static void project(const orc_state *s, int i, int k, double *dels, double *delt, double *deln)
{
    double delx = s->x[k] - s->x[i];
    double dely = s->y[k] - s->y[i];
    double delz = s->z[k] - s->z[i];
    *dels = delx * F3(s->tan1, i, 0) + dely * F3(s->tan1, i, 1) + delz * F3(s->tan1, i, 2);
    *delt = delx * F3(s->tan2, i, 0) + dely * F3(s->tan2, i, 1) + delz * F3(s->tan2, i, 2);
    *deln = delx * F3(s->nor, i, 0) + dely * F3(s->nor, i, 1) + delz * F3(s->nor, i, 2);
}

/* append k to a split list with the reference's raw column-major offset */
static void push(int *nbhs, int *list, int i, int k1)
{
    nbhs[i] += 1;
    int count = nbhs[i];
    list[(size_t)i * ORC_SPLIT_COLS + (count - 1)] = k1;
}

void orc_generate_split_stencils(orc_state *s)
{
    /* generate_connectivity.f90:18-31: interior, then wall, then outer */
    for (int m = 0; m < s->interior_points; ++m) {
        int i = s->interior_points_index[m] - 1;
        /* :51-56 */
        s->xpos_nbhs[i] = s->xneg_nbhs[i] = 0;
        s->ypos_nbhs[i] = s->yneg_nbhs[i] = 0;
        s->zpos_nbhs[i] = s->zneg_nbhs[i] = 0;
        for (int r = 0; r < s->nbhs[i]; ++r) {
            int k1 = CONN(s, i, r);
            double dels, delt, deln;
            project(s, i, k1 - 1, &dels, &delt, &deln);
            /* :72-106  (<= 0 / > 0) */
            if (dels <= 0.0) push(s->xpos_nbhs, s->xpos_conn, i, k1);
            if (dels > 0.0) push(s->xneg_nbhs, s->xneg_conn, i, k1);
            if (delt <= 0.0) push(s->ypos_nbhs, s->ypos_conn, i, k1);
            if (delt > 0.0) push(s->yneg_nbhs, s->yneg_conn, i, k1);
            if (deln <= 0.0) push(s->zpos_nbhs, s->zpos_conn, i, k1);
            if (deln > 0.0) push(s->zneg_nbhs, s->zneg_conn, i, k1);
        }
    }
    for (int m = 0; m < s->wall_points; ++m) {
        int i = s->wall_points_index[m] - 1;
        /* :128-132 (zpos_nbhs is not reset for wall points) */
        s->xpos_nbhs[i] = s->xneg_nbhs[i] = 0;
        s->ypos_nbhs[i] = s->yneg_nbhs[i] = 0;
        s->zneg_nbhs[i] = 0;
        for (int r = 0; r < s->nbhs[i]; ++r) {
            int k1 = CONN(s, i, r);
            double dels, delt, deln;
            project(s, i, k1 - 1, &dels, &delt, &deln);
            (void)deln;
            /* :148-174  (<= 0 / >= 0; zneg takes every neighbour) */
            if (dels <= 0.0) push(s->xpos_nbhs, s->xpos_conn, i, k1);
            if (dels >= 0.0) push(s->xneg_nbhs, s->xneg_conn, i, k1);
            if (delt <= 0.0) push(s->ypos_nbhs, s->ypos_conn, i, k1);
            if (delt >= 0.0) push(s->yneg_nbhs, s->yneg_conn, i, k1);
            push(s->zneg_nbhs, s->zneg_conn, i, k1);
        }
    }
    for (int m = 0; m < s->outer_points; ++m) {
        int i = s->outer_points_index[m] - 1;
        /* :196-200 (zpos_nbhs is incremented from its previous value) */
        s->xpos_nbhs[i] = s->xneg_nbhs[i] = 0;
        s->ypos_nbhs[i] = s->yneg_nbhs[i] = 0;
        s->zneg_nbhs[i] = 0;
        for (int r = 0; r < s->nbhs[i]; ++r) {
            int k1 = CONN(s, i, r);
            double dels, delt, deln;
            project(s, i, k1 - 1, &dels, &delt, &deln);
            (void)deln;
            /* :216-242 */
            if (dels <= 0.0) push(s->xpos_nbhs, s->xpos_conn, i, k1);
            if (dels >= 0.0) push(s->xneg_nbhs, s->xneg_conn, i, k1);
            if (delt <= 0.0) push(s->ypos_nbhs, s->ypos_conn, i, k1);
            if (delt >= 0.0) push(s->yneg_nbhs, s->yneg_conn, i, k1);
            push(s->zpos_nbhs, s->zpos_conn, i, k1);
        }
    }
}

void orc_aliasing(orc_state *s)
{
    /* implicit_aliasing.f90:13-44.  alias(1)=1 but count starts at 0, so the
     * first new neighbour of point 1 also receives alias 1 and replaces point
     * 1 in point_with_alias(1).  Entries of point_with_alias past the last
     * reached point stay 0; the reference then reads nbhs(0), which lies in
     * the allocator's header and reads 0, so those entries visit nothing. */
    int n = s->n;
    for (int i = 0; i < n; ++i) {
        s->alias[i] = 0;
        s->point_with_alias[i] = 0;
    }
    s->alias[0] = 1;
    s->point_with_alias[0] = 1;
    int count = 0;
    for (int i = 0; i < n; ++i) {
        int p = s->point_with_alias[i];
        if (p == 0) continue;
        for (int j = 0; j < s->nbhs[p - 1]; ++j) {
            int k = CONN(s, p - 1, j);
            if (s->alias[k - 1] == 0) {
                count += 1;
                s->alias[k - 1] = count;
                s->point_with_alias[count - 1] = k;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* per-point state                                                          */
/* ------------------------------------------------------------------------ */

void orc_compute_conserved_vector(orc_state *s)
{
    /* compute_conserved_vector.f90:20-32 */
    for (int k = 0; k < s->n; ++k) {
        const double *p = &P5(s->prim, k, 0);
        double *U = &P5(s->U, k, 0);
        U[0] = p[0];
        double temp1 = p[0] * p[1];
        double temp2 = p[0] * p[2];
        double temp3 = p[0] * p[3];
        U[1] = temp1 * F3(s->tan1, k, 0) + temp2 * F3(s->tan1, k, 1) + temp3 * F3(s->tan1, k, 2);
        U[2] = temp1 * F3(s->tan2, k, 0) + temp2 * F3(s->tan2, k, 1) + temp3 * F3(s->tan2, k, 2);
        U[3] = temp1 * F3(s->nor, k, 0) + temp2 * F3(s->nor, k, 1) + temp3 * F3(s->nor, k, 2);
        temp1 = p[1] * p[1] + p[2] * p[2] + p[3] * p[3];
        U[4] = 2.5 * p[4] + 0.5 * U[0] * temp1;
    }
}

void orc_eval_q_variables(orc_state *s)
{
    /* q_variables.f90:18-38 */
    for (int k = 0; k < s->n; ++k) {
        const double *p = &P5(s->prim, k, 0);
        double *q = &P5(s->q, k, 0);
        double rho = p[0], u1 = p[1], u2 = p[2], u3 = p[3], pr = p[4];
        double beta = 0.5 * rho / pr;
        q[0] = log(rho) + (log(beta) * 2.5) - beta * (u1 * u1 + u2 * u2 + u3 * u3);
        double two_times_beta = 2.0 * beta;
        q[1] = two_times_beta * u1;
        q[2] = two_times_beta * u2;
        q[3] = two_times_beta * u3;
        q[4] = -two_times_beta;
    }
}

/* 3x3 Cramer pieces shared by every weighted-least-squares solve in the
 * reference (q_derivatives.f90:85-107, *_flux_dG*.f90:115-123,
 * implicit_D_inv_*.f90:99-105). */
typedef struct {
    double xx, yy, zz, xy, yz, zx;
} geo6;

static double cramer_det(const geo6 *g)
{
    return g->xx * (g->yy * g->zz - g->yz * g->yz)
         - g->xy * (g->xy * g->zz - g->yz * g->zx)
         + g->zx * (g->xy * g->yz - g->yy * g->zx);
}

/* numerator of component 0/1/2 of the solution for right-hand side (a,b,c) */
static double cramer_num(const geo6 *g, int col, double a, double b, double c)
{
    if (col == 0)
        return a * (g->yy * g->zz - g->yz * g->yz)
             - b * (g->xy * g->zz - g->zx * g->yz)
             + c * (g->xy * g->yz - g->zx * g->yy);
    if (col == 1)
        return g->xx * (b * g->zz - g->yz * c)
             - g->xy * (a * g->zz - g->zx * c)
             + g->zx * (a * g->yz - g->zx * b);
    return g->xx * (g->yy * c - b * g->yz)
         - g->xy * (g->xy * c - a * g->yz)
         + g->zx * (g->xy * b - a * g->yy);
}

/* dist**power (q_derivatives.f90:63, interior_flux_dGxpos.f90:78, ...) as
 * the flang build evaluates it.  parameter.f90:21 makes power a compile-time
 * constant, and flang's LLVM reduces the constant exponents 0, 0.5, 1, 2 and
 * 3 to 1, sqrt(d), d, d*d and (d*d)*d (x86 code of the reference objects
 * built here); any other exponent calls libm's pow.  glibc's pow(d, 2) is
 * not always d*d rounded once, so the reduction matters bit for bit. */
static double ref_pow(double d, double power)
{
    if (power == 0.0) return 1.0;
    if (power == 0.5) return sqrt(d);
    if (power == 1.0) return d;
    if (power == 2.0) return d * d;
    if (power == 3.0) return d * d * d;
    return pow(d, power);
}

void orc_eval_q_derivatives(orc_state *s)
{
    /* q_derivatives.f90:17-110 */
    for (int i = 0; i < s->n; ++i) {
        double x_i = s->x[i], y_i = s->y[i], z_i = s->z[i];
        geo6 g = {0, 0, 0, 0, 0, 0};
        double sx[5] = {0}, sy[5] = {0}, sz[5] = {0};
        for (int r = 0; r < 5; ++r) {
            QM(s, i, r, 0) = P5(s->q, i, r);
            QM(s, i, r, 1) = P5(s->q, i, r);
        }
        for (int kk = 0; kk < s->nbhs[i]; ++kk) {
            int nbh = CONN(s, i, kk) - 1;
            for (int r = 0; r < 5; ++r) {
                if (P5(s->q, nbh, r) > QM(s, i, r, 0)) QM(s, i, r, 0) = P5(s->q, nbh, r);
                if (P5(s->q, nbh, r) < QM(s, i, r, 1)) QM(s, i, r, 1) = P5(s->q, nbh, r);
            }
            double delx = s->x[nbh] - x_i;
            double dely = s->y[nbh] - y_i;
            double delz = s->z[nbh] - z_i;
            double dist = sqrt(delx * delx + dely * dely + delz * delz);
            double weights = 1.0 / ref_pow(dist, s->power);
            g.xx = g.xx + delx * delx * weights;
            g.yy = g.yy + dely * dely * weights;
            g.zz = g.zz + delz * delz * weights;
            g.xy = g.xy + delx * dely * weights;
            g.yz = g.yz + dely * delz * weights;
            g.zx = g.zx + delz * delx * weights;
            for (int r = 0; r < 5; ++r) {
                double t = P5(s->q, nbh, r) - P5(s->q, i, r);
                sx[r] = sx[r] + weights * delx * t;
                sy[r] = sy[r] + weights * dely * t;
                sz[r] = sz[r] + weights * delz * t;
            }
        }
        double det = cramer_det(&g);
        double one_by_det = 1.0 / det;
        for (int c = 0; c < 3; ++c)
            for (int r = 0; r < 5; ++r)
                DQ(s, i, r, c) = cramer_num(&g, c, sx[r], sy[r], sz[r]) * one_by_det;
    }

    /* q_derivatives.f90:116-201 (inner_iterations > 0 only) */
    for (int it = 0; it < s->inner_iterations; ++it) {
        for (int i = 0; i < s->n; ++i) {
            double x_i = s->x[i], y_i = s->y[i], z_i = s->z[i];
            geo6 g = {0, 0, 0, 0, 0, 0};
            double sx[5] = {0}, sy[5] = {0}, sz[5] = {0};
            for (int kk = 0; kk < s->nbhs[i]; ++kk) {
                int nbh = CONN(s, i, kk) - 1;
                double delx = s->x[nbh] - x_i;
                double dely = s->y[nbh] - y_i;
                double delz = s->z[nbh] - z_i;
                double dist = sqrt(delx * delx + dely * dely + delz * delz);
                double weights = 1.0 / ref_pow(dist, s->power);
                g.xx = g.xx + delx * delx * weights;
                g.yy = g.yy + dely * dely * weights;
                g.zz = g.zz + delz * delz * weights;
                g.xy = g.xy + delx * dely * weights;
                g.yz = g.yz + dely * delz * weights;
                g.zx = g.zx + delz * delx * weights;
                for (int r = 0; r < 5; ++r) {
                    double qti = P5(s->q, i, r) - 0.5 * (delx * DQ(s, i, r, 0) + dely * DQ(s, i, r, 1) + delz * DQ(s, i, r, 2));
                    double qtn = P5(s->q, nbh, r) - 0.5 * (delx * DQ(s, nbh, r, 0) + dely * DQ(s, nbh, r, 1) + delz * DQ(s, nbh, r, 2));
                    double t = qtn - qti;
                    sx[r] = sx[r] + weights * delx * t;
                    sy[r] = sy[r] + weights * dely * t;
                    sz[r] = sz[r] + weights * delz * t;
                }
            }
            double det = cramer_det(&g);
            double one_by_det = 1.0 / det;
            for (int c = 0; c < 3; ++c)
                for (int r = 0; r < 5; ++r)
                    TMP(s, i, r, c) = cramer_num(&g, c, sx[r], sy[r], sz[r]) * one_by_det;
        }
        for (int i = 0; i < s->n; ++i)
            for (int r = 0; r < 5; ++r)
                for (int c = 0; c < 3; ++c) DQ(s, i, r, c) = TMP(s, i, r, c);
    }
}

void orc_timestep_delt(orc_state *s)
{
    /* timestep_delt.f90:13-38 */
    for (int i = 0; i < s->n; ++i) {
        double min_delt = 1.0;
        for (int r = 0; r < s->nbhs[i]; ++r) {
            int k = CONN(s, i, r) - 1;
            double rho = P5(s->prim, k, 0), u1 = P5(s->prim, k, 1), u2 = P5(s->prim, k, 2);
            double u3 = P5(s->prim, k, 3), pr = P5(s->prim, k, 4);
            double delx = s->x[k] - s->x[i];
            double dely = s->y[k] - s->y[i];
            double delz = s->z[k] - s->z[i];
            double dist = sqrt(delx * delx + dely * dely + delz * delz);
            double mod_u = sqrt(u1 * u1 + u2 * u2 + u3 * u3);
            double delt = s->cfl * dist / (mod_u + 3.0 * sqrt(pr / rho));
            if (delt < min_delt) min_delt = delt;
        }
        s->delt[i] = min_delt;
    }
}

/* ------------------------------------------------------------------------ */
/* kinetic split fluxes                                                     */
/* ------------------------------------------------------------------------ */

/*
 * Half-range KFVS flux along axis (0: tan1, 1: tan2, 2: nor) for the
 * half-space sign (+1 positive / -1 negative), split_fluxes.f90:11-308.
 * Every "- X" of the negative variants equals "+ (-X)" in IEEE arithmetic,
 * so the sign factors below reproduce the reference bit for bit.
 */
void orc_flux_half(int axis, int sign, const double t1[3], const double t2[3],
                   const double n[3], const double prim[5], double pi, double G[5])
{
    double rho = prim[0], u1 = prim[1], u2 = prim[2], u3 = prim[3], pr = prim[4];
    double ut1 = t1[0] * u1 + t1[1] * u2 + t1[2] * u3;
    double ut2 = t2[0] * u1 + t2[1] * u2 + t2[2] * u3;
    double un = n[0] * u1 + n[1] * u2 + n[2] * u3;
    double beta = 0.5 * rho / pr;
    double ua = axis == 0 ? ut1 : (axis == 1 ? ut2 : un);
    double S = ua * sqrt(beta);
    double B = 0.5 * exp(-S * S) / sqrt(pi * beta);
    double e = erf(S);
    double A = 0.5 * (sign > 0 ? 1.0 + e : 1.0 - e);
    double Bs = sign > 0 ? B : -B;
    double u_sqr = ut1 * ut1 + ut2 * ut2 + un * un;
    double temp1 = (ua * A + Bs);
    G[0] = rho * temp1;
    G[1] = rho * ut1 * temp1;
    G[2] = rho * ut2 * temp1;
    G[3] = rho * un * temp1;
    G[1 + axis] = pr * A + rho * ua * temp1;
    temp1 = 2.5 * pr + 0.5 * rho * u_sqr;
    double tail = (temp1 + 0.5 * pr) * B;
    G[4] = (temp1 + pr) * ua * A + (sign > 0 ? tail : -tail);
}

/*
 * Quarter-range ("octant") fluxes of octant_fluxes.f90:9-442:
 *   axis 0/1 (tan1/tan2), sign_a the half-space along that axis,
 *   sign_n = -1 for the wall family Gw (A3neg, "- B3"),
 *   sign_n = +1 for the outer family Go (A3pos, "+ B3").
 */
void orc_flux_octant(int axis, int sign_a, int sign_n, const double t1[3],
                     const double t2[3], const double n[3], const double prim[5],
                     double pi, double G[5])
{
    double rho = prim[0], u1 = prim[1], u2 = prim[2], u3 = prim[3], pr = prim[4];
    double ut1 = t1[0] * u1 + t1[1] * u2 + t1[2] * u3;
    double ut2 = t2[0] * u1 + t2[1] * u2 + t2[2] * u3;
    double un = n[0] * u1 + n[1] * u2 + n[2] * u3;
    double beta = 0.5 * rho / pr;
    double ua = axis == 0 ? ut1 : ut2;
    double Sa = ua * sqrt(beta);
    double S3 = un * sqrt(beta);
    double Ba = 0.5 * exp(-Sa * Sa) / sqrt(pi * beta);
    double B3 = 0.5 * exp(-S3 * S3) / sqrt(pi * beta);
    double ea = erf(Sa);
    double Aa = 0.5 * (sign_a > 0 ? 1.0 + ea : 1.0 - ea);
    double e3 = erf(S3);
    double A3 = 0.5 * (sign_n > 0 ? 1.0 + e3 : 1.0 - e3);
    double u_sqr = ut1 * ut1 + ut2 * ut2 + un * un;
    double temp1 = (ua * Aa + (sign_a > 0 ? Ba : -Ba));
    G[0] = rho * temp1 * A3;
    if (axis == 0) {
        G[1] = (pr * Aa + rho * ut1 * temp1) * A3;
        G[2] = rho * ut2 * temp1 * A3;
    } else {
        G[1] = rho * ut1 * temp1 * A3;
        G[2] = (pr * Aa + rho * ut2 * temp1) * A3;
    }
    G[3] = rho * temp1 * (un * A3 + (sign_n > 0 ? B3 : -B3));
    double temp2 = 2.5 * pr + 0.5 * rho * u_sqr;
    double tb = (temp2 + 0.5 * pr) * Ba;
    double temp3 = (temp2 + pr) * ua * Aa + (sign_a > 0 ? tb : -tb);
    double tn = 0.5 * rho * un * B3 * temp1;
    G[4] = temp3 * A3 + (sign_n > 0 ? tn : -tn);
}

/* ------------------------------------------------------------------------ */
/* limiter and q -> primitive                                               */
/* ------------------------------------------------------------------------ */

void orc_venkat_limiter(const orc_state *s, const double qtilde[5], double phi[5], int k)
{
    /* limiters.f90:18-40 */
    for (int r = 0; r < 5; ++r) {
        double q = P5(s->q, k, r);
        double del_neg = qtilde[r] - q;
        if (fabs(del_neg) <= LIMITER_THRESHOLD) {
            phi[r] = 1.0;
        } else if (fabs(del_neg) > LIMITER_THRESHOLD) {
            double del_pos = 0.0;
            if (del_neg > 0.0) del_pos = QM(s, k, r, 0) - q;
            else if (del_neg < 0.0) del_pos = QM(s, k, r, 1) - q;
            double epsi = s->vl_const * s->min_dist[k];
            epsi = epsi * epsi * epsi; /* epsi**3.0d0 lowers to (e*e)*e */
            double num = (del_pos * del_pos) + (epsi * epsi);
            num = num * del_neg + 2.0 * del_neg * del_neg * del_pos;
            double den = del_pos * del_pos + 2.0 * del_neg * del_neg;
            den = den + del_neg * del_pos + epsi * epsi;
            den = den * del_neg;
            double temp = num / den;
            phi[r] = temp < 1.0 ? temp : 1.0;
        }
    }
}

void orc_qtilde_to_primitive(const double q[5], double prim[5])
{
    /* qvariables_to_primitive_variables.f90:16-30 */
    double beta = -q[4] * 0.5;
    double temp = 0.5 / beta;
    prim[1] = q[1] * temp;
    prim[2] = q[2] * temp;
    prim[3] = q[3] * temp;
    double temp1 = q[0] + beta * (prim[1] * prim[1] + prim[2] * prim[2] + prim[3] * prim[3]);
    double temp2 = temp1 - (log(beta) * 2.5);
    prim[0] = exp(temp2);
    prim[4] = prim[0] * temp;
}

/* ------------------------------------------------------------------------ */
/* flux stencils                                                            */
/* ------------------------------------------------------------------------ */

enum flux_kind { FK_HALF, FK_WALL, FK_OUTER };

typedef struct {
    const int *list;   /* split list (stride 22) or conn (stride 27) */
    int stride;
    int count;
    enum flux_kind kind;
    int axis;          /* 0: x (tan1), 1: y (tan2), 2: z (nor) */
    int sign;          /* +1 pos, -1 neg */
    int by_one_by_det; /* outer stencils: G = temp*(1/det) */
} stencil_t;

static void eval_flux(const stencil_t *st, const double t1[3], const double t2[3],
                      const double n[3], const double prim[5], double pi, double G[5])
{
    if (st->kind == FK_HALF)
        orc_flux_half(st->axis, st->sign, t1, t2, n, prim, pi, G);
    else
        orc_flux_octant(st->axis, st->sign, st->kind == FK_OUTER ? 1 : -1, t1, t2, n, prim, pi, G);
}

/* reconstruct q at one end of edge (i,k) with the limiter of point `at` and
 * evaluate the directional flux in the frame of i.
 * interior_flux_dGxpos.f90:92-104 */
static void edge_end_flux(const orc_state *s, const stencil_t *st, int at, double delx,
                          double dely, double delz, const double t1[3], const double t2[3],
                          const double n[3], double G[5])
{
    double temp[5], qtilde[5], phi[5] = {1, 1, 1, 1, 1}, prim[5];
    for (int r = 0; r < 5; ++r) {
        temp[r] = delx * DQ(s, at, r, 0) + dely * DQ(s, at, r, 1) + delz * DQ(s, at, r, 2);
        qtilde[r] = P5(s->q, at, r) - 0.5 * temp[r];
    }
    orc_venkat_limiter(s, qtilde, phi, at);
    for (int r = 0; r < 5; ++r) qtilde[r] = P5(s->q, at, r) - 0.5 * phi[r] * temp[r];
    orc_qtilde_to_primitive(qtilde, prim);
    eval_flux(st, t1, t2, n, prim, s->pi, G);
}

/* one directional flux derivative, e.g. interior_dGx_pos
 * (interior_flux_dGxpos.f90:20-126); wall and outer variants differ only in
 * the flux family, the neighbour list and the final division (outer:
 * temp*one_by_det, outer_flux_dGxpos.f90:118-124). */
static void stencil_dG(const orc_state *s, int i, const stencil_t *st, double G[5])
{
    geo6 g = {0, 0, 0, 0, 0, 0};
    double sx[5] = {0}, sy[5] = {0}, sz[5] = {0};
    double x_i = s->x[i], y_i = s->y[i], z_i = s->z[i];
    double t1[3] = {F3(s->tan1, i, 0), F3(s->tan1, i, 1), F3(s->tan1, i, 2)};
    double t2[3] = {F3(s->tan2, i, 0), F3(s->tan2, i, 1), F3(s->tan2, i, 2)};
    double n[3] = {F3(s->nor, i, 0), F3(s->nor, i, 1), F3(s->nor, i, 2)};
    for (int j = 0; j < st->count; ++j) {
        int k = st->list[(size_t)i * st->stride + j] - 1;
        double delx = s->x[k] - x_i;
        double dely = s->y[k] - y_i;
        double delz = s->z[k] - z_i;
        double dels = delx * t1[0] + dely * t1[1] + delz * t1[2];
        double delt = delx * t2[0] + dely * t2[1] + delz * t2[2];
        double deln = delx * n[0] + dely * n[1] + delz * n[2];
        double dist = sqrt(dels * dels + delt * delt + deln * deln);
        double weights = 1.0 / ref_pow(dist, s->power);
        double dels_w = dels * weights, delt_w = delt * weights, deln_w = deln * weights;
        g.xx = g.xx + dels * dels_w;
        g.yy = g.yy + delt * delt_w;
        g.zz = g.zz + deln * deln_w;
        g.xy = g.xy + dels * delt_w;
        g.yz = g.yz + delt * deln_w;
        g.zx = g.zx + deln * dels_w;
        double G_i[5], G_k[5];
        edge_end_flux(s, st, i, delx, dely, delz, t1, t2, n, G_i);
        edge_end_flux(s, st, k, delx, dely, delz, t1, t2, n, G_k);
        for (int r = 0; r < 5; ++r) {
            double t = G_k[r] - G_i[r];
            sx[r] = sx[r] + t * dels_w;
            sy[r] = sy[r] + t * delt_w;
            sz[r] = sz[r] + t * deln_w;
        }
    }
    double det = cramer_det(&g);
    if (st->by_one_by_det) {
        double one_by_det = 1.0 / det;
        for (int r = 0; r < 5; ++r) G[r] = cramer_num(&g, st->axis, sx[r], sy[r], sz[r]) * one_by_det;
    } else {
        for (int r = 0; r < 5; ++r) G[r] = cramer_num(&g, st->axis, sx[r], sy[r], sz[r]) / det;
    }
}

static stencil_t split_stencil(const orc_state *s, int i, int axis, int sign, enum flux_kind kind)
{
    stencil_t st;
    const int *lists[3][2] = {{s->xpos_conn, s->xneg_conn}, {s->ypos_conn, s->yneg_conn},
                              {s->zpos_conn, s->zneg_conn}};
    const int *counts[3][2] = {{s->xpos_nbhs, s->xneg_nbhs}, {s->ypos_nbhs, s->yneg_nbhs},
                               {s->zpos_nbhs, s->zneg_nbhs}};
    int which = sign > 0 ? 0 : 1;
    st.list = lists[axis][which];
    st.stride = ORC_SPLIT_COLS;
    st.count = counts[axis][which][i];
    st.kind = kind;
    st.axis = axis;
    st.sign = sign;
    st.by_one_by_det = kind == FK_OUTER || (kind == FK_HALF && 0);
    return st;
}

static stencil_t full_conn_stencil(const orc_state *s, int i, int axis, int sign, enum flux_kind kind)
{
    stencil_t st;
    st.list = s->conn;
    st.stride = ORC_CONN_COLS;
    st.count = s->nbhs[i];
    st.kind = kind;
    st.axis = axis;
    st.sign = sign;
    st.by_one_by_det = 0;
    return st;
}

void orc_eval_flux_residual(orc_state *s)
{
    double Gxp[5], Gxn[5], Gyp[5], Gyn[5], Gzp[5], Gzn[5];
    /* flux_residual.f90:40-52 wall */
    for (int m = 0; m < s->wall_points; ++m) {
        int k = s->wall_points_index[m] - 1;
        stencil_t st;
        st = split_stencil(s, k, 0, +1, FK_WALL); stencil_dG(s, k, &st, Gxp);
        st = split_stencil(s, k, 0, -1, FK_WALL); stencil_dG(s, k, &st, Gxn);
        st = split_stencil(s, k, 1, +1, FK_WALL); stencil_dG(s, k, &st, Gyp);
        st = split_stencil(s, k, 1, -1, FK_WALL); stencil_dG(s, k, &st, Gyn);
        /* wall_flux_dGzneg.f90:61-65 iterates the full conn with flux_Gzn */
        st = full_conn_stencil(s, k, 2, -1, FK_HALF); stencil_dG(s, k, &st, Gzn);
        for (int r = 0; r < 5; ++r)
            P5(s->flux_res, k, r) = 2.0 * (Gxp[r] + Gxn[r] + Gyp[r] + Gyn[r] + Gzn[r]);
    }
    /* :54-66 outer */
    for (int m = 0; m < s->outer_points; ++m) {
        int k = s->outer_points_index[m] - 1;
        stencil_t st;
        st = split_stencil(s, k, 0, +1, FK_OUTER); stencil_dG(s, k, &st, Gxp);
        st = split_stencil(s, k, 0, -1, FK_OUTER); stencil_dG(s, k, &st, Gxn);
        st = split_stencil(s, k, 1, +1, FK_OUTER); stencil_dG(s, k, &st, Gyp);
        st = split_stencil(s, k, 1, -1, FK_OUTER); stencil_dG(s, k, &st, Gyn);
        /* outer_flux_dGzpos.f90: zpos list with flux_Gzp and temp*one_by_det */
        st = split_stencil(s, k, 2, +1, FK_HALF); st.by_one_by_det = 1;
        stencil_dG(s, k, &st, Gzp);
        for (int r = 0; r < 5; ++r)
            P5(s->flux_res, k, r) = (Gxp[r] + Gxn[r] + Gyp[r] + Gyn[r] + Gzp[r]);
    }
    /* :68-81 interior */
    for (int m = 0; m < s->interior_points; ++m) {
        int k = s->interior_points_index[m] - 1;
        stencil_t st;
        st = split_stencil(s, k, 0, +1, FK_HALF); stencil_dG(s, k, &st, Gxp);
        st = split_stencil(s, k, 0, -1, FK_HALF); stencil_dG(s, k, &st, Gxn);
        st = split_stencil(s, k, 1, +1, FK_HALF); stencil_dG(s, k, &st, Gyp);
        st = split_stencil(s, k, 1, -1, FK_HALF); stencil_dG(s, k, &st, Gyn);
        st = split_stencil(s, k, 2, +1, FK_HALF); stencil_dG(s, k, &st, Gzp);
        st = split_stencil(s, k, 2, -1, FK_HALF); stencil_dG(s, k, &st, Gzn);
        for (int r = 0; r < 5; ++r)
            P5(s->flux_res, k, r) = (Gxp[r] + Gxn[r] + Gyp[r] + Gyn[r] + Gzp[r] + Gzn[r]);
    }
}

/* ------------------------------------------------------------------------ */
/* implicit LU-SGS                                                          */
/* ------------------------------------------------------------------------ */

/* one direction's D_inv contribution: implicit_D_inv_interior.f90:51-108
 * (positive half) and :113-169 (negative half); c = prim(2..4) + or - a */
static double dinv_dir(const orc_state *s, int i, const stencil_t *st, double as)
{
    geo6 g = {0, 0, 0, 0, 0, 0};
    double sx = 0.0, sy = 0.0, sz = 0.0;
    double x_i = s->x[i], y_i = s->y[i], z_i = s->z[i];
    double t1[3] = {F3(s->tan1, i, 0), F3(s->tan1, i, 1), F3(s->tan1, i, 2)};
    double t2[3] = {F3(s->tan2, i, 0), F3(s->tan2, i, 1), F3(s->tan2, i, 2)};
    double n[3] = {F3(s->nor, i, 0), F3(s->nor, i, 1), F3(s->nor, i, 2)};
    for (int j = 0; j < st->count; ++j) {
        int k = st->list[(size_t)i * st->stride + j] - 1;
        double delx = s->x[k] - x_i;
        double dely = s->y[k] - y_i;
        double delz = s->z[k] - z_i;
        double dels = delx * t1[0] + dely * t1[1] + delz * t1[2];
        double delt = delx * t2[0] + dely * t2[1] + delz * t2[2];
        double deln = delx * n[0] + dely * n[1] + delz * n[2];
        double dist = sqrt(dels * dels + delt * delt + deln * deln);
        double weights = 1.0 / ref_pow(dist, s->power);
        double dels_w = dels * weights, delt_w = delt * weights, deln_w = deln * weights;
        g.xx = g.xx + dels * dels_w;
        g.yy = g.yy + delt * delt_w;
        g.zz = g.zz + deln * deln_w;
        g.xy = g.xy + dels * delt_w;
        g.yz = g.yz + delt * deln_w;
        g.zx = g.zx + deln * dels_w;
        sx = sx + dels_w;
        sy = sy + delt_w;
        sz = sz + deln_w;
    }
    double det = cramer_det(&g);
    double temp = cramer_num(&g, st->axis, sx, sy, sz);
    /* global-frame velocity component with the sound speed (Q8) */
    double temp1;
    if (st->sign > 0) {
        temp1 = P5(s->prim, i, 1 + st->axis) + as;
        return 0.50 * fabs(temp1 + fabs(temp1)) * temp / det;
    }
    temp1 = P5(s->prim, i, 1 + st->axis) - as;
    return -(0.50 * fabs(temp1 - fabs(temp1)) * temp / det);
}

double orc_implicit_D_inv(const orc_state *s, int i)
{
    double delta_t = s->delt[i];
    double as = sqrt(1.40 * P5(s->prim, i, 4) / P5(s->prim, i, 0));
    stencil_t st;
    if (s->status[i] == 1) {
        /* implicit_D_inv_wall.f90:45-357 */
        double d[5];
        st = split_stencil(s, i, 0, +1, FK_WALL); d[0] = dinv_dir(s, i, &st, as);
        st = split_stencil(s, i, 0, -1, FK_WALL); d[1] = dinv_dir(s, i, &st, as);
        st = split_stencil(s, i, 1, +1, FK_WALL); d[2] = dinv_dir(s, i, &st, as);
        st = split_stencil(s, i, 1, -1, FK_WALL); d[3] = dinv_dir(s, i, &st, as);
        st = full_conn_stencil(s, i, 2, -1, FK_HALF); d[4] = dinv_dir(s, i, &st, as);
        double temp = (1.0 / delta_t) - 2.0 * (d[0] + d[1] + d[2] + d[3] + d[4]);
        return 1.0 / temp;
    }
    /* implicit_D_inv_interior.f90:46-417 */
    double d[6];
    st = split_stencil(s, i, 0, +1, FK_HALF); d[0] = dinv_dir(s, i, &st, as);
    st = split_stencil(s, i, 0, -1, FK_HALF); d[1] = dinv_dir(s, i, &st, as);
    st = split_stencil(s, i, 1, +1, FK_HALF); d[2] = dinv_dir(s, i, &st, as);
    st = split_stencil(s, i, 1, -1, FK_HALF); d[3] = dinv_dir(s, i, &st, as);
    st = split_stencil(s, i, 2, +1, FK_HALF); d[4] = dinv_dir(s, i, &st, as);
    st = split_stencil(s, i, 2, -1, FK_HALF); d[5] = dinv_dir(s, i, &st, as);
    double temp = (1.0 / delta_t) - (d[0] + d[1] + d[2] + d[3] + d[4] + d[5]);
    return 1.0 / temp;
}

/* implicit_D_inv_outer.f90:17-363, reached only from the alternative sweep
 * driver (implicit_forward_and_reverse_sweeps_backup.f90).  x/y over the
 * split lists, z over the full conn; the negative halves enter with a plus
 * sign (:171, :295) and the sum is (1/delt) - dGxp + dGxn - dGyp + dGyn - dGzp
 * (:360).  The routine's local `delt` is both the time step (:46) and the
 * loop's tan2 projection (:73, ...), so the 1/delt of :360 is the projection
 * of the last conn neighbour's offset (the time step when nbhs = 0). */
static double implicit_D_inv_outer(const orc_state *s, int i)
{
    double delt = s->delt[i];
    double as = sqrt(1.40 * P5(s->prim, i, 4) / P5(s->prim, i, 0));
    stencil_t st;
    st = split_stencil(s, i, 0, +1, FK_OUTER); double dGxp = dinv_dir(s, i, &st, as);
    st = split_stencil(s, i, 0, -1, FK_OUTER); double dGxn = -dinv_dir(s, i, &st, as);
    st = split_stencil(s, i, 1, +1, FK_OUTER); double dGyp = dinv_dir(s, i, &st, as);
    st = split_stencil(s, i, 1, -1, FK_OUTER); double dGyn = -dinv_dir(s, i, &st, as);
    st = full_conn_stencil(s, i, 2, +1, FK_HALF); double dGzp = dinv_dir(s, i, &st, as);
    if (s->nbhs[i] > 0) {
        int k = CONN(s, i, s->nbhs[i] - 1) - 1;
        double delx = s->x[k] - s->x[i];
        double dely = s->y[k] - s->y[i];
        double delz = s->z[k] - s->z[i];
        delt = delx * F3(s->tan2, i, 0) + dely * F3(s->tan2, i, 1) + delz * F3(s->tan2, i, 2);
    }
    double temp = (1.0 / delt) - dGxp + dGxn - dGyp + dGyn - dGzp;
    return 1.0 / temp;
}

/* get_primitive_vector.f90:10-41: local-frame U of point k -> global prim */
static void get_primitive_vector(const orc_state *s, int k, double U[5], double prim[5])
{
    prim[0] = U[0];
    double temp = 1.0 / U[0];
    double U2r = U[1], U3r = U[2], U4r = U[3];
    U[1] = F3(s->tan1, k, 0) * U2r + F3(s->tan2, k, 0) * U3r + F3(s->nor, k, 0) * U4r;
    U[2] = F3(s->tan1, k, 1) * U2r + F3(s->tan2, k, 1) * U3r + F3(s->nor, k, 1) * U4r;
    U[3] = F3(s->tan1, k, 2) * U2r + F3(s->tan2, k, 2) * U3r + F3(s->nor, k, 2) * U4r;
    prim[1] = temp * U[1];
    prim[2] = temp * U[2];
    prim[3] = temp * U[3];
    temp = prim[1] * prim[1] + prim[2] * prim[2] + prim[3] * prim[3];
    prim[4] = 0.4 * (U[4] - 0.5 * prim[0] * temp);
}

/* one direction of LdelU / UdelU:
 * implicit_LdelU_forward_sweep_interior.f90:51-117 (forward: alias_k <
 * alias_i, delUp) and implicit_UdelU_backward_sweep_interior.f90:51-121
 * (backward: alias_k > alias_i, delUn).  The geometric sums use every list
 * entry; the flux differences only the qualifying neighbours. */
static void sweep_dir(const orc_state *s, int i, const stencil_t *st, int forward, double dG[5])
{
    geo6 g = {0, 0, 0, 0, 0, 0};
    double sx[5] = {0}, sy[5] = {0}, sz[5] = {0};
    double x_i = s->x[i], y_i = s->y[i], z_i = s->z[i];
    double t1[3] = {F3(s->tan1, i, 0), F3(s->tan1, i, 1), F3(s->tan1, i, 2)};
    double t2[3] = {F3(s->tan2, i, 0), F3(s->tan2, i, 1), F3(s->tan2, i, 2)};
    double n[3] = {F3(s->nor, i, 0), F3(s->nor, i, 1), F3(s->nor, i, 2)};
    const double *dU = forward ? s->delUp : s->delUn;
    for (int j = 0; j < st->count; ++j) {
        int k = st->list[(size_t)i * st->stride + j] - 1;
        double delx = s->x[k] - x_i;
        double dely = s->y[k] - y_i;
        double delz = s->z[k] - z_i;
        double dels = delx * t1[0] + dely * t1[1] + delz * t1[2];
        double delt = delx * t2[0] + dely * t2[1] + delz * t2[2];
        double deln = delx * n[0] + dely * n[1] + delz * n[2];
        double dist = sqrt(dels * dels + delt * delt + deln * deln);
        double weights = 1.0 / ref_pow(dist, s->power);
        double dels_w = dels * weights, delt_w = delt * weights, deln_w = deln * weights;
        g.xx = g.xx + dels * dels_w;
        g.yy = g.yy + delt * delt_w;
        g.zz = g.zz + deln * deln_w;
        g.xy = g.xy + dels * delt_w;
        g.yz = g.yz + delt * deln_w;
        g.zx = g.zx + deln * dels_w;
        int alias_of_i = s->alias[i];
        int alias_of_k = s->alias[k];
        int take = forward ? (alias_of_k < alias_of_i) : (alias_of_k > alias_of_i);
        if (take) {
            double prim[5], Gold[5], Gnew[5], U[5];
            for (int r = 0; r < 5; ++r) prim[r] = P5(s->prim, k, r);
            eval_flux(st, t1, t2, n, prim, s->pi, Gold);
            for (int r = 0; r < 5; ++r) U[r] = P5(dU, k, r) + P5(s->U, k, r);
            get_primitive_vector(s, k, U, prim);
            eval_flux(st, t1, t2, n, prim, s->pi, Gnew);
            for (int r = 0; r < 5; ++r) {
                double t = Gnew[r] - Gold[r];
                sx[r] = sx[r] + t * dels_w;
                sy[r] = sy[r] + t * delt_w;
                sz[r] = sz[r] + t * deln_w;
            }
        }
    }
    double det = cramer_det(&g);
    for (int r = 0; r < 5; ++r) dG[r] = cramer_num(&g, st->axis, sx[r], sy[r], sz[r]) / det;
}

static void sweep_term(const orc_state *s, int i, int forward, double out[5])
{
    stencil_t st;
    if (s->status[i] == 1) {
        /* implicit_LdelU_forward_sweep_wall.f90:50-405 */
        double d[5][5];
        st = split_stencil(s, i, 0, +1, FK_WALL); sweep_dir(s, i, &st, forward, d[0]);
        st = split_stencil(s, i, 0, -1, FK_WALL); sweep_dir(s, i, &st, forward, d[1]);
        st = split_stencil(s, i, 1, +1, FK_WALL); sweep_dir(s, i, &st, forward, d[2]);
        st = split_stencil(s, i, 1, -1, FK_WALL); sweep_dir(s, i, &st, forward, d[3]);
        st = full_conn_stencil(s, i, 2, -1, FK_HALF); sweep_dir(s, i, &st, forward, d[4]);
        for (int r = 0; r < 5; ++r)
            out[r] = 2.0 * (d[0][r] + d[1][r] + d[2][r] + d[3][r] + d[4][r]);
        return;
    }
    /* implicit_LdelU_forward_sweep_interior.f90:49-474 */
    double d[6][5];
    st = split_stencil(s, i, 0, +1, FK_HALF); sweep_dir(s, i, &st, forward, d[0]);
    st = split_stencil(s, i, 0, -1, FK_HALF); sweep_dir(s, i, &st, forward, d[1]);
    st = split_stencil(s, i, 1, +1, FK_HALF); sweep_dir(s, i, &st, forward, d[2]);
    st = split_stencil(s, i, 1, -1, FK_HALF); sweep_dir(s, i, &st, forward, d[3]);
    st = split_stencil(s, i, 2, +1, FK_HALF); sweep_dir(s, i, &st, forward, d[4]);
    st = split_stencil(s, i, 2, -1, FK_HALF); sweep_dir(s, i, &st, forward, d[5]);
    for (int r = 0; r < 5; ++r)
        out[r] = d[0][r] + d[1][r] + d[2][r] + d[3][r] + d[4][r] + d[5][r];
}

/* implicit_LdelU_forward_sweep_outer.f90:20-406 and
 * implicit_UdelU_backward_sweep_outer.f90:19-406 (alternative driver only):
 * x/y octant fluxes Go over the split lists, z flux_Gzp over the full conn,
 * each dG = temp/det, summed without a factor */
static void sweep_term_outer(const orc_state *s, int i, int forward, double out[5])
{
    double d[5][5];
    stencil_t st;
    st = split_stencil(s, i, 0, +1, FK_OUTER); sweep_dir(s, i, &st, forward, d[0]);
    st = split_stencil(s, i, 0, -1, FK_OUTER); sweep_dir(s, i, &st, forward, d[1]);
    st = split_stencil(s, i, 1, +1, FK_OUTER); sweep_dir(s, i, &st, forward, d[2]);
    st = split_stencil(s, i, 1, -1, FK_OUTER); sweep_dir(s, i, &st, forward, d[3]);
    st = full_conn_stencil(s, i, 2, +1, FK_HALF); sweep_dir(s, i, &st, forward, d[4]);
    for (int r = 0; r < 5; ++r) out[r] = d[0][r] + d[1][r] + d[2][r] + d[3][r] + d[4][r];
}

/* nearest neighbour of a supersonic point (sweeps.f90:39-50); skip_wall
 * adds the status /= 1 filter of the alternative driver and
 * state_update_implicit.f90:72,89 */
static int nearest_neighbour_f(const orc_state *s, int p, double start, int skip_wall)
{
    double min_dist = start;
    int q = p;
    for (int r = 0; r < s->nbhs[p]; ++r) {
        int nbh = CONN(s, p, r) - 1;
        double dx = s->x[nbh] - s->x[p];
        double dy = s->y[nbh] - s->y[p];
        double dz = s->z[nbh] - s->z[p];
        double ds = sqrt(dx * dx + dy * dy + dz * dz);
        if (ds < min_dist && (!skip_wall || s->status[nbh] != 1)) {
            min_dist = ds;
            q = nbh;
        }
    }
    return q;
}

static int nearest_neighbour(const orc_state *s, int p, double start)
{
    return nearest_neighbour_f(s, p, start, 0);
}

void orc_implicit_sweeps(orc_state *s)
{
    int n = s->n;
    /* forward sweep, implicit_forward_and_reverse_sweeps.f90:26-53.  Entries
     * of point_with_alias equal to 0 are not visited (the reference visits
     * point 0, outside every array; those writes never reach a real point). */
    for (int ii = 0; ii < n; ++ii) {
        int p = s->point_with_alias[ii] - 1;
        if (p < 0) continue;
        int st = s->status[p];
        if (st == 1 || st == 0) {
            double D_inv = orc_implicit_D_inv(s, p);
            double LdelU[5];
            sweep_term(s, p, 1, LdelU);
            for (int r = 0; r < 5; ++r)
                P5(s->delUp, p, r) = -(D_inv * (P5(s->flux_res, p, r) + LdelU[r]));
        } else if (st == 5 || st == 6) {
            int q = nearest_neighbour(s, p, 1000000.0);
            for (int r = 0; r < 5; ++r) P5(s->delUp, p, r) = P5(s->delUp, q, r);
        }
    }
    /* backward sweep, :58-84 (D_inv recomputed identically) */
    for (int ii = n - 1; ii >= 0; --ii) {
        int p = s->point_with_alias[ii] - 1;
        if (p < 0) continue;
        int st = s->status[p];
        if (st == 1 || st == 0) {
            double D_inv = orc_implicit_D_inv(s, p);
            double UdelU[5];
            sweep_term(s, p, 0, UdelU);
            for (int r = 0; r < 5; ++r)
                P5(s->delUn, p, r) = P5(s->delUp, p, r) - D_inv * UdelU[r];
        } else if (st == 5 || st == 6) {
            int q = nearest_neighbour(s, p, 1000000.0);
            for (int r = 0; r < 5; ++r) P5(s->delUn, p, r) = P5(s->delUn, q, r);
        }
    }
    /* :86-88 */
    for (int i = 0; i < n; ++i)
        for (int r = 0; r < 5; ++r) P5(s->delUp, i, r) = P5(s->delUn, i, r);
}

/* implicit_forward_and_reverse_sweeps_backup.f90:16-141: the forward sweep
 * visits wall, interior, then outer points, each class in index order, then
 * the supersonic inlet and outlet points (nearest non-wall neighbour's
 * delUp); the backward sweep visits the same classes in reverse order.  The
 * alias comparisons inside LdelU/UdelU are unchanged, so a neighbour with a
 * smaller alias that this order visits later contributes its delUp from
 * before the sweep.  No delUp = delUn copy: state_update_implicit.f90:98-100
 * makes it. */
void orc_implicit_sweeps_backup(orc_state *s)
{
    double D_inv, T[5];
    for (int m = 0; m < s->wall_points; ++m) {
        int k = s->wall_points_index[m] - 1;
        D_inv = orc_implicit_D_inv(s, k);
        sweep_term(s, k, 1, T);
        for (int r = 0; r < 5; ++r) P5(s->delUp, k, r) = -(D_inv * (P5(s->flux_res, k, r) + T[r]));
    }
    for (int m = 0; m < s->interior_points; ++m) {
        int k = s->interior_points_index[m] - 1;
        D_inv = orc_implicit_D_inv(s, k);
        sweep_term(s, k, 1, T);
        for (int r = 0; r < 5; ++r) P5(s->delUp, k, r) = -(D_inv * (P5(s->flux_res, k, r) + T[r]));
    }
    for (int m = 0; m < s->outer_points; ++m) {
        int k = s->outer_points_index[m] - 1;
        D_inv = implicit_D_inv_outer(s, k);
        sweep_term_outer(s, k, 1, T);
        for (int r = 0; r < 5; ++r) P5(s->delUp, k, r) = -(D_inv * (P5(s->flux_res, k, r) + T[r]));
    }
    for (int m = 0; m < s->supersonic_inlet_points; ++m) {
        int k = s->supersonic_inlet_points_index[m] - 1;
        int p = nearest_neighbour_f(s, k, 100000.0, 1);
        for (int r = 0; r < 5; ++r) P5(s->delUp, k, r) = P5(s->delUp, p, r);
    }
    for (int m = 0; m < s->supersonic_outlet_points; ++m) {
        int k = s->supersonic_outlet_points_index[m] - 1;
        int p = nearest_neighbour_f(s, k, 100000.0, 1);
        for (int r = 0; r < 5; ++r) P5(s->delUp, k, r) = P5(s->delUp, p, r);
    }
    for (int m = s->wall_points - 1; m >= 0; --m) {
        int k = s->wall_points_index[m] - 1;
        D_inv = orc_implicit_D_inv(s, k);
        sweep_term(s, k, 0, T);
        for (int r = 0; r < 5; ++r) P5(s->delUn, k, r) = P5(s->delUp, k, r) - D_inv * T[r];
    }
    for (int m = s->interior_points - 1; m >= 0; --m) {
        int k = s->interior_points_index[m] - 1;
        D_inv = orc_implicit_D_inv(s, k);
        sweep_term(s, k, 0, T);
        for (int r = 0; r < 5; ++r) P5(s->delUn, k, r) = P5(s->delUp, k, r) - D_inv * T[r];
    }
    for (int m = s->outer_points - 1; m >= 0; --m) {
        int k = s->outer_points_index[m] - 1;
        D_inv = implicit_D_inv_outer(s, k);
        sweep_term_outer(s, k, 0, T);
        for (int r = 0; r < 5; ++r) P5(s->delUn, k, r) = P5(s->delUp, k, r) - D_inv * T[r];
    }
    for (int m = s->supersonic_inlet_points - 1; m >= 0; --m) {
        int k = s->supersonic_inlet_points_index[m] - 1;
        int p = nearest_neighbour_f(s, k, 100000.0, 1);
        for (int r = 0; r < 5; ++r) P5(s->delUn, k, r) = P5(s->delUn, p, r);
    }
    for (int m = s->supersonic_outlet_points - 1; m >= 0; --m) {
        int k = s->supersonic_outlet_points_index[m] - 1;
        int p = nearest_neighbour_f(s, k, 100000.0, 1);
        for (int r = 0; r < 5; ++r) P5(s->delUn, k, r) = P5(s->delUn, p, r);
    }
}

double orc_implicit_D_inv_outer(const orc_state *s, int i)
{
    return implicit_D_inv_outer(s, i);
}

void orc_explicit_sweep(orc_state *s)
{
    /* Explicit form of the update that state_update.f90:35,67 keeps as
     * comments (U = U - delt*flux_res), expressed through the sweep's output:
     * delUn = -delt*flux_res, then delUp = delUn as the sweep ends. */
    for (int k = 0; k < s->n; ++k)
        for (int r = 0; r < 5; ++r) {
            P5(s->delUn, k, r) = -(s->delt[k] * P5(s->flux_res, k, r));
            P5(s->delUp, k, r) = P5(s->delUn, k, r);
        }
}

/* ------------------------------------------------------------------------ */
/* state update                                                             */
/* ------------------------------------------------------------------------ */

/* conserved_to_primitive_vector.f90:10-42 */
static void conserved_to_primitive(orc_state *s, int k, double U[5])
{
    double *p = &P5(s->prim, k, 0);
    p[0] = U[0];
    double temp = 1.0 / U[0];
    double U2r = U[1], U3r = U[2], U4r = U[3];
    U[1] = F3(s->tan1, k, 0) * U2r + F3(s->tan2, k, 0) * U3r + F3(s->nor, k, 0) * U4r;
    U[2] = F3(s->tan1, k, 1) * U2r + F3(s->tan2, k, 1) * U3r + F3(s->nor, k, 1) * U4r;
    U[3] = F3(s->tan1, k, 2) * U2r + F3(s->tan2, k, 2) * U3r + F3(s->nor, k, 2) * U4r;
    p[1] = temp * U[1];
    p[2] = temp * U[2];
    p[3] = temp * U[3];
    temp = p[1] * p[1] + p[2] * p[2] + p[3] * p[3];
    p[4] = 0.4 * (U[4] - 0.5 * p[0] * temp);
}

void orc_conserved_vector_Ubar(const orc_state *s, int k, double Ubar[5])
{
    /* conserved_vector_Ubar.f90:19-58 */
    double pi = s->pi;
    double t1x = F3(s->tan1, k, 0), t1y = F3(s->tan1, k, 1), t1z = F3(s->tan1, k, 2);
    double t2x = F3(s->tan2, k, 0), t2y = F3(s->tan2, k, 1), t2z = F3(s->tan2, k, 2);
    double nx = F3(s->nor, k, 0), ny = F3(s->nor, k, 1), nz = F3(s->nor, k, 2);
    double u1_inf_rot = s->u1_inf * t1x + s->u2_inf * t1y + s->u3_inf * t1z;
    double u2_inf_rot = s->u1_inf * t2x + s->u2_inf * t2y + s->u3_inf * t2z;
    double u3_inf_rot = s->u1_inf * nx + s->u2_inf * ny + s->u3_inf * nz;
    double rho_e_inf = s->pr_inf * 2.5 + 0.5 * s->rho_inf *
        (u1_inf_rot * u1_inf_rot + u2_inf_rot * u2_inf_rot + u3_inf_rot * u3_inf_rot);
    double beta = (0.5 * s->rho_inf) / s->pr_inf;
    double S3 = u3_inf_rot * sqrt(beta);
    double B3_inf = exp(-S3 * S3) / (2.0 * sqrt(pi * beta));
    double A3n_inf = 0.5 * (1.0 - erf(S3));
    double rho = P5(s->prim, k, 0), u1 = P5(s->prim, k, 1), u2 = P5(s->prim, k, 2);
    double u3 = P5(s->prim, k, 3), pr = P5(s->prim, k, 4);
    double u1_rot = u1 * t1x + u2 * t1y + u3 * t1z;
    double u2_rot = u1 * t2x + u2 * t2y + u3 * t2z;
    double u3_rot = u1 * nx + u2 * ny + u3 * nz;
    double rho_e = pr * 2.5 + 0.5 * rho * (u1_rot * u1_rot + u2_rot * u2_rot + u3_rot * u3_rot);
    beta = (0.5 * rho) / pr;
    S3 = u3_rot * sqrt(beta);
    double B3 = exp(-S3 * S3) / (2.0 * sqrt(pi * beta));
    double A3p = 0.5 * (1.0 + erf(S3));
    Ubar[0] = (s->rho_inf * A3n_inf) + (rho * A3p);
    Ubar[1] = (s->rho_inf * u1_inf_rot * A3n_inf) + (rho * u1_rot * A3p);
    Ubar[2] = (s->rho_inf * u2_inf_rot * A3n_inf) + (rho * u2_rot * A3p);
    double temp1 = s->rho_inf * (u3_inf_rot * A3n_inf - B3_inf);
    double temp2 = rho * (u3_rot * A3p + B3);
    Ubar[3] = temp1 + temp2;
    temp1 = (rho_e_inf * A3n_inf - 0.5 * s->rho_inf * u3_inf_rot * B3_inf);
    temp2 = (rho_e * A3p + 0.5 * rho * u3_rot * B3);
    Ubar[4] = temp1 + temp2;
}

void orc_state_update(orc_state *s)
{
    /* state_update.f90:22-115 (canonical explicit-outer variant) */
    double sum_res_sqr = 0.0;
    s->max_res = 0.0;
    for (int m = 0; m < s->wall_points; ++m) {
        int k = s->wall_points_index[m] - 1;
        double temp = P5(s->U, k, 0);
        double U[5];
        for (int r = 0; r < 5; ++r) U[r] = P5(s->U, k, r) + P5(s->delUn, k, r);
        U[3] = 0.0;
        double res_sqr = (U[0] - temp) * (U[0] - temp);
        if (res_sqr > s->max_res) {
            s->max_res = res_sqr;
            s->max_res_point = k + 1;
        }
        sum_res_sqr = sum_res_sqr + res_sqr;
        conserved_to_primitive(s, k, U);
        for (int r = 0; r < 5; ++r) P5(s->delUp, k, r) = P5(s->delUn, k, r);
    }
    for (int m = 0; m < s->outer_points; ++m) {
        int k = s->outer_points_index[m] - 1;
        double U[5];
        orc_conserved_vector_Ubar(s, k, U);
        for (int r = 0; r < 5; ++r) U[r] = U[r] - s->delt[k] * P5(s->flux_res, k, r);
        conserved_to_primitive(s, k, U);
    }
    for (int m = 0; m < s->interior_points; ++m) {
        int k = s->interior_points_index[m] - 1;
        double temp = P5(s->U, k, 0);
        double U[5];
        for (int r = 0; r < 5; ++r) U[r] = P5(s->U, k, r) + P5(s->delUn, k, r);
        double res_sqr = (U[0] - temp) * (U[0] - temp);
        if (res_sqr > s->max_res) {
            s->max_res = res_sqr;
            s->max_res_point = k + 1;
        }
        sum_res_sqr = sum_res_sqr + res_sqr;
        conserved_to_primitive(s, k, U);
        for (int r = 0; r < 5; ++r) P5(s->delUp, k, r) = P5(s->delUn, k, r);
    }
    /* :80-114 supersonic outlet, then inlet: copy the nearest neighbour */
    for (int m = 0; m < s->supersonic_outlet_points; ++m) {
        int k = s->supersonic_outlet_points_index[m] - 1;
        int p = nearest_neighbour(s, k, 100000.0);
        for (int r = 0; r < 5; ++r) P5(s->prim, k, r) = P5(s->prim, p, r);
    }
    for (int m = 0; m < s->supersonic_inlet_points; ++m) {
        int k = s->supersonic_inlet_points_index[m] - 1;
        int p = nearest_neighbour(s, k, 100000.0);
        for (int r = 0; r < 5; ++r) P5(s->prim, k, r) = P5(s->prim, p, r);
    }
    s->res_new = sqrt(sum_res_sqr) / s->n;
}

/* state_update_implicit.f90:14-102, the update paired with the alternative
 * sweep: wall and interior U + delUn without the wall's U(4) = 0 and without
 * a per-point delUp copy, outer Ubar + delUn (:45-47), supersonic points
 * copy the nearest non-wall neighbour's prim (:62-96), then delUp = delUn
 * for every point (:98-100). */
void orc_state_update_implicit(orc_state *s)
{
    double sum_res_sqr = 0.0;
    s->max_res = 0.0;
    for (int m = 0; m < s->wall_points; ++m) {
        int k = s->wall_points_index[m] - 1;
        double temp = P5(s->U, k, 0);
        double U[5];
        for (int r = 0; r < 5; ++r) U[r] = P5(s->U, k, r) + P5(s->delUn, k, r);
        double res_sqr = (U[0] - temp) * (U[0] - temp);
        if (res_sqr > s->max_res) {
            s->max_res = res_sqr;
            s->max_res_point = k + 1;
        }
        sum_res_sqr = sum_res_sqr + res_sqr;
        conserved_to_primitive(s, k, U);
    }
    for (int m = 0; m < s->outer_points; ++m) {
        int k = s->outer_points_index[m] - 1;
        double U[5];
        orc_conserved_vector_Ubar(s, k, U);
        for (int r = 0; r < 5; ++r) U[r] = U[r] + P5(s->delUn, k, r);
        conserved_to_primitive(s, k, U);
    }
    for (int m = 0; m < s->interior_points; ++m) {
        int k = s->interior_points_index[m] - 1;
        double temp = P5(s->U, k, 0);
        double U[5];
        for (int r = 0; r < 5; ++r) U[r] = P5(s->U, k, r) + P5(s->delUn, k, r);
        double res_sqr = (U[0] - temp) * (U[0] - temp);
        if (res_sqr > s->max_res) {
            s->max_res = res_sqr;
            s->max_res_point = k + 1;
        }
        sum_res_sqr = sum_res_sqr + res_sqr;
        conserved_to_primitive(s, k, U);
    }
    for (int m = 0; m < s->supersonic_outlet_points; ++m) {
        int k = s->supersonic_outlet_points_index[m] - 1;
        int p = nearest_neighbour_f(s, k, 100000.0, 1);
        for (int r = 0; r < 5; ++r) P5(s->prim, k, r) = P5(s->prim, p, r);
    }
    for (int m = 0; m < s->supersonic_inlet_points; ++m) {
        int k = s->supersonic_inlet_points_index[m] - 1;
        int p = nearest_neighbour_f(s, k, 100000.0, 1);
        for (int r = 0; r < 5; ++r) P5(s->prim, k, r) = P5(s->prim, p, r);
    }
    s->res_new = sqrt(sum_res_sqr) / s->n;
    for (int i = 0; i < s->n; ++i)
        for (int r = 0; r < 5; ++r) P5(s->delUp, i, r) = P5(s->delUn, i, r);
}

/* ------------------------------------------------------------------------ */
/* the iteration                                                            */
/* ------------------------------------------------------------------------ */

void orc_fpi_solver(orc_state *s, int t, int implicit_mode)
{
    /* fpi_solver.f90:27-43 */
    orc_compute_conserved_vector(s);
    orc_eval_q_variables(s);
    orc_eval_q_derivatives(s);
    orc_timestep_delt(s);
    orc_eval_flux_residual(s);
    if (implicit_mode == ORC_IMPLICIT_BACKUP) {
        orc_implicit_sweeps_backup(s);
        orc_state_update_implicit(s);
    } else {
        if (implicit_mode) orc_implicit_sweeps(s);
        else orc_explicit_sweep(s);
        orc_state_update(s);
    }
    if (t <= 2) {
        s->res_old = s->res_new;
        s->residue = 0.0;
    } else {
        s->residue = log10(s->res_new / s->res_old);
    }
}

void orc_run(orc_state *s, int t0, int n_iters, int implicit_mode,
             double *res_new_hist, double *residue_hist)
{
    for (int it = 0; it < n_iters; ++it) {
        orc_fpi_solver(s, t0 + it, implicit_mode);
        if (res_new_hist) res_new_hist[it] = s->res_new;
        if (residue_hist) residue_hist[it] = s->residue;
    }
}
