// kernels_fp64.hip -- FP64 kernels of the q-LSKUM pseudo-time iteration in
// the reference's evaluation order.
//
// Streaming kernels take one thread per point; the gathering kernels a group
// of lanes per point (q-derivatives: 4 lanes splitting conn; list-by-list
// flux: one lane per direction list).  Every list-walking routine below
// follows the reference routine named at it with the same association of
// each product; the file is compiled with -ffp-contract=off.  The fused flux
// kernels (kernels_flux_fused.hip) give the flux residual of every point
// whose walk fits in 64 entries; this file keeps the list-by-list form for
// the rest and for the reference-order mode.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "device_state.h"
#include "kernels.h"
#include "qlskum_math.h"

namespace mfcfd {

namespace {

// ---------------------------------------------------------------------------
// boundary permutations
// ---------------------------------------------------------------------------
__global__ void k_gather(double* dst, const double* src, const int* perm, int n, int m) {
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= (long long)n * m) return;
    const long long k = e / m;
    const int c = (int)(e - k * m);
    dst[e] = src[(long long)perm[k] * m + c];
}

__global__ void k_scatter(double* dst, const double* src, const int* perm, int n, int m) {
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= (long long)n * m) return;
    const long long k = e / m;
    const int c = (int)(e - k * m);
    dst[(long long)perm[k] * m + c] = src[e];
}

// ---------------------------------------------------------------------------
// Cramer weights of every direction list: the geometric sums of the list in
// list order (interior_flux_dGxpos.f90:84-90) and the cofactors of the column
// the direction solves for, divided by the determinant
// ---------------------------------------------------------------------------
template <bool UW>
__global__ void __launch_bounds__(kBlock) k_stencil_coef(DevGeom g, double* coef, double power) {
    const int i = xcd_index();
    if (i >= g.n) return;
    const Frame<double> f = load_frame(g.frame, i);
    for (int d = 0; d < 6; ++d) {
        const int cnt = g.dir_cnt[(long long)d * g.n + i];
        const int* list = g.lst + g.dir_off[(long long)d * g.n + i];
        list_cofactors<UW>(g, i, f, list, cnt, d >> 1, power, coef + 18 * (long long)i + 3 * d);
    }
}

// Bind-time check of every least-squares system the iteration solves: the
// direction lists the flux stencils walk and each point's conn, whose 3x3
// weighted normal matrices the reference inverts by Cramer's rule
// (interior_flux_dGxpos.f90:84-95, q_derivatives.f90:44-61).  A coplanar or
// collinear stencil makes that matrix singular and the reference divides by
// a zero determinant.  Flags the smallest (point, list) whose determinant is
// not above 1e-12 of its scale ((trace/3)^3, invariant under the frame's
// rotation), NaN included: 8*point + list, list 6 = conn.  Empty walked lists
// are refused on the host.
template <bool UW>
__global__ void __launch_bounds__(kBlock) k_stencil_check(DevGeom g, double power, int* first_bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    const Frame<double> f = load_frame(g.frame, i);
    const double4 pi4 = g.pos[i];
    for (int d = 0; d <= 6; ++d) {
        const int cnt = d < 6 ? g.dir_cnt[(long long)d * g.n + i] : g.nbhs[i];
        const int* list = d < 6 ? g.lst + g.dir_off[(long long)d * g.n + i] : g.conn + kConnCols * (long long)i;
        if (cnt == 0) continue;
        Geo6<double> geo = {0, 0, 0, 0, 0, 0};
        for (int j = 0; j < cnt; ++j) {
            const double4 pk = g.pos[list[j]];
            const double delx = pk.x - pi4.x, dely = pk.y - pi4.y, delz = pk.z - pi4.z;
            const double dels = delx * f.t1[0] + dely * f.t1[1] + delz * f.t1[2];
            const double delt = delx * f.t2[0] + dely * f.t2[1] + delz * f.t2[2];
            const double deln = delx * f.n[0] + dely * f.n[1] + delz * f.n[2];
            const double w = wls_weight<UW>(dels, delt, deln, power);
            geo.xx += dels * dels * w;
            geo.yy += delt * delt * w;
            geo.zz += deln * deln * w;
            geo.xy += dels * delt * w;
            geo.yz += delt * deln * w;
            geo.zx += deln * dels * w;
        }
        const double tr = (geo.xx + geo.yy + geo.zz) / 3.0;
        const double det = cramer_det(geo);
        if (!(fabs(det) > 1e-12 * tr * tr * tr)) atomicMin(first_bad, 8 * i + d);
    }
}

// ---------------------------------------------------------------------------
// compute_conserved_vector.f90:20-32 and q_variables.f90:18-38
// ---------------------------------------------------------------------------
__device__ __forceinline__ void conserved_point(const DevState& s, long long k, const double p[5],
                                                const Frame<double>& f) {
    double U[5];
    U[0] = p[0];
    double temp1 = p[0] * p[1];
    const double temp2 = p[0] * p[2];
    const double temp3 = p[0] * p[3];
    U[1] = temp1 * f.t1[0] + temp2 * f.t1[1] + temp3 * f.t1[2];
    U[2] = temp1 * f.t2[0] + temp2 * f.t2[1] + temp3 * f.t2[2];
    U[3] = temp1 * f.n[0] + temp2 * f.n[1] + temp3 * f.n[2];
    temp1 = p[1] * p[1] + p[2] * p[2] + p[3] * p[3];
    U[4] = 2.5 * p[4] + 0.5 * U[0] * temp1;
#pragma unroll
    for (int r = 0; r < 5; ++r) s.U[5 * k + r] = U[r];
}

template <typename QT>
__device__ __forceinline__ void q_point(const DevState& s, long long k, const double p[5]) {
    const double rho = p[0], u1 = p[1], u2 = p[2], u3 = p[3], pr = p[4];
    const double beta = 0.5 * rho / pr;
    double q[5];
    q[0] = log(rho) + (log(beta) * 2.5) - beta * (u1 * u1 + u2 * u2 + u3 * u3);
    const double two_times_beta = 2.0 * beta;
    q[1] = two_times_beta * u1;
    q[2] = two_times_beta * u2;
    q[3] = two_times_beta * u3;
    q[4] = -two_times_beta;
    QT* qo = qarrays<QT>(s).qp(k);
#pragma unroll
    for (int r = 0; r < 5; ++r) qo[r] = QT(q[r]);
    // timestep_delt.f90:40-42: the neighbour-only part of the local step
    const double den = __builtin_sqrt(u1 * u1 + u2 * u2 + u3 * u3) + 3.0 * __builtin_sqrt(pr / rho);
    s.tsden[k] = den;
    if (sizeof(QT) == 4) qo[5] = QT(den);
}

template <bool DO_U, bool DO_Q, typename QT>
__global__ void __launch_bounds__(kBlock) k_conserved_q(DevGeom g, DevState s) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g.n) return;
    double p[5];
#pragma unroll
    for (int r = 0; r < 5; ++r) p[r] = s.prim[5 * (long long)k + r];
    if (DO_U) conserved_point(s, k, p, load_frame(g.frame, k));
    if (DO_Q) q_point<QT>(s, k, p);
}

// the time step's denominators alone (time step called outside an iteration)
__global__ void __launch_bounds__(kBlock) k_timestep_den(DevGeom g, DevState s) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g.n) return;
    const double* p = s.prim + 5 * (long long)k;
    const double rho = p[0], u1 = p[1], u2 = p[2], u3 = p[3], pr = p[4];
    s.tsden[k] = __builtin_sqrt(u1 * u1 + u2 * u2 + u3 * u3) + 3.0 * __builtin_sqrt(pr / rho);
}

// ---------------------------------------------------------------------------
// q_derivatives.f90:17-110 (first pass) and timestep_delt.f90:20-52
// ---------------------------------------------------------------------------
// QT: type of q, qm, dq and of the gradient right-hand sides (double: the
// FP64 form; float: the mixed form); the geometric sums and the 3x3 solve
// stay FP64
// WAVES: minimum waves per SIMD of the compiled variant (register cap)
// A point is handled by kQdGroup lanes, each taking every kQdGroup-th
// neighbour; the group combines its partial sums, maxima and minima at the
// end.  Fewer points in flight per XCD keep their neighbours' q, prim and
// positions resident in its L2.
constexpr int kQdGroup = 4;
template <bool DO_DQ, bool DO_DT, bool UW, typename QT, int WAVES = 4>
__global__ void __launch_bounds__(kBlock, WAVES) k_qderiv_timestep(DevGeom g, DevState s, FlowParams fp) {
    const unsigned t = (unsigned)xcd_index();
    const int i = (int)(t / kQdGroup);
    const int lane = (int)(t % kQdGroup);
    if (i >= g.n) return;  // whole groups: blocks hold a multiple of kQdGroup lanes
    const QArrays<QT> qa = qarrays<QT>(s);
    const double4 pi4 = g.pos[i];
    const int nb = g.nbhs[i];
    const int* cn = g.conn + kConnCols * (long long)i;
    QT qi[5], qmax[5], qmin[5];
    QT sx[5], sy[5], sz[5];
    Geo6<double> geo = {0, 0, 0, 0, 0, 0};
    if (DO_DQ) {
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            qi[r] = qa.qp(i)[r];
            qmax[r] = qi[r];
            qmin[r] = qi[r];
            sx[r] = sy[r] = sz[r] = QT(0);
        }
    }
    double min_delt = 1.0;
    for (int j = lane; j < nb; j += kQdGroup) {
        const int k = cn[j];
        const double4 pk = g.pos[k];
        const double delx = pk.x - pi4.x;
        const double dely = pk.y - pi4.y;
        const double delz = pk.z - pi4.z;
        if (DO_DQ) {
            QT qk[5];
#pragma unroll
            for (int r = 0; r < 5; ++r) qk[r] = qa.qp(k)[r];
#pragma unroll
            for (int r = 0; r < 5; ++r) {
                if (qk[r] > qmax[r]) qmax[r] = qk[r];
                if (qk[r] < qmin[r]) qmin[r] = qk[r];
            }
            const double weights = wls_weight<UW>(delx, dely, delz, fp.power);
            geo.xx = geo.xx + delx * delx * weights;
            geo.yy = geo.yy + dely * dely * weights;
            geo.zz = geo.zz + delz * delz * weights;
            geo.xy = geo.xy + delx * dely * weights;
            geo.yz = geo.yz + dely * delz * weights;
            geo.zx = geo.zx + delz * delx * weights;
            const QT wx = QT(weights * delx), wy = QT(weights * dely), wz = QT(weights * delz);
#pragma unroll
            for (int r = 0; r < 5; ++r) {
                const QT t = qk[r] - qi[r];
                sx[r] = sx[r] + wx * t;
                sy[r] = sy[r] + wy * t;
                sz[r] = sz[r] + wz * t;
            }
        }
        if (DO_DT) {
            // timestep_delt.f90:34-47, the neighbour's denominator read, not recomputed
            const double dist = __builtin_sqrt(delx * delx + dely * dely + delz * delz);
            // the mixed form reads its FP32 copy, next to q
            const double den = sizeof(QT) == 4 && DO_DQ ? (double)qa.qp(k)[5] : s.tsden[k];
            const double delt = fp.cfl * dist / den;
            if (delt < min_delt) min_delt = delt;
        }
    }
    if (DO_DQ) {
        geo.xx = group_sum<kQdGroup>(geo.xx);
        geo.yy = group_sum<kQdGroup>(geo.yy);
        geo.zz = group_sum<kQdGroup>(geo.zz);
        geo.xy = group_sum<kQdGroup>(geo.xy);
        geo.yz = group_sum<kQdGroup>(geo.yz);
        geo.zx = group_sum<kQdGroup>(geo.zx);
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            sx[r] = group_sum<kQdGroup>(sx[r]);
            sy[r] = group_sum<kQdGroup>(sy[r]);
            sz[r] = group_sum<kQdGroup>(sz[r]);
            qmax[r] = group_max<kQdGroup>(qmax[r]);
            qmin[r] = group_min<kQdGroup>(qmin[r]);
        }
    }
    if (DO_DT) min_delt = group_min<kQdGroup>(min_delt);
    if (lane != 0) return;
    if (DO_DQ) {
        const double det = cramer_det(geo);
        const double one_by_det = 1.0 / det;
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            qa.qmp(i)[r * 2 + 0] = qmax[r];
            qa.qmp(i)[r * 2 + 1] = qmin[r];
#pragma unroll
            for (int c = 0; c < 3; ++c)
                qa.dqp(i)[r * 3 + c] =
                    QT(cramer_num(geo, c, (double)sx[r], (double)sy[r], (double)sz[r]) * one_by_det);
        }
    }
    if (DO_DT) s.delt[i] = min_delt;
}

// q_derivatives.f90:116-201: inner iterations (second-order q-derivatives)
template <bool UW>
__global__ void __launch_bounds__(kBlock) k_qderiv_inner(DevGeom g, DevState s, FlowParams fp,
                                                         double* temp) {
    const int i = xcd_index();
    if (i >= g.n) return;
    const double4 pi4 = g.pos[i];
    const int nb = g.nbhs[i];
    const int* cn = g.conn + kConnCols * (long long)i;
    double sx[5] = {0, 0, 0, 0, 0}, sy[5] = {0, 0, 0, 0, 0}, sz[5] = {0, 0, 0, 0, 0};
    Geo6<double> geo = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < nb; ++j) {
        const int k = cn[j];
        const double4 pk = g.pos[k];
        const double delx = pk.x - pi4.x;
        const double dely = pk.y - pi4.y;
        const double delz = pk.z - pi4.z;
        const double weights = wls_weight<UW>(delx, dely, delz, fp.power);
        geo.xx = geo.xx + delx * delx * weights;
        geo.yy = geo.yy + dely * dely * weights;
        geo.zz = geo.zz + delz * delz * weights;
        geo.xy = geo.xy + delx * dely * weights;
        geo.yz = geo.yz + dely * delz * weights;
        geo.zx = geo.zx + delz * delx * weights;
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            const double* dqi = s.dq + (5 * (long long)i + r) * 3;
            const double* dqk = s.dq + (5 * (long long)k + r) * 3;
            const double qti = s.q[5 * (long long)i + r] - 0.5 * (delx * dqi[0] + dely * dqi[1] + delz * dqi[2]);
            const double qtn = s.q[5 * (long long)k + r] - 0.5 * (delx * dqk[0] + dely * dqk[1] + delz * dqk[2]);
            const double t = qtn - qti;
            sx[r] = sx[r] + weights * delx * t;
            sy[r] = sy[r] + weights * dely * t;
            sz[r] = sz[r] + weights * delz * t;
        }
    }
    const double det = cramer_det(geo);
    const double one_by_det = 1.0 / det;
#pragma unroll
    for (int r = 0; r < 5; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c)
            temp[(5 * (long long)i + r) * 3 + c] = cramer_num(geo, c, sx[r], sy[r], sz[r]) * one_by_det;
}

// m values per point between an FP64 array of m per point and an FP32 array
// of stride32 >= m per point
__global__ void k_to_float(float* dst, int stride32, const double* src, int m, long long n) {
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= m * n) return;
    const long long k = e / m;
    dst[stride32 * k + (e - k * m)] = (float)src[e];
}

__global__ void k_to_double(double* dst, const float* src, int stride32, int m, long long n) {
    const long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (e >= m * n) return;
    const long long k = e / m;
    dst[e] = (double)src[stride32 * k + (e - k * m)];
}

__global__ void k_copy(double* dst, const double* src, long long n) {
    const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// ---------------------------------------------------------------------------
// flux stencils in list order: interior_flux_dG*.f90, wall_flux_dG*.f90,
// outer_flux_dG*.f90
// ---------------------------------------------------------------------------
struct Stencil {
    const int* list;
    int count;
    int family;   // 0 half-range, 1 wall octant, 2 outer octant
    int axis;
    int sign;
    bool by_one_by_det;
};

// the list a sweep walks: the flux stencil's, except an outer point's z
// term, which walks conn (implicit_LdelU_forward_sweep_outer.f90:346-347,
// reached only from the alternative sweep driver)
__device__ __forceinline__ Stencil sweep_stencil(const DevGeom& g, int cls, int d, int i);

__device__ __forceinline__ Stencil dir_stencil(const DevGeom& g, int cls, int d, int i) {
    Stencil st;
    st.list = g.lst + g.dir_off[(long long)d * g.n + i];
    st.count = g.dir_cnt[(long long)d * g.n + i];
    st.family = dir_family(cls, d);
    st.axis = d >> 1;
    st.sign = (d & 1) ? -1 : 1;
    st.by_one_by_det = cls == 2;
    return st;
}

__device__ __forceinline__ Stencil sweep_stencil(const DevGeom& g, int cls, int d, int i) {
    Stencil st = dir_stencil(g, cls, d, i);
    if (cls == 2 && d == ZPOS) {
        st.list = g.conn + kConnCols * (long long)i;
        st.count = g.nbhs[i];
    }
    st.by_one_by_det = false;
    return st;
}

// T: arithmetic of the reconstruction and fluxes (double: FP64 form; float:
// mixed form); the least-squares sums and the 3x3 solve stay FP64
template <bool UW, typename T>
__device__ void stencil_dG(const DevGeom& g, const DevState& s, const FlowParams& fp, int i,
                           const Stencil& st, const Frame<double>& f, double G[5]) {
    const QArrays<T> qa = qarrays<T>(s);
    const Frame<T> ft = frame_as<T>(f);
    Geo6<double> geo = {0, 0, 0, 0, 0, 0};
    double sx[5] = {0, 0, 0, 0, 0}, sy[5] = {0, 0, 0, 0, 0}, sz[5] = {0, 0, 0, 0, 0};
    const double4 pi4 = g.pos[i];
    T qi[5];
#pragma unroll
    for (int r = 0; r < 5; ++r) qi[r] = qa.qp(i)[r];
    const T epsi2_i = T(limiter_epsi2(fp.vl_const, pi4.w));
    for (int j = 0; j < st.count; ++j) {
        const int k = st.list[j];
        const double4 pk = g.pos[k];
        const double delx = pk.x - pi4.x;
        const double dely = pk.y - pi4.y;
        const double delz = pk.z - pi4.z;
        const double dels = delx * f.t1[0] + dely * f.t1[1] + delz * f.t1[2];
        const double delt = delx * f.t2[0] + dely * f.t2[1] + delz * f.t2[2];
        const double deln = delx * f.n[0] + dely * f.n[1] + delz * f.n[2];
        const double weights = wls_weight<UW>(dels, delt, deln, fp.power);
        const double dels_w = dels * weights, delt_w = delt * weights, deln_w = deln * weights;
        geo.xx = geo.xx + dels * dels_w;
        geo.yy = geo.yy + delt * delt_w;
        geo.zz = geo.zz + deln * deln_w;
        geo.xy = geo.xy + dels * delt_w;
        geo.yz = geo.yz + delt * deln_w;
        geo.zx = geo.zx + deln * dels_w;
        T prim[5], Gi[5], Gk[5], qk[5];
        reconstruct_end<T>(qi, qa.dqp(i), qa.qmp(i), epsi2_i, T(delx), T(dely),
                           T(delz), prim);
        flux_eval<T>(st.family, st.axis, st.sign, ft, prim, T(fp.pi), Gi);
#pragma unroll
        for (int r = 0; r < 5; ++r) qk[r] = qa.qp(k)[r];
        reconstruct_end<T>(qk, qa.dqp(k), qa.qmp(k),
                           T(limiter_epsi2(fp.vl_const, pk.w)), T(delx), T(dely), T(delz), prim);
        flux_eval<T>(st.family, st.axis, st.sign, ft, prim, T(fp.pi), Gk);
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            const double t = (double)(Gk[r] - Gi[r]);
            sx[r] = sx[r] + t * dels_w;
            sy[r] = sy[r] + t * delt_w;
            sz[r] = sz[r] + t * deln_w;
        }
    }
    const double det = cramer_det(geo);
    if (st.by_one_by_det) {
        const double one_by_det = 1.0 / det;
#pragma unroll
        for (int r = 0; r < 5; ++r) G[r] = cramer_num(geo, st.axis, sx[r], sy[r], sz[r]) * one_by_det;
    } else {
#pragma unroll
        for (int r = 0; r < 5; ++r) G[r] = cramer_num(geo, st.axis, sx[r], sy[r], sz[r]) / det;
    }
}

// flux_residual.f90:40-81 for the points of idx (any class, from status).
// A point is handled by kListGroup lanes: lane l evaluates the point's l-th
// direction list (x+, x-, y+, y-, then z+ and z- for interior points, z-
// over conn for wall points, z+ for outer points), and every lane of the
// group then adds the list results in that order, as flux_residual.f90 does.
// The lists of a point run side by side instead of one after another.
constexpr int kListGroup = 8;
template <bool UW, typename T>
__global__ void __launch_bounds__(kBlock) k_flux_residual_lists(DevGeom g, DevState s, FlowParams fp,
                                                                const int* idx, int count) {
    const unsigned t = (unsigned)xcd_index();
    const int m = (int)(t / kListGroup);
    const int lane = (int)(t % kListGroup);
    if (m >= count) return;  // whole groups: blocks hold a multiple of kListGroup lanes
    const int i = idx[m];
    const int cls = g.status[i];
    const int n_lists = cls == 0 ? 6 : (cls == 1 || cls == 2 ? 5 : 4);
    double G[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    if (lane < n_lists) {
        const int d = lane < 4 || cls == 0 ? lane : (cls == 1 ? ZNEG : ZPOS);
        const Frame<double> f = load_frame(g.frame, i);
        stencil_dG<UW, T>(g, s, fp, i, dir_stencil(g, cls, d, i), f, G);
    }
    double acc[5], v4[5], v5[5];
#pragma unroll
    for (int r = 0; r < 5; ++r) {
        acc[r] = __shfl(G[r], 0, kListGroup);
#pragma unroll
        for (int l = 1; l < 4; ++l) acc[r] = acc[r] + __shfl(G[r], l, kListGroup);
        v4[r] = __shfl(G[r], 4, kListGroup);
        v5[r] = __shfl(G[r], 5, kListGroup);
    }
    if (lane != 0) return;
#pragma unroll
    for (int r = 0; r < 5; ++r) {
        if (cls == 0)
            acc[r] = (acc[r] + v4[r]) + v5[r];
        else if (cls == 1)
            acc[r] = 2.0 * (acc[r] + v4[r]);
        else if (cls == 2)
            acc[r] = acc[r] + v4[r];
        s.flux_res[5 * (long long)i + r] = acc[r];
    }
}

// explicit update through the sweep's outputs: delUn = -delt*R, delUp = delUn
__global__ void __launch_bounds__(kBlock) k_explicit_sweep(DevGeom g, DevState s) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= g.n) return;
    const double dt = s.delt[k];
#pragma unroll
    for (int r = 0; r < 5; ++r) {
        const double v = -(dt * s.flux_res[5 * (long long)k + r]);
        s.delUn[5 * (long long)k + r] = v;
        s.delUp[5 * (long long)k + r] = v;
    }
}

// ---------------------------------------------------------------------------
// implicit LU-SGS: implicit_D_inv_{interior,wall}.f90 and the alias-ordered
// sweeps of implicit_forward_and_reverse_sweeps.f90:26-88
// ---------------------------------------------------------------------------
// The geometric part of one direction's D_inv term
// (implicit_D_inv_interior.f90): over the direction list, the normal matrix of
// the weighted least-squares fit and the right-hand side (sum dels*w,
// sum delt*w, sum deln*w); returns the Cramer numerator of the list's axis
// and the determinant.  Positions, frames and weights only, so it is computed
// once, at bind.
template <bool UW>
__device__ void dinv_geometry(const DevGeom& g, double power, int i, const Stencil& st, const Frame<double>& f,
                              double& temp, double& det) {
    Geo6<double> geo = {0, 0, 0, 0, 0, 0};
    double sx = 0.0, sy = 0.0, sz = 0.0;
    const double4 pi4 = g.pos[i];
    for (int j = 0; j < st.count; ++j) {
        const double4 pk = g.pos[st.list[j]];
        const double delx = pk.x - pi4.x;
        const double dely = pk.y - pi4.y;
        const double delz = pk.z - pi4.z;
        const double dels = delx * f.t1[0] + dely * f.t1[1] + delz * f.t1[2];
        const double delt = delx * f.t2[0] + dely * f.t2[1] + delz * f.t2[2];
        const double deln = delx * f.n[0] + dely * f.n[1] + delz * f.n[2];
        const double weights = wls_weight<UW>(dels, delt, deln, power);
        const double dels_w = dels * weights, delt_w = delt * weights, deln_w = deln * weights;
        geo.xx = geo.xx + dels * dels_w;
        geo.yy = geo.yy + delt * delt_w;
        geo.zz = geo.zz + deln * deln_w;
        geo.xy = geo.xy + dels * delt_w;
        geo.yz = geo.yz + delt * deln_w;
        geo.zx = geo.zx + deln * dels_w;
        sx = sx + dels_w;
        sy = sy + delt_w;
        sz = sz + deln_w;
    }
    det = cramer_det(geo);
    temp = cramer_num(geo, st.axis, sx, sy, sz);
}

// setup: the numerator and determinant of all six direction lists of every
// point, dgeo[12 i + 2 d] and dgeo[12 i + 2 d + 1], over the lists the sweeps
// walk (sweep_stencil).  Outer points, swept only by the alternative driver,
// take ZPOS over conn (implicit_D_inv_outer.f90:306-353), and their ZNEG
// slot holds what implicit_D_inv_outer.f90:360 divides 1 by: its local delt
// is also the loop's tan2 projection (:73 ...), so after the conn loop it is
// the last conn neighbour's tan2 offset -- (that offset, 1), or (0, 0) when
// the point has no neighbours and delt keeps the time step.
template <bool UW>
__global__ void __launch_bounds__(kBlock) k_dinv_geometry(DevGeom g, double* dgeo, double power) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.n) return;
    const Frame<double> f = load_frame(g.frame, i);
    const int cls = g.status[i];
    for (int d = XPOS; d <= ZNEG; ++d) {
        double temp = 0.0, det = 0.0;
        if (cls == 2 && d == ZNEG) {
            const int nb = g.nbhs[i];
            if (nb > 0) {
                const double4 pi4 = g.pos[i];
                const double4 pk = g.pos[g.conn[kConnCols * (long long)i + nb - 1]];
                const double delx = pk.x - pi4.x, dely = pk.y - pi4.y, delz = pk.z - pi4.z;
                temp = delx * f.t2[0] + dely * f.t2[1] + delz * f.t2[2];
                det = 1.0;
            }
        } else {
            dinv_geometry<UW>(g, power, i, sweep_stencil(g, cls, d, i), f, temp, det);
        }
        dgeo[12 * (long long)i + 2 * d] = temp;
        dgeo[12 * (long long)i + 2 * d + 1] = det;
    }
}

// One direction's D_inv term from its precomputed geometry: the split
// spectral radius of the global-frame velocity component with the sound
// speed (implicit_D_inv_interior.f90:107-108,168-169) times numerator /
// determinant, in the reference's operation order
__device__ __forceinline__ double dinv_dir(const DevState& s, int i, int d, const double* gd, double as) {
    const double temp = gd[2 * d], det = gd[2 * d + 1];
    const double vel = s.prim[5 * (long long)i + 1 + (d >> 1)];
    if ((d & 1) == 0) {
        const double temp1 = vel + as;
        return 0.50 * fabs(temp1 + fabs(temp1)) * temp / det;
    }
    const double temp1 = vel - as;
    return -(0.50 * fabs(temp1 - fabs(temp1)) * temp / det);
}

__global__ void __launch_bounds__(kBlock) k_dinv(DevGeom g, DevState s, const int* idx, int count) {
    const int m = xcd_index();
    if (m >= count) return;
    const int i = idx[m];
    const int cls = g.status[i];
    const double* gd = g.dgeo + 12 * (long long)i;
    const double delta_t = s.delt[i];
    const double as = __builtin_sqrt(1.40 * s.prim[5 * (long long)i + 4] / s.prim[5 * (long long)i]);
    double d[6];
    double temp;
    if (cls == 2) {
        // implicit_D_inv_outer.f90:44-362: the negative halves enter with a
        // plus sign (:171, :295), the sum is taken left to right (:360)
        const double dGxp = dinv_dir(s, i, XPOS, gd, as);
        const double dGxn = -dinv_dir(s, i, XNEG, gd, as);
        const double dGyp = dinv_dir(s, i, YPOS, gd, as);
        const double dGyn = -dinv_dir(s, i, YNEG, gd, as);
        const double dGzp = dinv_dir(s, i, ZPOS, gd, as);
        const double delt = gd[2 * ZNEG + 1] != 0.0 ? gd[2 * ZNEG] : delta_t;
        temp = (1.0 / delt) - dGxp + dGxn - dGyp + dGyn - dGzp;
    } else if (cls == 1) {
        for (int dd = XPOS; dd <= YNEG; ++dd) d[dd] = dinv_dir(s, i, dd, gd, as);
        d[4] = dinv_dir(s, i, ZNEG, gd, as);
        temp = (1.0 / delta_t) - 2.0 * (d[0] + d[1] + d[2] + d[3] + d[4]);
    } else {
        for (int dd = XPOS; dd <= ZNEG; ++dd) d[dd] = dinv_dir(s, i, dd, gd, as);
        temp = (1.0 / delta_t) - (d[0] + d[1] + d[2] + d[3] + d[4] + d[5]);
    }
    s.dinv[i] = 1.0 / temp;
}

// one direction of LdelU (FORWARD) or UdelU
// (implicit_LdelU_forward_sweep_interior.f90:51-117)
template <bool FORWARD, bool UW>
__device__ void sweep_dir(const DevGeom& g, const DevState& s, const FlowParams& fp, int i,
                          const Stencil& st, const Frame<double>& f, double dG[5]) {
    Geo6<double> geo = {0, 0, 0, 0, 0, 0};
    double sx[5] = {0, 0, 0, 0, 0}, sy[5] = {0, 0, 0, 0, 0}, sz[5] = {0, 0, 0, 0, 0};
    const double4 pi4 = g.pos[i];
    const int alias_i = g.alias[i];
    const double* dU = FORWARD ? s.delUp : s.delUn;
    for (int j = 0; j < st.count; ++j) {
        const int k = st.list[j];
        const double4 pk = g.pos[k];
        const double delx = pk.x - pi4.x;
        const double dely = pk.y - pi4.y;
        const double delz = pk.z - pi4.z;
        const double dels = delx * f.t1[0] + dely * f.t1[1] + delz * f.t1[2];
        const double delt = delx * f.t2[0] + dely * f.t2[1] + delz * f.t2[2];
        const double deln = delx * f.n[0] + dely * f.n[1] + delz * f.n[2];
        const double weights = wls_weight<UW>(dels, delt, deln, fp.power);
        const double dels_w = dels * weights, delt_w = delt * weights, deln_w = deln * weights;
        geo.xx = geo.xx + dels * dels_w;
        geo.yy = geo.yy + delt * delt_w;
        geo.zz = geo.zz + deln * deln_w;
        geo.xy = geo.xy + dels * delt_w;
        geo.yz = geo.yz + delt * deln_w;
        geo.zx = geo.zx + deln * dels_w;
        const int alias_k = g.alias[k];
        const bool take = FORWARD ? (alias_k < alias_i) : (alias_k > alias_i);
        if (take) {
            double prim[5], Gold[5], Gnew[5], U[5];
#pragma unroll
            for (int r = 0; r < 5; ++r) prim[r] = s.prim[5 * (long long)k + r];
            flux_eval(st.family, st.axis, st.sign, f, prim, fp.pi, Gold);
#pragma unroll
            for (int r = 0; r < 5; ++r) U[r] = dU[5 * (long long)k + r] + s.U[5 * (long long)k + r];
            const Frame<double> fk = load_frame(g.frame, k);
            conserved_to_prim(U, fk, prim);
            flux_eval(st.family, st.axis, st.sign, f, prim, fp.pi, Gnew);
#pragma unroll
            for (int r = 0; r < 5; ++r) {
                const double t = Gnew[r] - Gold[r];
                sx[r] = sx[r] + t * dels_w;
                sy[r] = sy[r] + t * delt_w;
                sz[r] = sz[r] + t * deln_w;
            }
        }
    }
    const double det = cramer_det(geo);
#pragma unroll
    for (int r = 0; r < 5; ++r) dG[r] = cramer_num(geo, st.axis, sx[r], sy[r], sz[r]) / det;
}

template <bool FORWARD, bool UW>
__device__ void sweep_term(const DevGeom& g, const DevState& s, const FlowParams& fp, int i,
                           const Frame<double>& f, double out[5]) {
    double d[5];
    if (g.status[i] == 2) {
        // implicit_LdelU_forward_sweep_outer.f90:50-406: Go over the split
        // x/y lists, Gzp over conn, summed without a factor
        sweep_dir<FORWARD, UW>(g, s, fp, i, sweep_stencil(g, 2, XPOS, i), f, out);
#pragma unroll 1
        for (int dd = XNEG; dd <= ZPOS; ++dd) {
            sweep_dir<FORWARD, UW>(g, s, fp, i, sweep_stencil(g, 2, dd, i), f, d);
#pragma unroll
            for (int r = 0; r < 5; ++r) out[r] = out[r] + d[r];
        }
        return;
    }
    if (g.status[i] == 1) {
        sweep_dir<FORWARD, UW>(g, s, fp, i, dir_stencil(g, 1, XPOS, i), f, out);
#pragma unroll 1
        for (int dd = XNEG; dd <= YNEG; ++dd) {
            sweep_dir<FORWARD, UW>(g, s, fp, i, dir_stencil(g, 1, dd, i), f, d);
#pragma unroll
            for (int r = 0; r < 5; ++r) out[r] = out[r] + d[r];
        }
        sweep_dir<FORWARD, UW>(g, s, fp, i, dir_stencil(g, 1, ZNEG, i), f, d);
#pragma unroll
        for (int r = 0; r < 5; ++r) out[r] = 2.0 * (out[r] + d[r]);
        return;
    }
    sweep_dir<FORWARD, UW>(g, s, fp, i, dir_stencil(g, 0, XPOS, i), f, out);
#pragma unroll 1
    for (int dd = XNEG; dd <= ZNEG; ++dd) {
        sweep_dir<FORWARD, UW>(g, s, fp, i, dir_stencil(g, 0, dd, i), f, d);
#pragma unroll
        for (int r = 0; r < 5; ++r) out[r] = out[r] + d[r];
    }
}

// nearest neighbour of a supersonic point (implicit_forward_and_reverse_
// sweeps.f90:39-50, state_update.f90:84-94); skip_wall: the alternative
// driver's status /= 1 filter (implicit_forward_and_reverse_sweeps_backup.f90:
// 56, state_update_implicit.f90:72,89)
__device__ int nearest_neighbour_dev(const DevGeom& g, int k, double start, bool skip_wall) {
    const double4 pk = g.pos[k];
    double min_dist = start;
    int p = k;
    for (int r = 0; r < g.nbhs[k]; ++r) {
        const int nbh = g.conn[kConnCols * (long long)k + r];
        const double4 pn = g.pos[nbh];
        const double dx = pn.x - pk.x, dy = pn.y - pk.y, dz = pn.z - pk.z;
        const double ds = __builtin_sqrt(dx * dx + dy * dy + dz * dz);
        if (ds < min_dist && !(skip_wall && g.status[nbh] == 1)) {
            min_dist = ds;
            p = nbh;
        }
    }
    return p;
}

// one level of a sweep, in place: the level schedule orders every point
// after the points whose new values it reads and after the points that
// read its old value, so in-place reads see what the reference's visiting
// order sees
template <bool FORWARD, bool UW>
__global__ void __launch_bounds__(64) k_sweep_level(DevGeom g, DevState s, FlowParams fp,
                                                    const int* points, int begin, int end) {
    const int m = begin + blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= end) return;
    const int p = points[m];
    const int cls = g.status[p];
    if (cls == 5 || cls == 6) {
        // implicit_forward_and_reverse_sweeps.f90:38-52, 69-83
        const int q = nearest_neighbour_dev(g, p, fp.backup ? 100000.0 : 1000000.0, fp.backup != 0);
        double* dU = FORWARD ? s.delUp : s.delUn;
#pragma unroll
        for (int r = 0; r < 5; ++r) dU[5 * (long long)p + r] = dU[5 * (long long)q + r];
        return;
    }
    const Frame<double> f = load_frame(g.frame, p);
    double T5[5];
    sweep_term<FORWARD, UW>(g, s, fp, p, f, T5);
    const double D_inv = s.dinv[p];
    if (FORWARD) {
#pragma unroll
        for (int r = 0; r < 5; ++r)
            s.delUp[5 * (long long)p + r] = -(D_inv * (s.flux_res[5 * (long long)p + r] + T5[r]));
    } else {
#pragma unroll
        for (int r = 0; r < 5; ++r)
            s.delUn[5 * (long long)p + r] = s.delUp[5 * (long long)p + r] - D_inv * T5[r];
    }
}

// ---------------------------------------------------------------------------
// state_update.f90:22-115 with the residual of state_update.f90:39-44,115
// ---------------------------------------------------------------------------
__device__ void conserved_vector_Ubar(const DevGeom& g, const DevState& s, FlowParams fp, int k,
                                      const Frame<double>& f, double Ubar[5]) {
    // conserved_vector_Ubar.f90:19-58; a point of a batch takes its cloud's freestream
    if (g.inst) {
        const int b = g.inst[k];
        fp.u1_inf = g.inst_fs[2 * b];
        fp.u2_inf = g.inst_fs[2 * b + 1];
    }
    const double pi = fp.pi;
    const double u1_inf_rot = fp.u1_inf * f.t1[0] + fp.u2_inf * f.t1[1] + fp.u3_inf * f.t1[2];
    const double u2_inf_rot = fp.u1_inf * f.t2[0] + fp.u2_inf * f.t2[1] + fp.u3_inf * f.t2[2];
    const double u3_inf_rot = fp.u1_inf * f.n[0] + fp.u2_inf * f.n[1] + fp.u3_inf * f.n[2];
    const double rho_e_inf = fp.pr_inf * 2.5 + 0.5 * fp.rho_inf *
        (u1_inf_rot * u1_inf_rot + u2_inf_rot * u2_inf_rot + u3_inf_rot * u3_inf_rot);
    double beta = (0.5 * fp.rho_inf) / fp.pr_inf;
    double S3 = u3_inf_rot * __builtin_sqrt(beta);
    const double B3_inf = exp(-S3 * S3) / (2.0 * __builtin_sqrt(pi * beta));
    const double A3n_inf = 0.5 * (1.0 - erf(S3));
    const double* p = s.prim + 5 * (long long)k;
    const double rho = p[0], u1 = p[1], u2 = p[2], u3 = p[3], pr = p[4];
    const double u1_rot = u1 * f.t1[0] + u2 * f.t1[1] + u3 * f.t1[2];
    const double u2_rot = u1 * f.t2[0] + u2 * f.t2[1] + u3 * f.t2[2];
    const double u3_rot = u1 * f.n[0] + u2 * f.n[1] + u3 * f.n[2];
    const double rho_e = pr * 2.5 + 0.5 * rho * (u1_rot * u1_rot + u2_rot * u2_rot + u3_rot * u3_rot);
    beta = (0.5 * rho) / pr;
    S3 = u3_rot * __builtin_sqrt(beta);
    const double B3 = exp(-S3 * S3) / (2.0 * __builtin_sqrt(pi * beta));
    const double A3p = 0.5 * (1.0 + erf(S3));
    Ubar[0] = (fp.rho_inf * A3n_inf) + (rho * A3p);
    Ubar[1] = (fp.rho_inf * u1_inf_rot * A3n_inf) + (rho * u1_rot * A3p);
    Ubar[2] = (fp.rho_inf * u2_inf_rot * A3n_inf) + (rho * u2_rot * A3p);
    double temp1 = fp.rho_inf * (u3_inf_rot * A3n_inf - B3_inf);
    double temp2 = rho * (u3_rot * A3p + B3);
    Ubar[3] = temp1 + temp2;
    temp1 = (rho_e_inf * A3n_inf - 0.5 * fp.rho_inf * u3_inf_rot * B3_inf);
    temp2 = (rho_e * A3p + 0.5 * rho * u3_rot * B3);
    Ubar[4] = temp1 + temp2;
}

struct MaxPair {
    double v;
    int pos;   // position in the reference's wall-then-interior order
    int point;
};

__device__ __forceinline__ MaxPair max_pair(MaxPair a, MaxPair b) {
    if (b.v > a.v) return b;
    if (b.v == a.v && b.pos < a.pos) return b;
    return a;
}

// EXPLICIT: the explicit form's sweep (k_explicit_sweep: delUn = delUp =
// -delt*R for every point) is done here first, with the same arithmetic, so
// delUn and delUp are written once instead of written and read back
// BACKUP: state_update_implicit.f90:22-60, paired with the alternative
// sweep driver: wall points keep U(4), nothing is copied into delUp here,
// outer points take Ubar + delUn
template <bool EXPLICIT, bool BACKUP = false>
__global__ void __launch_bounds__(kBlock) k_state_update(DevGeom g, DevState s, FlowParams fp,
                                                         double* partial, int* partial_pos,
                                                         int* partial_pt, double* partial_max,
                                                         double* ressq) {
    constexpr int kWaves = kBlock / 64;
    __shared__ double ssum[kWaves];
    __shared__ double smax[kWaves];
    __shared__ int spos[kWaves];
    __shared__ int spt[kWaves];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    double res_sqr = 0.0;
    MaxPair mp = {0.0, 0x7fffffff, 0};
    if (k < g.n) {
        const int st = g.status[k];
        double dU[5];
        if (EXPLICIT) {
            const double dt = s.delt[k];
#pragma unroll
            for (int r = 0; r < 5; ++r) {
                dU[r] = -(dt * s.flux_res[5 * (long long)k + r]);
                s.delUn[5 * (long long)k + r] = dU[r];
                s.delUp[5 * (long long)k + r] = dU[r];
            }
        }
        if (st == 0 || st == 1) {
            const Frame<double> f = load_frame(g.frame, k);
            const double temp = s.U[5 * (long long)k];
            if (!EXPLICIT) {
#pragma unroll
                for (int r = 0; r < 5; ++r) dU[r] = s.delUn[5 * (long long)k + r];
            }
            double U[5];
#pragma unroll
            for (int r = 0; r < 5; ++r) U[r] = s.U[5 * (long long)k + r] + dU[r];
            if (st == 1 && !BACKUP) U[3] = 0.0;
            res_sqr = (U[0] - temp) * (U[0] - temp);
            mp.v = res_sqr;
            mp.pos = g.respos[k];
            mp.point = g.perm[k] + 1;
            double prim[5];
            conserved_to_prim(U, f, prim);
#pragma unroll
            for (int r = 0; r < 5; ++r) {
                s.prim[5 * (long long)k + r] = prim[r];
                if (!EXPLICIT && !BACKUP) s.delUp[5 * (long long)k + r] = dU[r];
            }
        } else if (st == 2) {
            const Frame<double> f = load_frame(g.frame, k);
            double U[5], prim[5];
            conserved_vector_Ubar(g, s, fp, k, f, U);
            if (BACKUP) {
                // state_update_implicit.f90:45-47
#pragma unroll
                for (int r = 0; r < 5; ++r) U[r] = U[r] + s.delUn[5 * (long long)k + r];
            } else {
                const double dt = s.delt[k];
#pragma unroll
                for (int r = 0; r < 5; ++r) U[r] = U[r] - dt * s.flux_res[5 * (long long)k + r];
            }
            conserved_to_prim(U, f, prim);
#pragma unroll
            for (int r = 0; r < 5; ++r) s.prim[5 * (long long)k + r] = prim[r];
        }
    }
    if (ressq && k < g.n) ressq[k] = res_sqr;
    // block partials: a butterfly over each wave (every lane ends with the
    // same, order-fixed sum), then the waves' results in wave order
    double sum = res_sqr;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        sum = sum + __shfl_xor(sum, o, 64);
        const MaxPair b = {__shfl_xor(mp.v, o, 64), __shfl_xor(mp.pos, o, 64), __shfl_xor(mp.point, o, 64)};
        mp = max_pair(mp, b);
    }
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        ssum[wave] = sum;
        smax[wave] = mp.v;
        spos[wave] = mp.pos;
        spt[wave] = mp.point;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double total = ssum[0];
        MaxPair best = {smax[0], spos[0], spt[0]};
#pragma unroll
        for (int w = 1; w < kWaves; ++w) {
            total = total + ssum[w];
            best = max_pair(best, MaxPair{smax[w], spos[w], spt[w]});
        }
        partial[blockIdx.x] = total;
        partial_max[blockIdx.x] = best.v;
        partial_pos[blockIdx.x] = best.pos;
        partial_pt[blockIdx.x] = best.point;
    }
}

// state_update.f90:80-114: supersonic points copy the nearest neighbour's
// state after every other point is updated (outlet list, then inlet list,
// each in order, so a single thread walks them)
__global__ void k_supersonic(DevGeom g, DevState s, const int* out_idx, int n_out, const int* in_idx,
                             int n_in, int skip_wall) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    for (int pass = 0; pass < 2; ++pass) {
        const int* idx = pass == 0 ? out_idx : in_idx;
        const int cnt = pass == 0 ? n_out : n_in;
        for (int m = 0; m < cnt; ++m) {
            const int k = idx[m];
            const int p = nearest_neighbour_dev(g, k, 100000.0, skip_wall != 0);
            for (int r = 0; r < 5; ++r) s.prim[5 * (long long)k + r] = s.prim[5 * (long long)p + r];
        }
    }
}

__global__ void __launch_bounds__(1024) k_residual_final(const double* partial, const int* partial_pos,
                                                         const int* partial_pt, const double* partial_max,
                                                         int n_partial, int n_points,
                                                         ResidualRecord* rec, double* hist, int hist_index,
                                                         const int* hist_counter) {
    __shared__ double ssum[1024];
    __shared__ double smax[1024];
    __shared__ int spos[1024];
    __shared__ int spt[1024];
    double acc = 0.0;
    MaxPair mp = {0.0, 0x7fffffff, 0};
    // fixed assignment of partials to threads and a fixed tree: deterministic
    for (int b = threadIdx.x; b < n_partial; b += blockDim.x) {
        acc = acc + partial[b];
        MaxPair o = {partial_max[b], partial_pos[b], partial_pt[b]};
        mp = max_pair(mp, o);
    }
    ssum[threadIdx.x] = acc;
    smax[threadIdx.x] = mp.v;
    spos[threadIdx.x] = mp.pos;
    spt[threadIdx.x] = mp.point;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            ssum[threadIdx.x] = ssum[threadIdx.x] + ssum[threadIdx.x + w];
            MaxPair a = {smax[threadIdx.x], spos[threadIdx.x], spt[threadIdx.x]};
            MaxPair b = {smax[threadIdx.x + w], spos[threadIdx.x + w], spt[threadIdx.x + w]};
            a = max_pair(a, b);
            smax[threadIdx.x] = a.v;
            spos[threadIdx.x] = a.pos;
            spt[threadIdx.x] = a.point;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double sum = ssum[0];
        rec->sum_res_sqr = sum;
        rec->res_new = __builtin_sqrt(sum) / n_points;
        rec->max_res = smax[0];
        // the reference keeps the previous max_res_point when no residual is
        // positive; 0 here means "unchanged"
        rec->max_res_point = smax[0] > 0.0 ? spt[0] : 0;
        if (hist) hist[hist_counter ? *hist_counter : hist_index] = rec->res_new;
    }
}

// residual of each cloud of a batch (state_update.f90:39-44,115 over the
// cloud's own points): block b sums its cloud's per-point terms in a fixed
// order; the history row is the iteration counter's
// Per-cloud residuals of a batch in two fixed-order passes: every chunk of a
// cloud's range (in order per thread, then a pairwise tree), then every
// cloud over its chunks' sums the same way.  The chunks spread each cloud
// over many workgroups, where one workgroup per cloud left the largest
// cloud's sum latency-bound.
__global__ void __launch_bounds__(kBlock) k_chunk_sums(const double* ressq, const int* chunk_beg,
                                                       const int* chunk_end, double* chunk_sum) {
    __shared__ double ssum[kBlock];
    const int c = blockIdx.x;
    double acc = 0.0;
    for (int k = chunk_beg[c] + threadIdx.x; k < chunk_end[c]; k += kBlock) acc = acc + ressq[k];
    ssum[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) ssum[threadIdx.x] = ssum[threadIdx.x] + ssum[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) chunk_sum[c] = ssum[0];
}

__global__ void __launch_bounds__(kBlock) k_instance_residual(const double* chunk_sum, const int* chunk_ptr,
                                                             const double* norm, int n_inst, double* inst_hist,
                                                             const int* counter) {
    __shared__ double ssum[kBlock];
    const int b = blockIdx.x;
    double acc = 0.0;
    for (int c = chunk_ptr[b] + threadIdx.x; c < chunk_ptr[b + 1]; c += kBlock) acc = acc + chunk_sum[c];
    ssum[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) ssum[threadIdx.x] = ssum[threadIdx.x] + ssum[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) inst_hist[(long long)(*counter) * n_inst + b] = __builtin_sqrt(ssum[0]) / norm[b];
}

__global__ void k_bump(int* counter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *counter = *counter + 1;
}

}  // namespace

// implemented in kernels_flux_fused.hip
void launch_flux_fused(const DevGeom& g, const DevState& s, const FlowParams& p, const TypeLists& tl,
                       hipStream_t st, int parts, bool mixed, const LaunchShapes& ls);

// ---------------------------------------------------------------------------
// launch wrappers
// ---------------------------------------------------------------------------
void launch_stencil_coefficients(const DevGeom& g, double* coef, double power, hipStream_t st) {
    if (g.n <= 0) return;
    if (power == 0.0)
        hipLaunchKernelGGL(k_stencil_coef<true>, dim3(grid_for(g.n, kBlock)), dim3(kBlock), 0, st, g, coef, power);
    else
        hipLaunchKernelGGL(k_stencil_coef<false>, dim3(grid_for(g.n, kBlock)), dim3(kBlock), 0, st, g, coef,
                           power);
}
void launch_stencil_check(const DevGeom& g, double power, int* first_bad, hipStream_t st) {
    if (g.n <= 0) return;
    if (power == 0.0)
        hipLaunchKernelGGL(k_stencil_check<true>, dim3(grid_for(g.n, kBlock)), dim3(kBlock), 0, st, g, power,
                           first_bad);
    else
        hipLaunchKernelGGL(k_stencil_check<false>, dim3(grid_for(g.n, kBlock)), dim3(kBlock), 0, st, g, power,
                           first_bad);
}
void launch_gather(double* dst, const double* src, const int* perm, int n, int m, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_gather, dim3(grid_for((long long)n * m, kBlock)), dim3(kBlock), 0, st, dst, src, perm,
                       n, m);
}
void launch_to_float(float* dst, int stride32, const double* src, int m, long long n, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_to_float, dim3(grid_for(m * n, kBlock)), dim3(kBlock), 0, st, dst, stride32, src, m, n);
}
void launch_to_double(double* dst, const float* src, int stride32, int m, long long n, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_to_double, dim3(grid_for(m * n, kBlock)), dim3(kBlock), 0, st, dst, src, stride32, m, n);
}
void launch_scatter(double* dst, const double* src, const int* perm, int n, int m, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_scatter, dim3(grid_for((long long)n * m, kBlock)), dim3(kBlock), 0, st, dst, src,
                       perm, n, m);
}
void launch_conserved_vector(const DevGeom& g, const DevState& s, hipStream_t st) {
    if (g.n <= 0) return;
    hipLaunchKernelGGL((k_conserved_q<true, false, double>), dim3(grid_for(g.n, kBlock)), dim3(kBlock), 0, st, g,
                       s);
}
void launch_q_variables(const DevGeom& g, const DevState& s, const FlowParams& p, hipStream_t st) {
    if (g.n <= 0) return;
    if (p.mixed)
        hipLaunchKernelGGL((k_conserved_q<false, true, float>), dim3(grid_for(g.n, kBlock)), dim3(kBlock), 0, st,
                           g, s);
    else
        hipLaunchKernelGGL((k_conserved_q<false, true, double>), dim3(grid_for(g.n, kBlock)), dim3(kBlock), 0, st,
                           g, s);
}
void launch_conserved_and_q(const DevGeom& g, const DevState& s, const FlowParams& p, hipStream_t st) {
    if (g.n <= 0) return;
    if (p.mixed)
        hipLaunchKernelGGL((k_conserved_q<true, true, float>), dim3(grid_for(g.n, kBlock)), dim3(kBlock), 0, st,
                           g, s);
    else
        hipLaunchKernelGGL((k_conserved_q<true, true, double>), dim3(grid_for(g.n, kBlock)), dim3(kBlock), 0, st,
                           g, s);
}
void launch_q_derivatives(const DevGeom& g, const DevState& s, const FlowParams& p, double* temp,
                          hipStream_t st) {
    if (g.n <= 0) return;
    const dim3 grid(grid_for((long long)g.n * kQdGroup, kBlock));
    if (p.mixed) {
        if (p.power == 0.0)
            hipLaunchKernelGGL((k_qderiv_timestep<true, false, true, float>), grid, dim3(kBlock), 0, st, g, s, p);
        else
            hipLaunchKernelGGL((k_qderiv_timestep<true, false, false, float>), grid, dim3(kBlock), 0, st, g, s, p);
        return;  // second-order inner iterations are FP64 only
    }
    if (p.power == 0.0)
        hipLaunchKernelGGL((k_qderiv_timestep<true, false, true, double>), grid, dim3(kBlock), 0, st, g, s, p);
    else
        hipLaunchKernelGGL((k_qderiv_timestep<true, false, false, double>), grid, dim3(kBlock), 0, st, g, s, p);
    for (int it = 0; it < p.inner_iterations; ++it) {
        if (p.power == 0.0)
            hipLaunchKernelGGL(k_qderiv_inner<true>, dim3(grid_for(g.n, kBlock)), dim3(kBlock), 0, st, g, s, p,
                               temp);
        else
            hipLaunchKernelGGL(k_qderiv_inner<false>, dim3(grid_for(g.n, kBlock)), dim3(kBlock), 0, st, g, s, p,
                               temp);
        hipLaunchKernelGGL(k_copy, dim3(grid_for(15LL * g.n, kBlock)), dim3(kBlock), 0, st, s.dq,
                           (const double*)temp, 15LL * g.n);
    }
}
void launch_timestep(const DevGeom& g, const DevState& s, const FlowParams& p, hipStream_t st) {
    if (g.n <= 0) return;
    hipLaunchKernelGGL(k_timestep_den, dim3(grid_for(g.n, kBlock)), dim3(kBlock), 0, st, g, s);
    // no least-squares weights in the time step
    hipLaunchKernelGGL((k_qderiv_timestep<false, true, true, double>), dim3(grid_for((long long)g.n * kQdGroup, kBlock)),
                       dim3(kBlock), 0, st, g, s, p);
}
template <typename QT>
void launch_qdt(const DevGeom& g, const DevState& s, const FlowParams& p, hipStream_t st, const KernelShape& ks) {
    int block = ks.block > 0 && ks.block <= kBlock ? ks.block : kBlock;
    block -= block % kQdGroup;
    const dim3 grid(grid_for((long long)g.n * kQdGroup, block)), blk(block);
    if (p.power != 0.0) {
        hipLaunchKernelGGL((k_qderiv_timestep<true, true, false, QT>), grid, blk, ks.lds, st, g, s, p);
        return;
    }
    switch (ks.waves) {
        case 2: hipLaunchKernelGGL((k_qderiv_timestep<true, true, true, QT, 2>), grid, blk, ks.lds, st, g, s, p); break;
        case 6: hipLaunchKernelGGL((k_qderiv_timestep<true, true, true, QT, 6>), grid, blk, ks.lds, st, g, s, p); break;
        default: hipLaunchKernelGGL((k_qderiv_timestep<true, true, true, QT, 4>), grid, blk, ks.lds, st, g, s, p); break;
    }
}
void launch_qderiv_timestep(const DevGeom& g, const DevState& s, const FlowParams& p, hipStream_t st,
                            const KernelShape& ks) {
    if (g.n <= 0) return;
    if (p.mixed)
        launch_qdt<float>(g, s, p, st, ks);
    else
        launch_qdt<double>(g, s, p, st, ks);
}
template <typename T>
void launch_lists(const DevGeom& g, const DevState& s, const FlowParams& p, const int* idx, int count,
                  hipStream_t st) {
    if (count <= 0) return;
    const dim3 grid(grid_for((long long)count * kListGroup, kBlock));
    if (p.power == 0.0)
        hipLaunchKernelGGL((k_flux_residual_lists<true, T>), grid, dim3(kBlock), 0, st, g, s, p, idx, count);
    else
        hipLaunchKernelGGL((k_flux_residual_lists<false, T>), grid, dim3(kBlock), 0, st, g, s, p, idx, count);
}
void launch_flux_residual(const DevGeom& g, const DevState& s, const FlowParams& p, const TypeLists& tl,
                          hipStream_t st, int parts, const LaunchShapes& ls) {
    launch_flux_fused(g, s, p, tl, st, parts, p.mixed != 0, ls);
    if (parts & 4) {
        if (p.mixed)
            launch_lists<float>(g, s, p, tl.generic, tl.n_generic, st);
        else
            launch_lists<double>(g, s, p, tl.generic, tl.n_generic, st);
    }
}
void launch_flux_residual_reference(const DevGeom& g, const DevState& s, const FlowParams& p,
                                    const TypeLists& tl, hipStream_t st) {
    const int* lists[3] = {tl.wall, tl.outer, tl.interior};
    const int counts[3] = {tl.n_wall, tl.n_outer, tl.n_interior};
    for (int c = 0; c < 3; ++c) launch_lists<double>(g, s, p, lists[c], counts[c], st);
}
void launch_explicit_sweep(const DevGeom& g, const DevState& s, hipStream_t st) {
    if (g.n <= 0) return;
    hipLaunchKernelGGL(k_explicit_sweep, dim3(grid_for(g.n, kBlock)), dim3(kBlock), 0, st, g, s);
}
void launch_dinv(const DevGeom& g, const DevState& s, const FlowParams& p, const TypeLists& tl,
                 hipStream_t st) {
    const int* lists[3] = {tl.wall, tl.interior, tl.outer};
    const int counts[3] = {tl.n_wall, tl.n_interior, p.backup ? tl.n_outer : 0};
    for (int c = 0; c < 3; ++c) {
        if (counts[c] <= 0) continue;
        hipLaunchKernelGGL(k_dinv, dim3(grid_for(counts[c], kBlock)), dim3(kBlock), 0, st, g, s, lists[c], counts[c]);
    }
}
void launch_dinv_geometry(const DevGeom& g, double* dgeo, double power, hipStream_t st) {
    if (g.n <= 0) return;
    if (power == 0.0)
        hipLaunchKernelGGL(k_dinv_geometry<true>, dim3(grid_for(g.n, kBlock)), dim3(kBlock), 0, st, g, dgeo, power);
    else
        hipLaunchKernelGGL(k_dinv_geometry<false>, dim3(grid_for(g.n, kBlock)), dim3(kBlock), 0, st, g, dgeo, power);
}
void launch_sweeps_levels(const DevGeom& g, const DevState& s, const FlowParams& p, const SweepSchedule& sch,
                          hipStream_t st) {
    for (int l = 0; l < sch.n_levels_fwd; ++l) {
        const int b = sch.fwd_ptr[l], e = sch.fwd_ptr[l + 1];
        if (p.power == 0.0)
            hipLaunchKernelGGL((k_sweep_level<true, true>), dim3(grid_for(e - b, 64)), dim3(64), 0, st, g, s, p,
                               sch.fwd_points, b, e);
        else
            hipLaunchKernelGGL((k_sweep_level<true, false>), dim3(grid_for(e - b, 64)), dim3(64), 0, st, g, s, p,
                               sch.fwd_points, b, e);
    }
    for (int l = 0; l < sch.n_levels_bwd; ++l) {
        const int b = sch.bwd_ptr[l], e = sch.bwd_ptr[l + 1];
        if (p.power == 0.0)
            hipLaunchKernelGGL((k_sweep_level<false, true>), dim3(grid_for(e - b, 64)), dim3(64), 0, st, g, s, p,
                               sch.bwd_points, b, e);
        else
            hipLaunchKernelGGL((k_sweep_level<false, false>), dim3(grid_for(e - b, 64)), dim3(64), 0, st, g, s, p,
                               sch.bwd_points, b, e);
    }
    // implicit_forward_and_reverse_sweeps.f90:86-88; the alternative driver
    // leaves it to state_update_implicit.f90:98-100
    if (!p.backup)
        hipLaunchKernelGGL(k_copy, dim3(grid_for(5LL * g.n, kBlock)), dim3(kBlock), 0, st, s.delUp,
                           (const double*)s.delUn, 5LL * g.n);
}
int state_update_blocks(int n) { return grid_for(n, kBlock); }
void launch_state_update(const DevGeom& g, const DevState& s, const FlowParams& p, const TypeLists& tl,
                         double* partial, int* partial_pos, int n_partial, ResidualRecord* rec, double* hist,
                         int hist_index, hipStream_t st, const InstanceResidual* ir, bool with_explicit_sweep) {
    // partial buffer layout: [sum | max] doubles, [pos | point] ints
    double* partial_max = partial + n_partial;
    int* partial_pt = partial_pos + n_partial;
    const bool backup = p.backup && !with_explicit_sweep;
    if (with_explicit_sweep)
        hipLaunchKernelGGL((k_state_update<true, false>), dim3(n_partial), dim3(kBlock), 0, st, g, s, p, partial,
                           partial_pos, partial_pt, partial_max, ir ? ir->ressq : nullptr);
    else if (backup)
        hipLaunchKernelGGL((k_state_update<false, true>), dim3(n_partial), dim3(kBlock), 0, st, g, s, p, partial,
                           partial_pos, partial_pt, partial_max, ir ? ir->ressq : nullptr);
    else
        hipLaunchKernelGGL((k_state_update<false, false>), dim3(n_partial), dim3(kBlock), 0, st, g, s, p, partial,
                           partial_pos, partial_pt, partial_max, ir ? ir->ressq : nullptr);
    if (tl.n_sup_in + tl.n_sup_out > 0)
        hipLaunchKernelGGL(k_supersonic, dim3(1), dim3(1), 0, st, g, s, tl.sup_out, tl.n_sup_out, tl.sup_in,
                           tl.n_sup_in, backup ? 1 : 0);
    // state_update_implicit.f90:98-100
    if (backup)
        hipLaunchKernelGGL(k_copy, dim3(grid_for(5LL * g.n, kBlock)), dim3(kBlock), 0, st, s.delUp,
                           (const double*)s.delUn, 5LL * g.n);
    hipLaunchKernelGGL(k_residual_final, dim3(1), dim3(1024), 0, st, (const double*)partial,
                       (const int*)partial_pos, (const int*)partial_pt, (const double*)partial_max, n_partial,
                       g.n, rec, hist, hist_index, ir ? (const int*)ir->counter : (const int*)nullptr);
    if (ir) {
        if (ir->n_chunks > 0)
            hipLaunchKernelGGL(k_chunk_sums, dim3(ir->n_chunks), dim3(kBlock), 0, st, (const double*)ir->ressq,
                               ir->chunk_beg, ir->chunk_end, ir->chunk_sum);
        hipLaunchKernelGGL(k_instance_residual, dim3(ir->n_inst), dim3(kBlock), 0, st, (const double*)ir->chunk_sum,
                           ir->chunk_ptr, ir->norm, ir->n_inst, ir->inst_hist, (const int*)ir->counter);
        hipLaunchKernelGGL(k_bump, dim3(1), dim3(64), 0, st, ir->counter);
    }
}

}  // namespace mfcfd
