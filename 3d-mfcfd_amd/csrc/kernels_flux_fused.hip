// kernels_flux_fused.hip -- flux residual of interior, wall and outer points,
// one pass over each point's walk, in FP64 or in the mixed form.
//
// The reference walks each of a point's direction lists separately and
// recomputes, for every list an edge (i,k) belongs to, the same limited
// reconstruction at both ends (interior_flux_dGxpos.f90:92-104).  An
// interior neighbour belongs to exactly one list per axis, so each edge is
// reconstructed three times.  This kernel walks each point's walk once (conn
// with a membership bit per direction list, then the list entries that are
// not an in-order subsequence of conn), reconstructs each end
// once, evaluates the exponentials and error functions of each axis once
// per end, and forms the directional fluxes of that edge from them.  Every
// flux value is computed with the reference's own expressions, so it is
// bit-identical to the list-by-list kernel; only the least-squares solve is
// applied edge by edge, through the Cramer weights of each list
// (sum_j (G_k - G_i)*(c . (dels, delt, deln)) instead of the 3x3 solve of the
// accumulated sums), which changes the rounding, not the mathematics.
//
// A point is handled by a group of kFluxGroup lanes, each taking every
// kFluxGroup-th neighbour, and the group sums its partial residuals at the
// end.  Per lane this is the same work as one lane per point, but a resident
// wave holds 64 / kFluxGroup points instead of 64: the neighbour data of all
// points in flight on one XCD then fits its 4 MB L2.  With one lane per point
// at the 8 resident waves per CU the FP64 kernel's registers allow, each
// point's q, qm and dq were fetched from HBM about 19 times per launch.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "device_state.h"
#include "kernels.h"
#include "qlskum_math.h"



namespace mfcfd {

namespace {

constexpr int kFluxGroup = 4;

// quantities of one reconstructed end state shared by every flux of the edge
template <typename T>
struct EndState {
    T rho, pr, ut[3], sqrt_beta, sqrt_pi_beta, e5;
};

template <typename T>
__device__ __forceinline__ EndState<T> end_state(const T prim[5], const Frame<T>& f, T pi) {
    EndState<T> e;
    const T u1 = prim[1], u2 = prim[2], u3 = prim[3];
    e.rho = prim[0];
    e.pr = prim[4];
    e.ut[0] = f.t1[0] * u1 + f.t1[1] * u2 + f.t1[2] * u3;
    e.ut[1] = f.t2[0] * u1 + f.t2[1] * u2 + f.t2[2] * u3;
    e.ut[2] = f.n[0] * u1 + f.n[1] * u2 + f.n[2] * u3;
    const T beta = ddiv(T(0.5) * e.rho, e.pr);
    e.sqrt_beta = dsqrt_fast(beta);
    e.sqrt_pi_beta = dsqrt_fast(pi * beta);
    const T u_sqr = e.ut[0] * e.ut[0] + e.ut[1] * e.ut[1] + e.ut[2] * e.ut[2];
    e.e5 = T(2.5) * e.pr + T(0.5) * e.rho * u_sqr;
    return e;
}

// S = u_a*sqrt(beta), B = 0.5*exp(-S*S)/sqrt(pi*beta), erf(S)
template <int A, typename T>
__device__ __forceinline__ void axis_terms(const EndState<T>& e, T& B, T& er) {
    const T S = e.ut[A] * e.sqrt_beta;
    const T ex = dexp(-S * S);
    B = ddiv(T(0.5) * ex, e.sqrt_pi_beta);
    er = derf_e(S, ex);
}

// split_fluxes.f90 flux_G{x,y,z}{p,n}; sg = +1 or -1
template <int A, typename T>
__device__ __forceinline__ void half_flux(const EndState<T>& e, T sg, T B, T er, T G[5]) {
    const T ua = e.ut[A];
    const T Aa = T(0.5) * (T(1) + sg * er);
    const T temp1 = (ua * Aa + sg * B);
    G[0] = e.rho * temp1;
    G[1] = e.rho * e.ut[0] * temp1;
    G[2] = e.rho * e.ut[1] * temp1;
    G[3] = e.rho * e.ut[2] * temp1;
    G[1 + A] = e.pr * Aa + e.rho * ua * temp1;
    const T tail = (e.e5 + T(0.5) * e.pr) * B;
    G[4] = (e.e5 + e.pr) * ua * Aa + sg * tail;
}

// octant_fluxes.f90 flux_Gw{x,y}{p,n} (sn = -1) / flux_Go{x,y}{p,n} (sn = +1)
template <int A, typename T>
__device__ __forceinline__ void octant_flux(const EndState<T>& e, T sg, T sn, T Ba, T era, T B3, T er3,
                                            T G[5]) {
    const T ua = e.ut[A];
    const T Aa = T(0.5) * (T(1) + sg * era);
    const T A3 = T(0.5) * (T(1) + sn * er3);
    const T temp1 = (ua * Aa + sg * Ba);
    G[0] = e.rho * temp1 * A3;
    if (A == 0) {
        G[1] = (e.pr * Aa + e.rho * e.ut[0] * temp1) * A3;
        G[2] = e.rho * e.ut[1] * temp1 * A3;
    } else {
        G[1] = e.rho * e.ut[0] * temp1 * A3;
        G[2] = (e.pr * Aa + e.rho * e.ut[1] * temp1) * A3;
    }
    G[3] = e.rho * temp1 * (e.ut[2] * A3 + sn * B3);
    const T tb = (e.e5 + T(0.5) * e.pr) * Ba;
    const T temp3 = (e.e5 + e.pr) * ua * Aa + sg * tb;
    const T tn = T(0.5) * e.rho * e.ut[2] * B3 * temp1;
    G[4] = temp3 * A3 + sn * tn;
}

// flux differences in T, Cramer-weighted sums in FP64
template <typename T>
__device__ __forceinline__ void accumulate(double acc[5], const T Gi[5], const T Gk[5], double w) {
#pragma unroll
    for (int r = 0; r < 5; ++r) acc[r] = acc[r] + (double)(Gk[r] - Gi[r]) * w;
}

__device__ __forceinline__ double cramer_weight(const double* cf, int d, double wgt, double dels,
                                                double delt, double deln) {
    return wgt * (cf[3 * d] * dels + cf[3 * d + 1] * delt + cf[3 * d + 2] * deln);
}

// Half-range fluxes along axis A for the lists of that axis the entry is in:
// ab bit 0 = list 2A (positive half), bit 1 = list 2A+1.  An interior
// neighbour of conn is in exactly one of the two (generate_connectivity.f90:
// 72-106 splits with <= 0 and > 0); a spilled entry may be in either.
template <int A, typename T>
__device__ __forceinline__ void half_axis(const EndState<T>& ei, const EndState<T>& ek, unsigned ab, T Bi, T eri,
                                          T Bk, T erk, const double* cf, double wgt, double dels, double delt,
                                          double deln, double acc[5]) {
    T Gi[5], Gk[5];
    const bool pos = ab & 1u;
    const T sg = pos ? T(1) : T(-1);
    half_flux<A>(ei, sg, Bi, eri, Gi);
    half_flux<A>(ek, sg, Bk, erk, Gk);
    accumulate(acc, Gi, Gk, cramer_weight(cf, 2 * A + (pos ? 0 : 1), wgt, dels, delt, deln));
    if (ab == 3u) {
        half_flux<A>(ei, T(-1), Bi, eri, Gi);
        half_flux<A>(ek, T(-1), Bk, erk, Gk);
        accumulate(acc, Gi, Gk, cramer_weight(cf, 2 * A + 1, wgt, dels, delt, deln));
    }
}

// interior points on this kernel have every walk entry in exactly one list
// per axis (checked at bind)
template <int A, typename T>
__device__ __forceinline__ void interior_axis(const EndState<T>& ei, const EndState<T>& ek, unsigned bits,
                                              const double* cf, double wgt, double dels, double delt,
                                              double deln, double acc[5]) {
    const bool pos = (bits >> (2 * A)) & 1u;
    const T sg = pos ? T(1) : T(-1);
    T Bi, eri, Bk, erk, Gi[5], Gk[5];
    axis_terms<A>(ei, Bi, eri);
    axis_terms<A>(ek, Bk, erk);
    half_flux<A>(ei, sg, Bi, eri, Gi);
    half_flux<A>(ek, sg, Bk, erk, Gk);
    accumulate(acc, Gi, Gk, cramer_weight(cf, 2 * A + (pos ? 0 : 1), wgt, dels, delt, deln));
}

// x and y half-range lists of an interior point on the general path
template <typename T>
__device__ __forceinline__ void half_axes_xy(const EndState<T>& ei, const EndState<T>& ek, unsigned bits,
                                             const double* cf, double wgt, double dels, double delt, double deln,
                                             double acc[5]) {
    T Bi, eri, Bk, erk;
    const unsigned abx = bits & 3u, aby = (bits >> 2) & 3u;
    if (abx) {
        axis_terms<0>(ei, Bi, eri);
        axis_terms<0>(ek, Bk, erk);
        half_axis<0>(ei, ek, abx, Bi, eri, Bk, erk, cf, wgt, dels, delt, deln, acc);
    }
    if (aby) {
        axis_terms<1>(ei, Bi, eri);
        axis_terms<1>(ek, Bk, erk);
        half_axis<1>(ei, ek, aby, Bi, eri, Bk, erk, cf, wgt, dels, delt, deln, acc);
    }
}

// Quarter-range fluxes along tangent axis A of a wall point (sn = -1, Gw)
// or an outer point (sn = +1, Go).  Their x/y lists split with <= 0 and
// >= 0 (generate_connectivity.f90:148-170, 216-238): a neighbour with a zero
// offset sits in both.
template <int A, typename T>
__device__ __forceinline__ void octant_axis(const EndState<T>& ei, const EndState<T>& ek, unsigned bits, T sn,
                                            T B3i, T er3i, T B3k, T er3k, const double* cf, double wgt,
                                            double dels, double delt, double deln, double acc[5]) {
    const unsigned ab = (bits >> (2 * A)) & 3u;
    if (!ab) return;
    T Bi, eri, Bk, erk, Gi[5], Gk[5];
    axis_terms<A>(ei, Bi, eri);
    axis_terms<A>(ek, Bk, erk);
    const bool pos = ab & 1u;
    const T sg = pos ? T(1) : T(-1);
    octant_flux<A>(ei, sg, sn, Bi, eri, B3i, er3i, Gi);
    octant_flux<A>(ek, sg, sn, Bk, erk, B3k, er3k, Gk);
    accumulate(acc, Gi, Gk, cramer_weight(cf, 2 * A + (pos ? 0 : 1), wgt, dels, delt, deln));
    if (ab == 3u) {
        octant_flux<A>(ei, T(-1), sn, Bi, eri, B3i, er3i, Gi);
        octant_flux<A>(ek, T(-1), sn, Bk, erk, B3k, er3k, Gk);
        accumulate(acc, Gi, Gk, cramer_weight(cf, 2 * A + 1, wgt, dels, delt, deln));
    }
}

// CLASS 0: interior points whose walk is conn with every entry in exactly
// one list per axis (6 half-range lists).  CLASS 1: every other point, told
// apart by status: wall points take the Gw octant lists in x and y, the
// half-range z- list over conn and the factor 2 of flux_residual.f90:50;
// outer points the Go octant lists and the half-range z+ list (:54-66);
// interior points whose lists the 22-column spill altered the six
// half-range lists, each walk entry in the lists its bits name (:68-81).
// T: arithmetic of the reconstruction and fluxes (double: the FP64 form;
// float: the mixed form, reading the FP32 q, qm, dq); geometry, Cramer
// weights and the sums stay FP64.
// WAVES: minimum waves per SIMD the kernel is compiled for (its register
// cap); the launch picks one of the compiled variants and a block size
template <int CLASS, bool UW, typename T, int WAVES>
__global__ void __launch_bounds__(kBlock, WAVES) k_flux_fused(DevGeom g, DevState s, FlowParams fp, const int* idx,
                                                             int count) {
    const unsigned t = (unsigned)xcd_index();
    const int m = (int)(t / kFluxGroup);
    const int lane = (int)(t % kFluxGroup);
    if (m >= count) return;  // whole groups: blocks hold a multiple of kFluxGroup lanes
    const QArrays<T> qa = qarrays<T>(s);
    const int i = idx[m];
    const Frame<double> f = load_frame(g.frame, i);
    const Frame<T> ft = frame_as<T>(f);
    const T pi = T(fp.pi);
    const double4 pi4 = g.pos[i];
    // the point's own reconstruction inputs are re-read for every neighbour:
    // the four lanes of a group read the same addresses, which costs the
    // gather little, and holding them would cost registers and resident
    // waves (FP32: 116 instead of 139 VGPRs, 4 instead of 3 waves per SIMD)
    const T* qip = qa.qp(i);
    const T* dqi = qa.dqp(i);
    const T* qmi = qa.qmp(i);
    const T epsi2_i = T(limiter_epsi2(fp.vl_const, pi4.w));
    const double* cf = g.coef + 18 * (long long)i;
    const int nb = g.nwalk[i];
    const unsigned* wk = g.walk + kWalkCols * (long long)i;
    const int st_i = CLASS == 1 ? g.status[i] : 0;
    const T sn = st_i == 2 ? T(1) : T(-1);
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int j = lane; j < nb; j += kFluxGroup) {
        const unsigned e = wk[j];
        const unsigned bits = e & ((1u << kWalkMaskBits) - 1u);
        if (bits == 0u) continue;
        const int k = (int)(e >> kWalkMaskBits);
        const double4 pk = g.pos[k];
        const double delx = pk.x - pi4.x;
        const double dely = pk.y - pi4.y;
        const double delz = pk.z - pi4.z;
        const double dels = delx * f.t1[0] + dely * f.t1[1] + delz * f.t1[2];
        const double delt = delx * f.t2[0] + dely * f.t2[1] + delz * f.t2[2];
        const double deln = delx * f.n[0] + dely * f.n[1] + delz * f.n[2];
        const double wgt = wls_weight<UW>(dels, delt, deln, fp.power);
        T prim[5], qi[5];
#pragma unroll
        for (int r = 0; r < 5; ++r) qi[r] = qip[r];
        reconstruct_end<T>(qi, dqi, qmi, epsi2_i, T(delx), T(dely), T(delz), prim);
        const EndState<T> ei = end_state(prim, ft, pi);
        T qk[5];
#pragma unroll
        for (int r = 0; r < 5; ++r) qk[r] = qa.qp(k)[r];
        reconstruct_end<T>(qk, qa.dqp(k), qa.qmp(k), T(limiter_epsi2(fp.vl_const, pk.w)), T(delx), T(dely), T(delz),
                           prim);
        const EndState<T> ek = end_state(prim, ft, pi);
        if (CLASS == 0) {
            interior_axis<0>(ei, ek, bits, cf, wgt, dels, delt, deln, acc);
            interior_axis<1>(ei, ek, bits, cf, wgt, dels, delt, deln, acc);
            interior_axis<2>(ei, ek, bits, cf, wgt, dels, delt, deln, acc);
        } else {
            T B3i, er3i, B3k, er3k;
            axis_terms<2>(ei, B3i, er3i);
            axis_terms<2>(ek, B3k, er3k);
            if (st_i == 0) {
                half_axes_xy(ei, ek, bits, cf, wgt, dels, delt, deln, acc);
            } else {
                octant_axis<0>(ei, ek, bits, sn, B3i, er3i, B3k, er3k, cf, wgt, dels, delt, deln, acc);
                octant_axis<1>(ei, ek, bits, sn, B3i, er3i, B3k, er3k, cf, wgt, dels, delt, deln, acc);
            }
            const unsigned abz = (bits >> (2 * 2)) & 3u;
            if (abz) half_axis<2>(ei, ek, abz, B3i, er3i, B3k, er3k, cf, wgt, dels, delt, deln, acc);
        }
    }
#pragma unroll
    for (int r = 0; r < 5; ++r) acc[r] = group_sum<kFluxGroup>(acc[r]);
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < 5; ++r) s.flux_res[5 * (long long)i + r] = st_i == 1 ? 2.0 * acc[r] : acc[r];
    }
}

template <int CLASS, typename T>
void launch_class(const DevGeom& g, const DevState& s, const FlowParams& p, const int* idx, int count,
                  const KernelShape& ks, hipStream_t st) {
    if (count <= 0) return;
    int block = ks.block > 0 && ks.block <= kBlock ? ks.block : kBlock;
    block -= block % kFluxGroup;
    const dim3 grid(grid_for((long long)count * kFluxGroup, block)), blk(block);
    if (p.power != 0.0) {
        hipLaunchKernelGGL((k_flux_fused<CLASS, false, T, 2>), grid, blk, ks.lds, st, g, s, p, idx, count);
        return;
    }
    switch (ks.waves) {
        case 1: hipLaunchKernelGGL((k_flux_fused<CLASS, true, T, 1>), grid, blk, ks.lds, st, g, s, p, idx, count); break;
        case 3: hipLaunchKernelGGL((k_flux_fused<CLASS, true, T, 3>), grid, blk, ks.lds, st, g, s, p, idx, count); break;
        case 4: hipLaunchKernelGGL((k_flux_fused<CLASS, true, T, 4>), grid, blk, ks.lds, st, g, s, p, idx, count); break;
        case 5:
            if constexpr (sizeof(T) == 4) {
                hipLaunchKernelGGL((k_flux_fused<CLASS, true, T, 5>), grid, blk, ks.lds, st, g, s, p, idx, count);
                break;
            }
            [[fallthrough]];
        case 6:
            if constexpr (sizeof(T) == 4) {
                hipLaunchKernelGGL((k_flux_fused<CLASS, true, T, 6>), grid, blk, ks.lds, st, g, s, p, idx, count);
                break;
            }
            [[fallthrough]];
        default: hipLaunchKernelGGL((k_flux_fused<CLASS, true, T, 2>), grid, blk, ks.lds, st, g, s, p, idx, count); break;
    }
}

}  // namespace

void launch_flux_fused(const DevGeom& g, const DevState& s, const FlowParams& p, const TypeLists& tl,
                       hipStream_t st, int parts, bool mixed, const LaunchShapes& ls) {
    if (parts & 1) {
        if (mixed)
            launch_class<0, float>(g, s, p, tl.fused_interior, tl.n_fused_interior, ls.flux_interior, st);
        else
            launch_class<0, double>(g, s, p, tl.fused_interior, tl.n_fused_interior, ls.flux_interior, st);
    }
    if (parts & 2) {
        if (mixed)
            launch_class<1, float>(g, s, p, tl.fused_wall, tl.n_fused_wall, ls.flux_wall, st);
        else
            launch_class<1, double>(g, s, p, tl.fused_wall, tl.n_fused_wall, ls.flux_wall, st);
    }
}

}  // namespace mfcfd
