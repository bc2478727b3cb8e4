// kernels_flux_fused.hip -- flux residual of regular interior and wall points,
// one pass over the neighbour list.
//
// The reference walks each of a point's direction lists separately and
// recomputes, for every list an edge (i,k) belongs to, the same limited
// reconstruction at both ends (interior_flux_dGxpos.f90:92-104).  An
// interior neighbour belongs to exactly one list per axis, so each edge is
// reconstructed three times.  For a regular point (every list an in-order
// subsequence of conn) this kernel walks conn once, reconstructs each end
// once, evaluates the exponentials and error functions of each axis once
// per end, and forms the directional fluxes of that edge from them.  Every
// flux value is computed with the reference's own expressions, so it is
// bit-identical to the list-by-list kernel; only the least-squares solve is
// applied edge by edge, through the Cramer weights of each list
// (sum_j (G_k - G_i)*(c . (dels, delt, deln)) instead of the 3x3 solve of the
// accumulated sums), which changes the rounding, not the mathematics.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "device_state.h"
#include "kernels.h"
#include "qlskum_math.h"

// minimum waves per SIMD the fused kernels are compiled for (register cap)
#ifndef MFCFD_FUSED_WAVES
#define MFCFD_FUSED_WAVES 2
#endif


namespace mfcfd {

namespace {

// quantities of one reconstructed end state shared by every flux of the edge
struct EndState {
    double rho, pr, ut[3], sqrt_beta, sqrt_pi_beta, e5;
};

__device__ __forceinline__ EndState end_state(const double prim[5], const Frame<double>& f, double pi) {
    EndState e;
    const double u1 = prim[1], u2 = prim[2], u3 = prim[3];
    e.rho = prim[0];
    e.pr = prim[4];
    e.ut[0] = f.t1[0] * u1 + f.t1[1] * u2 + f.t1[2] * u3;
    e.ut[1] = f.t2[0] * u1 + f.t2[1] * u2 + f.t2[2] * u3;
    e.ut[2] = f.n[0] * u1 + f.n[1] * u2 + f.n[2] * u3;
    const double beta = 0.5 * e.rho / e.pr;
    e.sqrt_beta = __builtin_sqrt(beta);
    e.sqrt_pi_beta = __builtin_sqrt(pi * beta);
    const double u_sqr = e.ut[0] * e.ut[0] + e.ut[1] * e.ut[1] + e.ut[2] * e.ut[2];
    e.e5 = 2.5 * e.pr + 0.5 * e.rho * u_sqr;
    return e;
}

// S = u_a*sqrt(beta), B = 0.5*exp(-S*S)/sqrt(pi*beta), erf(S)
template <int A>
__device__ __forceinline__ void axis_terms(const EndState& e, double& B, double& er) {
    const double S = e.ut[A] * e.sqrt_beta;
    B = 0.5 * exp(-S * S) / e.sqrt_pi_beta;
    er = erf(S);
}

// split_fluxes.f90 flux_G{x,y,z}{p,n}; sg = +1.0 or -1.0
template <int A>
__device__ __forceinline__ void half_flux(const EndState& e, double sg, double B, double er, double G[5]) {
    const double ua = e.ut[A];
    const double Aa = 0.5 * (1.0 + sg * er);
    const double temp1 = (ua * Aa + sg * B);
    G[0] = e.rho * temp1;
    G[1] = e.rho * e.ut[0] * temp1;
    G[2] = e.rho * e.ut[1] * temp1;
    G[3] = e.rho * e.ut[2] * temp1;
    G[1 + A] = e.pr * Aa + e.rho * ua * temp1;
    const double tail = (e.e5 + 0.5 * e.pr) * B;
    G[4] = (e.e5 + e.pr) * ua * Aa + sg * tail;
}

// octant_fluxes.f90 flux_Gw{x,y}{p,n} (sn = -1) / flux_Go{x,y}{p,n} (sn = +1)
template <int A>
__device__ __forceinline__ void octant_flux(const EndState& e, double sg, double sn, double Ba, double era,
                                            double B3, double er3, double G[5]) {
    const double ua = e.ut[A];
    const double Aa = 0.5 * (1.0 + sg * era);
    const double A3 = 0.5 * (1.0 + sn * er3);
    const double temp1 = (ua * Aa + sg * Ba);
    G[0] = e.rho * temp1 * A3;
    if (A == 0) {
        G[1] = (e.pr * Aa + e.rho * e.ut[0] * temp1) * A3;
        G[2] = e.rho * e.ut[1] * temp1 * A3;
    } else {
        G[1] = e.rho * e.ut[0] * temp1 * A3;
        G[2] = (e.pr * Aa + e.rho * e.ut[1] * temp1) * A3;
    }
    G[3] = e.rho * temp1 * (e.ut[2] * A3 + sn * B3);
    const double tb = (e.e5 + 0.5 * e.pr) * Ba;
    const double temp3 = (e.e5 + e.pr) * ua * Aa + sg * tb;
    const double tn = 0.5 * e.rho * e.ut[2] * B3 * temp1;
    G[4] = temp3 * A3 + sn * tn;
}

__device__ __forceinline__ void accumulate(double acc[5], const double Gi[5], const double Gk[5], double w) {
#pragma unroll
    for (int r = 0; r < 5; ++r) acc[r] = acc[r] + (Gk[r] - Gi[r]) * w;
}

__device__ __forceinline__ double cramer_weight(const double* cf, int d, double wgt, double dels,
                                                double delt, double deln) {
    return wgt * (cf[3 * d] * dels + cf[3 * d + 1] * delt + cf[3 * d + 2] * deln);
}

template <int A>
__device__ __forceinline__ void interior_axis(const EndState& ei, const EndState& ek, unsigned bits,
                                              const double* cf, double wgt, double dels, double delt,
                                              double deln, double acc[5]) {
    // exactly one of the two lists of this axis holds the neighbour
    // (generate_connectivity.f90:72-106: <= 0 and > 0)
    const bool pos = (bits >> (2 * A)) & 1u;
    const double sg = pos ? 1.0 : -1.0;
    const int d = 2 * A + (pos ? 0 : 1);
    double Bi, eri, Bk, erk, Gi[5], Gk[5];
    axis_terms<A>(ei, Bi, eri);
    axis_terms<A>(ek, Bk, erk);
    half_flux<A>(ei, sg, Bi, eri, Gi);
    half_flux<A>(ek, sg, Bk, erk, Gk);
    accumulate(acc, Gi, Gk, cramer_weight(cf, d, wgt, dels, delt, deln));
}

template <int A>
__device__ __forceinline__ void wall_axis(const EndState& ei, const EndState& ek, unsigned bits,
                                          double B3i, double er3i, double B3k, double er3k,
                                          const double* cf, double wgt, double dels, double delt,
                                          double deln, double acc[5]) {
    // wall lists split with <= 0 and >= 0 (generate_connectivity.f90:148-170):
    // a neighbour with a zero offset sits in both
    const unsigned ab = (bits >> (2 * A)) & 3u;
    if (!ab) return;
    double Bi, eri, Bk, erk, Gi[5], Gk[5];
    axis_terms<A>(ei, Bi, eri);
    axis_terms<A>(ek, Bk, erk);
    const bool pos = ab & 1u;
    const double sg = pos ? 1.0 : -1.0;
    octant_flux<A>(ei, sg, -1.0, Bi, eri, B3i, er3i, Gi);
    octant_flux<A>(ek, sg, -1.0, Bk, erk, B3k, er3k, Gk);
    accumulate(acc, Gi, Gk, cramer_weight(cf, 2 * A + (pos ? 0 : 1), wgt, dels, delt, deln));
    if (ab == 3u) {
        octant_flux<A>(ei, -1.0, -1.0, Bi, eri, B3i, er3i, Gi);
        octant_flux<A>(ek, -1.0, -1.0, Bk, erk, B3k, er3k, Gk);
        accumulate(acc, Gi, Gk, cramer_weight(cf, 2 * A + 1, wgt, dels, delt, deln));
    }
}

// CLASS 0: interior (6 half-range lists); CLASS 1: wall (Gw octant lists in
// x and y, half-range Gzn over conn, factor 2 of flux_residual.f90:50)
template <int CLASS, bool UW>
__global__ void __launch_bounds__(kBlock, MFCFD_FUSED_WAVES) k_flux_fused(DevGeom g, DevState s, FlowParams fp, const int* idx,
                                                       int count) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= count) return;
    const int i = idx[m];
    const Frame<double> f = load_frame(g.frame, i);
    const double4 pi4 = g.pos[i];
    const double* qip = s.q + 5 * (long long)i;
    const double* dqi = s.dq + 15 * (long long)i;
    const double* qmi = s.qm + 10 * (long long)i;
    const double epsi2_i = limiter_epsi2(fp.vl_const, pi4.w);
    const double* cf = g.coef + 18 * (long long)i;
    const int nb = g.nbhs[i];
    const int* cn = g.conn + kConnCols * (long long)i;
    const unsigned char* mk = g.mask + kConnCols * (long long)i;
    double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int j = 0; j < nb; ++j) {
        const unsigned bits = mk[j];
        if (bits == 0u) continue;
        const int k = cn[j];
        const double4 pk = g.pos[k];
        const double delx = pk.x - pi4.x;
        const double dely = pk.y - pi4.y;
        const double delz = pk.z - pi4.z;
        const double dels = delx * f.t1[0] + dely * f.t1[1] + delz * f.t1[2];
        const double delt = delx * f.t2[0] + dely * f.t2[1] + delz * f.t2[2];
        const double deln = delx * f.n[0] + dely * f.n[1] + delz * f.n[2];
        const double wgt = wls_weight<UW>(dels, delt, deln, fp.power);
        double prim[5], qi[5], qk[5];
#pragma unroll
        for (int r = 0; r < 5; ++r) qi[r] = qip[r];
        reconstruct_end(qi, dqi, qmi, epsi2_i, delx, dely, delz, prim);
        const EndState ei = end_state(prim, f, fp.pi);
#pragma unroll
        for (int r = 0; r < 5; ++r) qk[r] = s.q[5 * (long long)k + r];
        reconstruct_end(qk, s.dq + 15 * (long long)k, s.qm + 10 * (long long)k,
                        limiter_epsi2(fp.vl_const, pk.w), delx, dely, delz, prim);
        const EndState ek = end_state(prim, f, fp.pi);
        if (CLASS == 0) {
            interior_axis<0>(ei, ek, bits, cf, wgt, dels, delt, deln, acc);
            interior_axis<1>(ei, ek, bits, cf, wgt, dels, delt, deln, acc);
            interior_axis<2>(ei, ek, bits, cf, wgt, dels, delt, deln, acc);
        } else {
            double B3i, er3i, B3k, er3k;
            axis_terms<2>(ei, B3i, er3i);
            axis_terms<2>(ek, B3k, er3k);
            wall_axis<0>(ei, ek, bits, B3i, er3i, B3k, er3k, cf, wgt, dels, delt, deln, acc);
            wall_axis<1>(ei, ek, bits, B3i, er3i, B3k, er3k, cf, wgt, dels, delt, deln, acc);
            if (bits & (1u << ZNEG)) {
                double Gi[5], Gk[5];
                half_flux<2>(ei, -1.0, B3i, er3i, Gi);
                half_flux<2>(ek, -1.0, B3k, er3k, Gk);
                accumulate(acc, Gi, Gk, cramer_weight(cf, ZNEG, wgt, dels, delt, deln));
            }
        }
    }
#pragma unroll
    for (int r = 0; r < 5; ++r) s.flux_res[5 * (long long)i + r] = CLASS == 1 ? 2.0 * acc[r] : acc[r];
}

}  // namespace

void launch_flux_fused(const DevGeom& g, const DevState& s, const FlowParams& p, const TypeLists& tl,
                       hipStream_t st, int parts) {
    const bool unit = p.power == 0.0;
    if ((parts & 1) && tl.n_fused_interior > 0) {
        const dim3 grid(grid_for(tl.n_fused_interior, kBlock));
        if (unit)
            hipLaunchKernelGGL((k_flux_fused<0, true>), grid, dim3(kBlock), 0, st, g, s, p, tl.fused_interior,
                               tl.n_fused_interior);
        else
            hipLaunchKernelGGL((k_flux_fused<0, false>), grid, dim3(kBlock), 0, st, g, s, p, tl.fused_interior,
                               tl.n_fused_interior);
    }
    if ((parts & 2) && tl.n_fused_wall > 0) {
        const dim3 grid(grid_for(tl.n_fused_wall, kBlock));
        if (unit)
            hipLaunchKernelGGL((k_flux_fused<1, true>), grid, dim3(kBlock), 0, st, g, s, p, tl.fused_wall,
                               tl.n_fused_wall);
        else
            hipLaunchKernelGGL((k_flux_fused<1, false>), grid, dim3(kBlock), 0, st, g, s, p, tl.fused_wall,
                               tl.n_fused_wall);
    }
}

}  // namespace mfcfd
