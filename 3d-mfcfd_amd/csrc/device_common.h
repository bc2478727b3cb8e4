// device_common.h -- helpers shared by the FP64 kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "device_state.h"
#include "qlskum_math.h"

namespace mfcfd {

constexpr int kBlock = 256;

inline int grid_for(long long n, int block) { return (int)((n + block - 1) / block); }

// MI355X deals workgroup b of a launch to XCD b % 8, and each XCD has its own
// L2.  Gather kernels renumber their blocks so that every XCD walks one
// contiguous eighth of the (Morton-ordered) points: a point's neighbours are
// then mostly read through the same L2.  The map is a bijection of
// [0, gridDim.x) for any grid size.
constexpr int kXcds = 8;

__device__ __forceinline__ int xcd_block() {
    const int nb = (int)gridDim.x;
    const int b = (int)blockIdx.x;
    const int x = b % kXcds, r = b / kXcds;
    const int q = nb / kXcds, rem = nb % kXcds;
    return x * q + (x < rem ? x : rem) + r;
}

// index of this thread when the launch covers count items, blocks renumbered per XCD
__device__ __forceinline__ int xcd_index() { return xcd_block() * (int)blockDim.x + (int)threadIdx.x; }

// The value of v in the lane whose index differs in bit O.  Within aligned
// groups of four lanes (O = 1, 2) this is a DPP quad permutation, one VALU
// move per 32-bit word; wider partners go through the LDS crossbar
// (__shfl_xor).  Every lane of each aligned group of four must be active.
template <int O>
__device__ __forceinline__ int lane_xor_b32(int v) {
    static_assert(O == 1 || O == 2, "quad permutations only");
    // quad_perm [1,0,3,2] and [2,3,0,1]
    return __builtin_amdgcn_update_dpp(v, v, O == 1 ? 0xB1 : 0x4E, 0xf, 0xf, false);
}
template <int O, int G, typename V>
__device__ __forceinline__ V xor_partner(V v) {
    if constexpr (O > 2) {
        return __shfl_xor(v, O, G);
    } else if constexpr (sizeof(V) == 8) {
        const long long b = __builtin_bit_cast(long long, v);
        const int lo = lane_xor_b32<O>((int)b), hi = lane_xor_b32<O>((int)(b >> 32));
        return __builtin_bit_cast(V, ((long long)hi << 32) | (long long)(unsigned)lo);
    } else {
        return __builtin_bit_cast(V, lane_xor_b32<O>(__builtin_bit_cast(int, v)));
    }
}

// Sum of v over the G consecutive lanes of an aligned lane group (G a power
// of two).  Every lane of the group forms the same pairwise tree, so every
// lane receives the same, order-determined value.
template <int G, int O = 1, typename V>
__device__ __forceinline__ V group_sum(V v) {
    if constexpr (O < G) {
        v += xor_partner<O, G>(v);
        return group_sum<G, 2 * O>(v);
    } else {
        return v;
    }
}

// group maximum / minimum with the comparisons of a sequential scan
// (a NaN operand never replaces a value)
template <int G, int O = 1, typename V>
__device__ __forceinline__ V group_max(V v) {
    if constexpr (O < G) {
        const V w = xor_partner<O, G>(v);
        if (w > v) v = w;
        return group_max<G, 2 * O>(v);
    } else {
        return v;
    }
}
template <int G, int O = 1, typename V>
__device__ __forceinline__ V group_min(V v) {
    if constexpr (O < G) {
        const V w = xor_partner<O, G>(v);
        if (w < v) v = w;
        return group_min<G, 2 * O>(v);
    } else {
        return v;
    }
}

template <typename T>
__device__ __forceinline__ Frame<T> frame_as(const Frame<double>& f) {
    Frame<T> o;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        o.t1[c] = T(f.t1[c]);
        o.t2[c] = T(f.t2[c]);
        o.n[c] = T(f.n[c]);
    }
    return o;
}

__device__ __forceinline__ Frame<double> load_frame(const double* frame, int k) {
    Frame<double> f;
    const double* p = frame + 9 * (long long)k;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        f.t1[c] = p[c];
        f.t2[c] = p[3 + c];
        f.n[c] = p[6 + c];
    }
    return f;
}

// dist**power as the reference build evaluates it: power is a compile-time
// PARAMETER (parameter.f90:21), and the constant exponents 0, 0.5, 1, 2 and
// 3 reduce to 1, sqrt(d), d, d*d and (d*d)*d (flang's x86 code of the
// reference); any other exponent is a pow call.  pow(d, 2) need not equal
// d*d rounded once, so the reduction is kept bit for bit.
__device__ __forceinline__ double ref_pow(double d, double power) {
    if (power == 0.0) return 1.0;
    if (power == 0.5) return __builtin_sqrt(d);
    if (power == 1.0) return d;
    if (power == 2.0) return d * d;
    if (power == 3.0) return d * d * d;
    return pow(d, power);
}

// least-squares weight 1/dist**power; UNIT_W selects the power = 0 kernels,
// where the weight is exactly 1 and no power function is compiled in
template <bool UNIT_W>
__device__ __forceinline__ double wls_weight(double a, double b, double c, double power) {
    if (UNIT_W) return 1.0;
    return 1.0 / ref_pow(__builtin_sqrt(a * a + b * b + c * c), power);
}

// Cramer weights of one direction list of point i: the geometric sums of the
// list in list order (interior_flux_dGxpos.f90:84-90) and the cofactors of
// column col (the list's axis) divided by the determinant, so that the
// list's derivative is sum_j (G_k - G_i) * (c0*dels_j + c1*delt_j + c2*deln_j)
// times the weight of entry j.  An empty list gives zeros.
template <bool UW>
__device__ __forceinline__ void list_cofactors(const DevGeom& g, int i, const Frame<double>& f, const int* list,
                                               int cnt, int col, double power, double c[3]) {
    const double4 pi4 = g.pos[i];
    Geo6<double> geo = {0, 0, 0, 0, 0, 0};
    for (int j = 0; j < cnt; ++j) {
        const double4 pk = g.pos[list[j]];
        const double delx = pk.x - pi4.x, dely = pk.y - pi4.y, delz = pk.z - pi4.z;
        const double dels = delx * f.t1[0] + dely * f.t1[1] + delz * f.t1[2];
        const double delt = delx * f.t2[0] + dely * f.t2[1] + delz * f.t2[2];
        const double deln = delx * f.n[0] + dely * f.n[1] + delz * f.n[2];
        const double w = wls_weight<UW>(dels, delt, deln, power);
        const double dels_w = dels * w, delt_w = delt * w, deln_w = deln * w;
        geo.xx = geo.xx + dels * dels_w;
        geo.yy = geo.yy + delt * delt_w;
        geo.zz = geo.zz + deln * deln_w;
        geo.xy = geo.xy + dels * delt_w;
        geo.yz = geo.yz + delt * deln_w;
        geo.zx = geo.zx + deln * dels_w;
    }
    if (cnt == 0) {
        c[0] = c[1] = c[2] = 0.0;
        return;
    }
    const double det = cramer_det(geo);
    c[0] = cramer_num(geo, col, 1.0, 0.0, 0.0) / det;
    c[1] = cramer_num(geo, col, 0.0, 1.0, 0.0) / det;
    c[2] = cramer_num(geo, col, 0.0, 0.0, 1.0) / det;
}

// Limited reconstruction of q at one end of an edge and conversion to
// primitive variables (interior_flux_dGxpos.f90:92-96,99-103):
//   qtilde = q - 0.5*phi*(delx*dq1 + dely*dq2 + delz*dq3)
template <typename T>
__device__ __forceinline__ void reconstruct_end(const T q[5], const T* dq, const T* qm, T epsi2, T delx,
                                                T dely, T delz, T prim[5]) {
    T temp[5], qt[5];
#pragma unroll
    for (int r = 0; r < 5; ++r) {
        temp[r] = delx * dq[3 * r] + dely * dq[3 * r + 1] + delz * dq[3 * r + 2];
        qt[r] = q[r] - T(0.5) * temp[r];
    }
#pragma unroll
    for (int r = 0; r < 5; ++r) {
        // FP64: the reference's qtilde - q.  FP32: the same quantity as
        // -0.5*temp; qtilde - q cancels catastrophically against the
        // limiter's 1e-5 switch in single precision
        const T del_neg = sizeof(T) == 8 ? qt[r] - q[r] : T(-0.5) * temp[r];
        const T phi = venkat_phi_dn(del_neg, q[r], qm[2 * r], qm[2 * r + 1], epsi2);
        qt[r] = q[r] - T(0.5) * phi * temp[r];
    }
    qtilde_to_primitive(qt, prim);
}

// epsi**2 of the limiter, epsi = (VL_CONST*min_dist)**3 (limiters.f90:34-35)
__device__ __forceinline__ double limiter_epsi2(double vl_const, double min_dist) {
    double e = vl_const * min_dist;
    e = e * e * e;
    return e * e;
}

// Flux family and final division of direction d for a point class
// (flux_residual.f90:40-81): interior half-range fluxes; wall quarter-range
// Gw for x/y and half-range Gzn over conn for z; outer quarter-range Go for
// x/y and half-range Gzp, both with temp*(1/det).
__device__ __forceinline__ int dir_family(int cls, int d) {
    if (cls == 1) return d <= YNEG ? 1 : 0;
    if (cls == 2) return d <= YNEG ? 2 : 0;
    return 0;
}

}  // namespace mfcfd
