// device_state.h -- the HBM-resident point cloud and flow state.
//
// Points are stored in device order: a Morton (Z-order) permutation of the
// host's point numbering, so the 25 neighbours of a point usually share
// cache lines with the neighbours of the points next to it.  The host's
// numbering survives only at the boundary (upload/download permute, the
// residual reports host point numbers).
//
// Per point k (device order, 0-based):
//   pos[k]            double4 (x, y, z, min_dist): one 32-byte gather per neighbour
//   frame[9k..9k+8]   tan1(3), tan2(3), nor(3)
//   conn[27k + j]     neighbours in the host's conn order (data_structure.f90:111)
//   dir_cnt[d*n + k], dir_off[d*n + k]
//                     direction d = xpos, xneg, ypos, yneg, zpos, zneg: the list
//                     each reference stencil of point k walks, resolved once at
//                     bind time from the host's (22,N) arrays exactly as the
//                     reference reads them (a list longer than 22 continues into
//                     the next point's column).  Wall points store their full
//                     conn as zneg (wall_flux_dGzneg.f90:61-65), outer points
//                     their zpos list; lst[dir_off .. dir_off + dir_cnt) holds it
//   conn[27k + j]     (as above) and walk[32k + j], nwalk[k]: the entries the
//                     fused flux kernel visits -- conn, then every list entry
//                     that does not continue an in-order match against it --
//                     each packing the neighbour's number above six bits,
//                     bit d set when the entry belongs to list d
//   coef[18k + 3d ..] Cramer weights of list d: the derivative of that stencil is
//                     sum_j (G_k - G_i) * (c0*dels_j + c1*delt_j + c2*deln_j)
//   prim/U/q/flux_res/delUp/delUn [5k + r], qm [(5k + r)*2 + s],
//   dq [(5k + r)*3 + c], delt [k]  -- the reference's (5,N), (2,5,N), (3,5,N)
//                     layouts, so transfers are plain permuted copies.
#pragma once

#include <hip/hip_runtime.h>

namespace mfcfd {

constexpr int kConnCols = 27;
// entries per point of the fused flux kernel's walk: conn, then the list
// entries that are not an in-order subsequence of conn (22-column spill)
constexpr int kWalkCols = 64;
// a walk entry packs the neighbour's device number above kWalkMaskBits bits
// of direction-list membership (bit d: the entry is in list d)
constexpr int kWalkMaskBits = 6;
constexpr int kMaxPoints = 1 << (32 - kWalkMaskBits);
constexpr int kSplitCols = 22;
constexpr int kSplitSlack = 32;

enum Dir { XPOS = 0, XNEG = 1, YPOS = 2, YNEG = 3, ZPOS = 4, ZNEG = 5 };

struct DevGeom {
    int n;
    const double4* pos;
    const double* frame;
    const int* status;
    const int* nbhs;
    const int* conn;
    const int* dir_cnt;          // [6][n]
    const long long* dir_off;    // [6][n]
    const int* lst;
    const unsigned* walk;        // [kWalkCols n] entries the fused flux kernel visits, packed
    const int* nwalk;            // [n] used entries of walk
    const double* coef;          // [18n]
    const double* dgeo;          // [12n] per direction list: D_inv's Cramer numerator and determinant
    const int* alias;
    const int* respos;           // position in the residual order (wall, then interior)
    const int* perm;             // device point -> host point (0-based)
    // batches of clouds: instance of each point and each instance's
    // freestream velocity (u1_inf, u2_inf); null for a single cloud
    const int* inst;
    const double* inst_fs;
};

struct DevState {
    double* prim;
    double* U;
    double* q;
    double* qm;
    double* dq;
    double* delt;
    double* tsden;   // [n] |u| + 3 sqrt(p/rho) of each point, the local time step's denominator
    double* flux_res;
    double* delUp;
    double* delUn;
    double* dinv;
    // mixed precision: q, qm, dq held in FP32 (the reference layouts, except
    // that q32 holds kQ32Stride floats per point: q, then the time step's
    // denominator tsden in FP32, which the q-derivative kernel then reads
    // with the same two loads as q)
    float* q32;
    float* qm32;
    float* dq32;
};

constexpr int kQ32Stride = 6;

// q-variable arrays of one precision: point k's q (5), qm (2,5) and dq (3,5)
// start at q + sq*k, qm + sqm*k, dq + sdq*k
template <typename T>
struct QArrays {
    T* q;
    T* qm;
    T* dq;
    int sq, sqm, sdq;
    __host__ __device__ T* qp(long long k) const { return q + sq * k; }
    __host__ __device__ T* qmp(long long k) const { return qm + sqm * k; }
    __host__ __device__ T* dqp(long long k) const { return dq + sdq * k; }
};

template <typename T>
__host__ __device__ inline QArrays<T> qarrays(const DevState& s);
template <>
__host__ __device__ inline QArrays<double> qarrays<double>(const DevState& s) {
    return {s.q, s.qm, s.dq, 5, 10, 15};
}
template <>
__host__ __device__ inline QArrays<float> qarrays<float>(const DevState& s) {
    return {s.q32, s.qm32, s.dq32, kQ32Stride, 10, 15};
}

struct FlowParams {
    double cfl, vl_const, power, pi;
    double rho_inf, u1_inf, u2_inf, u3_inf, pr_inf;
    int inner_iterations;
    int mixed;   // 1: q, qm, dq and the split fluxes in FP32
    int backup;  // 1: the alternative implicit pairing (implicit_forward_and_reverse_sweeps_backup.f90
                 // with state_update_implicit.f90)
};

// device-side residual record, state_update.f90:115 and data_structure.f90:58-59
struct ResidualRecord {
    double res_new;
    double sum_res_sqr;
    double max_res;
    int max_res_point;   // host numbering, 1-based (0: no positive residual)
    int stalls;          // sync-free sweep waits that hit their spin limit
};

}  // namespace mfcfd
