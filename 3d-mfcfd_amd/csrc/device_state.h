// device_state.h -- the HBM-resident point cloud and flow state.
//
// Layout (per point k, 0-based):
//   pos[k]            double4 (x, y, z, min_dist): one 32-byte gather per neighbour
//   frame[9k..9k+8]   tan1(3), tan2(3), nor(3)
//   conn[27k + j]     0-based neighbour numbers (data_structure.f90:111 order)
//   split[d][22k + j] 0-based split lists, addressed exactly like the
//                     reference's column-major (22,N) arrays so that lists
//                     longer than 22 read the next point's column (the
//                     buffers carry 32 entries of slack)
//   prim/U/q/flux_res/delUp/delUn [5k + r], qm [(5k + r)*2 + s],
//   dq [(5k + r)*3 + c], delt [k]  -- the reference's own (5,N), (2,5,N),
//                     (3,5,N) layouts, so host transfers are plain copies.
#pragma once

#include <hip/hip_runtime.h>

namespace mfcfd {

constexpr int kConnCols = 27;
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
    const int* split_cnt;   // [6][n]
    const int* split;       // [6][split_stride]
    long long split_stride;
    const int* alias;
    const int* respos;      // position of a point in the residual order (wall, then interior)
};

struct DevState {
    double* prim;
    double* U;
    double* q;
    double* qm;
    double* dq;
    double* delt;
    double* flux_res;
    double* delUp;
    double* delUn;
    double* dinv;
};

struct FlowParams {
    double cfl, vl_const, power, pi;
    double rho_inf, u1_inf, u2_inf, u3_inf, pr_inf;
    int inner_iterations;
};

// device-side residual record, state_update.f90:115 and data_structure.f90:58-59
struct ResidualRecord {
    double res_new;
    double sum_res_sqr;
    double max_res;
    int max_res_point;   // 1-based, as the reference stores it
    int pad;
};

}  // namespace mfcfd
