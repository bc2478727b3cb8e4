// preprocess.h -- host-side preprocessing (split stencils, alias order,
// sweep dependency levels).
#pragma once

#include <vector>

namespace mfcfd {

// generate_connectivity.f90:12-246.  cnt[d] (N) and lists[d] (22*N + 32)
// for d = xpos, xneg, ypos, yneg, zpos, zneg; point numbers 1-based.
void generate_split_stencils(int n, const double* x, const double* y, const double* z,
                             const double* tan1, const double* tan2, const double* nor,
                             const int* status, const int* nbhs, const int* conn, int* const cnt[6],
                             int* const lists[6]);

// implicit_aliasing.f90:13-44
void aliasing(int n, const int* nbhs, const int* conn, int* alias, int* pwa);

// Level schedule of the alias-ordered sweeps: a point's level is one more
// than the deepest level among the swept points it reads in that sweep.
struct Schedule {
    struct Inputs {
        int n;
        const int* status;
        const int* nbhs;
        const int* conn;          // (27,N), 1-based
        const int* split_cnt[6];  // (N)
        const int* split[6];      // (22,N)+slack, 1-based
        const int* alias;         // (N)
        const int* pwa;           // (N), 1-based, 0 = empty
        const unsigned char* swept;  // (N) 1 if the point is visited by the sweeps
    };
    std::vector<int> order;               // swept points in alias order (0-based)
    std::vector<int> fwd_ptr, fwd_points;
    std::vector<int> bwd_ptr, bwd_points;
    int n_levels_fwd = 0, n_levels_bwd = 0;
    // dependency lists by sweep position: forward position t is order[t],
    // backward position t is order[M-1-t]; deps hold 0-based point numbers
    std::vector<long long> fwd_dep_ptr, bwd_dep_ptr;
    std::vector<int> fwd_dep, bwd_dep;
    void build(const Inputs& in);
};

}  // namespace mfcfd
