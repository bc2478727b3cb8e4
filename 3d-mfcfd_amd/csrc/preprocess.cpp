// preprocess.cpp -- host-side preprocessing of the point cloud.
//
//  * split stencils   generate_connectivity.f90:12-246
//  * alias order      implicit_aliasing.f90:13-44
//  * sweep schedule   dependency levels of the alias-ordered Gauss-Seidel
//                     sweeps of implicit_forward_and_reverse_sweeps.f90:26-84
//
// The split lists are produced in the reference's (22,N) column-major
// storage: a list that grows past 22 entries keeps writing into the next
// point's column, and points are processed interior, then wall, then outer,
// so later points overwrite earlier spills exactly as in the reference.
#include "preprocess.h"

#include <algorithm>
#include <cstring>

namespace mfcfd {

namespace {

inline void project(const double* x, const double* y, const double* z, const double* tan1,
                    const double* tan2, const double* nor, int i, int k, double& dels, double& delt,
                    double& deln) {
    const double delx = x[k] - x[i];
    const double dely = y[k] - y[i];
    const double delz = z[k] - z[i];
    dels = delx * tan1[3 * i] + dely * tan1[3 * i + 1] + delz * tan1[3 * i + 2];
    delt = delx * tan2[3 * i] + dely * tan2[3 * i + 1] + delz * tan2[3 * i + 2];
    deln = delx * nor[3 * i] + dely * nor[3 * i + 1] + delz * nor[3 * i + 2];
}

}  // namespace

void generate_split_stencils(int n, const double* x, const double* y, const double* z,
                             const double* tan1, const double* tan2, const double* nor,
                             const int* status, const int* nbhs, const int* conn, int* const cnt[6],
                             int* const lists[6]) {
    auto push = [&](int d, int i, int k1) {
        cnt[d][i] += 1;
        lists[d][(long long)i * 22 + (cnt[d][i] - 1)] = k1;
    };
    for (int cls = 0; cls < 3; ++cls) {
        for (int i = 0; i < n; ++i) {
            if (status[i] != cls) continue;
            cnt[0][i] = cnt[1][i] = cnt[2][i] = cnt[3][i] = 0;
            cnt[5][i] = 0;
            if (cls == 0) cnt[4][i] = 0;  // wall and outer do not reset zpos
            for (int r = 0; r < nbhs[i]; ++r) {
                const int k1 = conn[27LL * i + r];
                double dels, delt, deln;
                project(x, y, z, tan1, tan2, nor, i, k1 - 1, dels, delt, deln);
                if (cls == 0) {
                    if (dels <= 0.0) push(0, i, k1);
                    if (dels > 0.0) push(1, i, k1);
                    if (delt <= 0.0) push(2, i, k1);
                    if (delt > 0.0) push(3, i, k1);
                    if (deln <= 0.0) push(4, i, k1);
                    if (deln > 0.0) push(5, i, k1);
                } else {
                    if (dels <= 0.0) push(0, i, k1);
                    if (dels >= 0.0) push(1, i, k1);
                    if (delt <= 0.0) push(2, i, k1);
                    if (delt >= 0.0) push(3, i, k1);
                    if (cls == 1) push(5, i, k1);
                    else push(4, i, k1);
                }
            }
        }
    }
}

void aliasing(int n, const int* nbhs, const int* conn, int* alias, int* pwa) {
    std::memset(alias, 0, sizeof(int) * (size_t)n);
    std::memset(pwa, 0, sizeof(int) * (size_t)n);
    if (n <= 0) return;
    alias[0] = 1;
    pwa[0] = 1;
    int count = 0;
    for (int i = 0; i < n; ++i) {
        const int p = pwa[i];
        if (p == 0) continue;  // the reference reads nbhs(0) = 0 here
        for (int j = 0; j < nbhs[p - 1]; ++j) {
            const int k = conn[27LL * (p - 1) + j];
            if (alias[k - 1] == 0) {
                count += 1;
                alias[k - 1] = count;
                pwa[count - 1] = k;
            }
        }
    }
}

// Dependencies of swept point p (0-based) in one sweep: every entry of the
// lists its LdelU/UdelU visits (x/y split lists, interior z split lists or
// the full conn for wall z) whose alias is below (forward) or above
// (backward) alias(p) and which is itself swept.
static void for_each_dep(const Schedule::Inputs& in, int p, bool forward, std::vector<int>& out,
                         std::vector<int>& mark, int stamp) {
    out.clear();
    const int ap = in.alias[p];
    auto consider = [&](int k) {
        if (k < 0 || k >= in.n) return;
        if (!in.swept[k]) return;
        const int ak = in.alias[k];
        if (forward ? !(ak < ap) : !(ak > ap)) return;
        if (mark[k] == stamp) return;
        mark[k] = stamp;
        out.push_back(k);
    };
    const int dirs = in.status[p] == 0 ? 6 : 4;
    for (int d = 0; d < dirs; ++d) {
        const int c = in.split_cnt[d][p];
        for (int j = 0; j < c; ++j) consider(in.split[d][(long long)p * 22 + j] - 1);
    }
    if (in.status[p] == 1)
        for (int j = 0; j < in.nbhs[p]; ++j) consider(in.conn[27LL * p + j] - 1);
}

void Schedule::build(const Inputs& in) {
    const int n = in.n;
    order.clear();
    for (int i = 0; i < n; ++i) {
        const int p = in.pwa[i] - 1;
        if (p < 0) continue;
        if (in.status[p] == 0 || in.status[p] == 1) order.push_back(p);
    }
    std::vector<int> lev(n, 0), mark(n, -1), deps;
    auto levels_for = [&](bool forward, std::vector<int>& ptr, std::vector<int>& pts, int& nlev,
                          std::vector<long long>& dptr, std::vector<int>& dep) {
        std::fill(lev.begin(), lev.end(), 0);
        std::fill(mark.begin(), mark.end(), -1);
        int maxl = 0;
        const int m = (int)order.size();
        dptr.assign(m + 1, 0);
        dep.clear();
        for (int t = 0; t < m; ++t) {
            const int p = forward ? order[t] : order[m - 1 - t];
            for_each_dep(in, p, forward, deps, mark, p);
            dep.insert(dep.end(), deps.begin(), deps.end());
            dptr[t + 1] = (long long)dep.size();
            int l = 0;
            for (int k : deps) l = std::max(l, lev[k]);
            lev[p] = l + 1;
            maxl = std::max(maxl, l + 1);
        }
        nlev = maxl;
        // level l (0-based) covers pts[ptr[l] .. ptr[l+1]); sweep order is
        // kept inside a level so the schedule is deterministic
        std::vector<int> counts(maxl + 1, 0);
        for (int p : order) counts[lev[p]]++;
        ptr.assign(maxl + 1, 0);
        for (int l = 0; l < maxl; ++l) ptr[l + 1] = ptr[l] + counts[l + 1];
        pts.assign(order.size(), 0);
        std::vector<int> fill(ptr.begin(), ptr.end() - 1);
        for (int t = 0; t < m; ++t) {
            const int p = forward ? order[t] : order[m - 1 - t];
            pts[fill[lev[p] - 1]++] = p;
        }
    };
    levels_for(true, fwd_ptr, fwd_points, n_levels_fwd, fwd_dep_ptr, fwd_dep);
    levels_for(false, bwd_ptr, bwd_points, n_levels_bwd, bwd_dep_ptr, bwd_dep);
}

}  // namespace mfcfd
