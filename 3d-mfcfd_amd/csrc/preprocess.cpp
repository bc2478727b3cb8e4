// preprocess.cpp -- host-side preprocessing of the point cloud.
//
//  * split stencils   generate_connectivity.f90:12-246
//  * alias order      implicit_aliasing.f90:13-44
//  * sweep plans      execution order, dependency levels and taken entries
//                     of the LU-SGS sweeps (implicit_forward_and_reverse_
//                     sweeps.f90:26-85 and its alternative, ..._backup.f90)
//
// The split lists are produced in the reference's (22,N) column-major
// storage: a list that grows past 22 entries keeps writing into the next
// point's column, and points are processed interior, then wall, then outer,
// so later points overwrite earlier spills exactly as in the reference.
#include "preprocess.h"

#include <algorithm>
#include <cmath>
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

void sweep_list(const SweepPlanInputs& in, int i, int d, const int*& list, int& count) {
    const int st = in.status[i];
    list = nullptr;
    count = 0;
    if (st == 1 && d == 4) return;  // ZPOS: the wall's z term walks conn as ZNEG
    if (st == 2 && d == 5) return;
    if (st == 2 && d == 4) {
        list = in.conn + 27LL * i;
        count = in.nbhs[i];
        return;
    }
    list = in.lst + in.off[(long long)d * in.n + i];
    count = in.cnt[(long long)d * in.n + i];
}

int nearest_neighbour(const SweepPlanInputs& in, int p, double start, bool skip_wall) {
    double min_dist = start;
    int q = p;
    for (int r = 0; r < in.nbhs[p]; ++r) {
        const int nbh = in.conn[27LL * p + r];
        const double dx = in.x[nbh] - in.x[p];
        const double dy = in.y[nbh] - in.y[p];
        const double dz = in.z[nbh] - in.z[p];
        const double ds = std::sqrt(dx * dx + dy * dy + dz * dz);
        if (ds < min_dist && !(skip_wall && in.status[nbh] == 1)) {
            min_dist = ds;
            q = nbh;
        }
    }
    return q;
}

namespace {

int build_dir(const SweepPlanInputs& in, bool forward, SweepDirPlan& out) {
    const int n = in.n;
    const std::vector<int>& visit = forward ? in.visit_fwd : in.visit_bwd;
    const int m = (int)visit.size();
    std::vector<int> at(n, -1);
    for (int u = 0; u < m; ++u) at[visit[u]] = u;
    // entries in visiting order, then levels
    std::vector<long long> vptr(m + 1, 0);
    std::vector<unsigned> vtake;
    vtake.reserve((size_t)m * 40);
    std::vector<int> level(n, 0);
    std::vector<std::vector<int>> war(n);   // earlier points that read k as old
    std::vector<char> is_old(n, 0), is_old_src(n, 0), is_wait_src(n, 0);
    const double start = in.backup ? 100000.0 : 1000000.0;
    int max_level = 0;
    for (int u = 0; u < m; ++u) {
        const int i = visit[u];
        const int st = in.status[i];
        int lev = 0;
        if (st == 5 || st == 6) {
            const int q = nearest_neighbour(in, i, start, in.backup);
            const bool wait = at[q] >= 0 && at[q] < u;
            vtake.push_back(((unsigned)q << kTakeShift) | (wait ? kTakeWait : 0u) | kTakeCopy);
            if (wait) {
                lev = std::max(lev, level[q]);
                is_wait_src[q] = 1;
            } else {
                is_old_src[q] = 1;
                if (at[q] > u) war[q].push_back(i);
            }
        } else {
            const int ai = in.alias[i];
            for (int d = 0; d < 6; ++d) {
                const int* list;
                int count;
                sweep_list(in, i, d, list, count);
                for (int j = 0; j < count; ++j) {
                    const int k = list[j];
                    const int ak = in.alias[k];
                    if (forward ? !(ak < ai) : !(ak > ai)) continue;
                    const bool wait = at[k] >= 0 && at[k] < u;
                    vtake.push_back(((unsigned)k << kTakeShift) | (wait ? kTakeWait : 0u) | (unsigned)d);
                    if (wait) {
                        lev = std::max(lev, level[k]);
                    } else {
                        is_old[k] = 1;
                        if (at[k] > u) war[k].push_back(i);
                    }
                }
            }
        }
        for (int j : war[i]) lev = std::max(lev, level[j]);
        level[i] = lev + 1;
        max_level = std::max(max_level, lev + 1);
        vptr[u + 1] = (long long)vtake.size();
    }
    // execution order: by level, visiting order inside a level
    std::vector<int> counts(max_level + 2, 0);
    for (int u = 0; u < m; ++u) counts[level[visit[u]]]++;
    out.level_ptr.assign(max_level + 1, 0);
    for (int l = 0; l < max_level; ++l) out.level_ptr[l + 1] = out.level_ptr[l] + counts[l + 1];
    std::vector<int> fill(out.level_ptr.begin(), out.level_ptr.end() - 1);
    std::vector<int> order(m);
    for (int u = 0; u < m; ++u) order[fill[level[visit[u]] - 1]++] = u;
    out.seq.resize(m);
    out.take_ptr.assign(m + 1, 0);
    out.late.assign(m, 0);
    out.take.clear();
    out.take.reserve(vtake.size());
    std::vector<int> exec_at(n, -1);
    for (int e = 0; e < m; ++e) {
        const int u = order[e];
        out.seq[e] = visit[u];
        exec_at[visit[u]] = e;
        out.take.insert(out.take.end(), vtake.begin() + vptr[u], vtake.begin() + vptr[u + 1]);
        out.take_ptr[e + 1] = (long long)out.take.size();
        // by neighbour: a neighbour sits in up to one list per axis (more
        // with the spill), and its entries then sit on adjacent lanes, of
        // which the first reads the neighbour's state for all of them
        const auto b = out.take.begin() + out.take_ptr[e];
        std::sort(b, out.take.end());
        // The wait dependency that executes last usually publishes last, well
        // after the others (on a 500k-point cloud: for 95% of positions, a
        // median 2.7 us after the one before it).  Its entries go last, and
        // the sweep kernel sums the others before it waits for this one.
        int k_last = -1, e_last = -1, n_last = 0;
        for (auto it = b; it != out.take.end(); ++it) {
            const unsigned x = *it;
            if (!(x & kTakeWait) || (x & 7u) == kTakeCopy) continue;
            const int k = (int)(x >> kTakeShift);
            if (exec_at[k] > e_last) e_last = exec_at[k], k_last = k;
        }
        if (k_last >= 0) {
            const auto is_last = [&](unsigned x) {
                return (x & kTakeWait) && (x & 7u) != kTakeCopy && (int)(x >> kTakeShift) == k_last;
            };
            std::stable_partition(b, out.take.end(), [&](unsigned x) { return !is_last(x); });
            n_last = (int)std::count_if(b, out.take.end(), is_last);
        }
        out.late[e] = (unsigned char)(n_last <= kMaxLate ? n_last : 0);
    }
    // every wait entry must read a point that executes earlier
    for (int e = 0; e < m; ++e)
        for (long long t = out.take_ptr[e]; t < out.take_ptr[e + 1]; ++t)
            if (out.take[t] & kTakeWait) {
                const int k = (int)(out.take[t] >> kTakeShift);
                if (exec_at[k] < 0 || exec_at[k] >= e) return -1;
            }
    out.old_pts.clear();
    out.old_src.clear();
    out.wait_src.clear();
    for (int k = 0; k < n; ++k) {
        if (is_old[k]) out.old_pts.push_back(k);
        if (is_old_src[k]) out.old_src.push_back(k);
        if (is_wait_src[k]) out.wait_src.push_back(k);
    }
    return 0;
}

}  // namespace

int SweepPlan::build(const SweepPlanInputs& in) {
    backup = in.backup;
    n_visited = (int)in.visit_fwd.size();
    if (build_dir(in, true, fwd)) return -1;
    if (build_dir(in, false, bwd)) return -1;
    return 0;
}

}  // namespace mfcfd
