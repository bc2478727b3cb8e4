// Test harness (not part of the library): the host preprocessing of
// 3d-mfcfd_amd/csrc/preprocess.cpp over a C ABI, built with AddressSanitizer
// and UndefinedBehaviorSanitizer by tests/sanitize/Makefile and driven by
// tests/sanitize/san_preprocess.py.  Point numbers are 1-based on input, as
// the reference holds them; the sweep plan is built with the identity device
// order, from the lists the flux stencils walk:
//   wall points: zneg walks conn (wall_flux_dGzneg.f90:61-65), zpos nothing;
//   outer points: zneg nothing;
//   supersonic points: nothing;
//   otherwise the split list through its raw (22,N) offset (the spill).
#include <cstddef>
#include <vector>

#include "../../3d-mfcfd_amd/csrc/preprocess.h"

extern "C" {

void san_split_stencils(int n, const double* x, const double* y, const double* z, const double* tan1,
                        const double* tan2, const double* nor, const int* status, const int* nbhs, const int* conn,
                        int* cnt, int* lists) {
    int* c[6];
    int* l[6];
    for (int d = 0; d < 6; ++d) {
        c[d] = cnt + (long long)d * n;
        l[d] = lists + (long long)d * (22LL * n + 32);
    }
    mfcfd::generate_split_stencils(n, x, y, z, tan1, tan2, nor, status, nbhs, conn, c, l);
}

void san_aliasing(int n, const int* nbhs, const int* conn, int* alias, int* pwa) {
    mfcfd::aliasing(n, nbhs, conn, alias, pwa);
}

// Builds the plan of the canonical (backup = 0) or alternative (1) driver.
// Outputs, per direction (0 forward, 1 backward): seq (<= n), take_ptr
// (<= n + 1) and take (<= cap entries); sizes[3 * dir + {0,1,2}] receive
// their lengths; late (when not null, <= n per direction) each position's
// count of trailing entries set apart as late.  Returns build()'s code, or
// -2 when cap is too small.
int san_sweep_plan_late(int n, const int* status, const int* nbhs, const int* conn, const int* cnt,
                        const int* lists, const int* alias, const int* pwa, const double* x, const double* y,
                        const double* z, int backup, long long cap, int* seq, long long* take_ptr, unsigned* take,
                        long long* sizes, unsigned char* late) {
    std::vector<int> dcnt(6LL * n, 0), conn0(27LL * n, -1), lst;
    std::vector<long long> off(6LL * n, 0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < nbhs[i]; ++j) conn0[27LL * i + j] = conn[27LL * i + j] - 1;
    for (int d = 0; d < 6; ++d)
        for (int i = 0; i < n; ++i) {
            const int st = status[i];
            const int* p = nullptr;
            int count = 0;
            if (st == 1 && d == 5) {
                p = conn + 27LL * i;
                count = nbhs[i];
            } else if ((st == 1 && d == 4) || (st == 2 && d == 5) || (st != 0 && st != 1 && st != 2)) {
                p = nullptr;
            } else {
                p = lists + (long long)d * (22LL * n + 32) + 22LL * i;
                count = cnt[(long long)d * n + i];
            }
            off[(long long)d * n + i] = (long long)lst.size();
            dcnt[(long long)d * n + i] = count;
            for (int e = 0; e < count; ++e) lst.push_back(p[e] - 1);
        }
    mfcfd::SweepPlanInputs in;
    in.n = n;
    in.status = status;
    in.alias = alias;
    in.nbhs = nbhs;
    in.conn = conn0.data();
    in.cnt = dcnt.data();
    in.off = off.data();
    in.lst = lst.data();
    in.x = x;
    in.y = y;
    in.z = z;
    in.backup = backup != 0;
    if (!backup) {
        for (int i = 0; i < n; ++i) {
            const int p = pwa[i] - 1;
            if (p < 0) continue;
            const int st = status[p];
            if (st == 0 || st == 1 || st == 5 || st == 6) in.visit_fwd.push_back(p);
        }
        in.visit_bwd.assign(in.visit_fwd.rbegin(), in.visit_fwd.rend());
    } else {
        const int order[5] = {1, 0, 2, 5, 6};
        for (int c : order) {
            for (int i = 0; i < n; ++i)
                if (status[i] == c) in.visit_fwd.push_back(i);
            for (int i = n - 1; i >= 0; --i)
                if (status[i] == c) in.visit_bwd.push_back(i);
        }
    }
    mfcfd::SweepPlan plan;
    const int rc = plan.build(in);
    if (rc) return rc;
    const mfcfd::SweepDirPlan* dirs[2] = {&plan.fwd, &plan.bwd};
    for (int d = 0; d < 2; ++d) {
        const mfcfd::SweepDirPlan& pd = *dirs[d];
        if ((long long)pd.take.size() > cap || (long long)pd.seq.size() > n) return -2;
        for (std::size_t u = 0; u < pd.seq.size(); ++u) seq[(long long)d * n + u] = pd.seq[u];
        for (std::size_t u = 0; u < pd.take_ptr.size(); ++u) take_ptr[(long long)d * (n + 1) + u] = pd.take_ptr[u];
        for (std::size_t u = 0; u < pd.take.size(); ++u) take[d * cap + u] = pd.take[u];
        if (late)
            for (std::size_t u = 0; u < pd.late.size(); ++u) late[(long long)d * n + u] = pd.late[u];
        sizes[3 * d] = (long long)pd.seq.size();
        sizes[3 * d + 1] = (long long)pd.take_ptr.size();
        sizes[3 * d + 2] = (long long)pd.take.size();
    }
    return 0;
}

int san_sweep_plan(int n, const int* status, const int* nbhs, const int* conn, const int* cnt, const int* lists,
                   const int* alias, const int* pwa, const double* x, const double* y, const double* z, int backup,
                   long long cap, int* seq, long long* take_ptr, unsigned* take, long long* sizes) {
    return san_sweep_plan_late(n, status, nbhs, conn, cnt, lists, alias, pwa, x, y, z, backup, cap, seq, take_ptr,
                               take, sizes, nullptr);
}
}
