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

// The LU-SGS sweeps as one plan per sweep direction, built on the host from
// the device-order lists.
//
// The reference visits points in a fixed order -- the alias order of
// point_with_alias (implicit_forward_and_reverse_sweeps.f90:26-85), or the
// type order wall, interior, outer, supersonic inlet, supersonic outlet of
// the alternative driver (implicit_forward_and_reverse_sweeps_backup.f90:
// 25-139) -- and each visited point reads its neighbours' delU in place.  An
// entry a point takes (alias smaller in the forward sweep, larger in the
// backward one) therefore reads this sweep's value of the neighbour when the
// neighbour was visited earlier ("wait" entry), and the value from before
// the sweep otherwise ("old" entry).  A supersonic point copies its nearest
// neighbour's delU under the same rule.
//
// Levels: a point's level is one more than the deepest of (a) the points its
// wait entries read and (b) the earlier-visited points that read it as an old
// entry (they must read before it writes).  Every order by level therefore
// reproduces the visiting order's results, in place (one launch per level) or
// with old values read from a snapshot (the sync-free sweep).
struct SweepDirPlan {
    std::vector<int> seq;             // visited points, device numbers, by level (stable)
    std::vector<int> level_ptr;       // level l covers seq[level_ptr[l] .. level_ptr[l+1])
    std::vector<long long> take_ptr;  // position u takes take[take_ptr[u] .. take_ptr[u+1])
    std::vector<unsigned> take;       // k << 4 | wait << 3 | direction (kTakeCopy: copy entry)
    std::vector<unsigned char> late;  // per position: its last entries, which read the wait
                                      // dependency that executes last (0..kMaxLate; 0: none)
    std::vector<int> old_pts;         // points some regular entry reads as old
    std::vector<int> old_src;         // copy sources read as old
    std::vector<int> wait_src;        // copy sources read as this sweep's value
};

constexpr unsigned kTakeWait = 8u;
// at most this many entries of a position are set apart as late (one lane each)
constexpr int kMaxLate = 4;
constexpr unsigned kTakeCopy = 7u;
constexpr int kTakeShift = 4;

struct SweepPlanInputs {
    int n;                       // points (device numbering 0..n-1)
    const int* status;           // [n]
    const int* alias;            // [n] the reference's alias of each point
    const int* nbhs;             // [n]
    const int* conn;             // [27 n] device numbers
    const int* cnt;              // [6][n] direction lists as the flux stencils walk them
    const long long* off;        // [6][n]
    const int* lst;
    const double* x;             // [n] coordinates, device order (supersonic nearest neighbour)
    const double* y;
    const double* z;
    std::vector<int> visit_fwd;  // the reference's visiting order of each sweep (device numbers)
    std::vector<int> visit_bwd;
    bool backup;                 // the alternative driver: outer points swept, status /= 1 filter
};

struct SweepPlan {
    bool backup = false;
    SweepDirPlan fwd, bwd;
    int n_visited = 0;
    // builds both directions; returns 0, or -1 when a wait entry would not
    // precede its point in the execution order (never, by construction;
    // checked anyway because the sync-free sweep relies on it)
    int build(const SweepPlanInputs& in);
};

// the list a sweep walks for point i and direction d (device numbers):
// the flux stencil's list, except that a wall point takes no ZPOS term and
// an outer point (alternative driver only) walks conn for ZPOS
// (implicit_LdelU_forward_sweep_outer.f90:346-347) and nothing for ZNEG
void sweep_list(const SweepPlanInputs& in, int i, int d, const int*& list, int& count);

// nearest neighbour of a supersonic point (implicit_forward_and_reverse_
// sweeps.f90:39-50; the alternative driver skips wall points, ..._backup.f90:
// 56): strict <, first hit in conn order, start distance as the reference
int nearest_neighbour(const SweepPlanInputs& in, int p, double start, bool skip_wall);

}  // namespace mfcfd
