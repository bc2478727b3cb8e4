// capi.hip -- C ABI of libmfcfd_hip.so (declared in include/mfcfd_hip.h).
//
// A context owns the HBM copies of the point cloud and the flow state.  At
// bind time the points are renumbered in Morton order (device order), each
// point's six direction lists are resolved from the host's (22,N) arrays
// exactly as the reference reads them, each point gets the walk the fused
// flux kernel visits (conn plus the entries the 22-column spill adds), and
// the Cramer weights of every list are computed.
// Entry points enqueue kernels on the context's stream and return; only
// fpi_solver/run/state_update/get_residual/download wait for the device.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <new>
#include <numeric>
#include <string>
#include <vector>

#include "../../include/mfcfd_hip.h"
#include "device_state.h"
#include "kernels.h"
#include "preprocess.h"

using namespace mfcfd;

namespace {

thread_local char g_last_error[512];

int fail(int code, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(g_last_error, sizeof(g_last_error), fmt, ap);
    va_end(ap);
    return code;
}

#define HIP_TRY(expr)                                                                      \
    do {                                                                                   \
        hipError_t e_ = (expr);                                                            \
        if (e_ != hipSuccess)                                                              \
            return fail(MFCFD_EHIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), __FILE__, \
                        __LINE__);                                                         \
    } while (0)

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    hipError_t alloc(size_t count) {
        release();
        n = count;
        if (count == 0) return hipSuccess;
        return hipMalloc((void**)&p, count * sizeof(T));
    }
    // fine-grained device memory: agent-scope atomic loads and stores to it
    // are coherent across the XCDs while a kernel runs (hipMalloc's
    // coarse-grained memory is coherent only at kernel boundaries)
    hipError_t alloc_fine(size_t count) {
        release();
        n = count;
        if (count == 0) return hipSuccess;
        return hipExtMallocWithFlags((void**)&p, count * sizeof(T), hipDeviceMallocFinegrained);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// 63-bit Morton code of a point in the cloud's bounding box, 21 bits per axis
uint64_t spread21(uint64_t v) {
    v &= 0x1FFFFFULL;
    v = (v | (v << 32)) & 0x1F00000000FFFFULL;
    v = (v | (v << 16)) & 0x1F0000FF0000FFULL;
    v = (v | (v << 8)) & 0x100F00F00F00F00FULL;
    v = (v | (v << 4)) & 0x10C30C30C30C30C3ULL;
    v = (v | (v << 2)) & 0x1249249249249249ULL;
    return v;
}

// inst (may be null): instance of each point; points are ordered by
// instance first, so every instance occupies one contiguous device range,
// and each instance's codes use its own bounding box, so its order is the
// one it has when bound alone
std::vector<int> morton_order(int n, const double* x, const double* y, const double* z,
                              const int* inst = nullptr) {
    int n_inst = 1;
    if (inst)
        for (int i = 0; i < n; ++i) n_inst = std::max(n_inst, inst[i] + 1);
    std::vector<double> lo(3 * (size_t)n_inst, 0.0), hi(3 * (size_t)n_inst, 0.0);
    std::vector<char> seen(n_inst, 0);
    for (int i = 0; i < n; ++i) {
        const int b = inst ? inst[i] : 0;
        const double v[3] = {x[i], y[i], z[i]};
        for (int c = 0; c < 3; ++c) {
            if (!seen[b] || v[c] < lo[3 * b + c]) lo[3 * b + c] = v[c];
            if (!seen[b] || v[c] > hi[3 * b + c]) hi[3 * b + c] = v[c];
        }
        seen[b] = 1;
    }
    std::vector<std::pair<std::pair<int, uint64_t>, int>> key(n);
    for (int i = 0; i < n; ++i) {
        const int b = inst ? inst[i] : 0;
        const double v[3] = {x[i], y[i], z[i]};
        uint64_t code = 0;
        for (int c = 0; c < 3; ++c) {
            const double span = hi[3 * b + c] > lo[3 * b + c] ? hi[3 * b + c] - lo[3 * b + c] : 1.0;
            double t = (v[c] - lo[3 * b + c]) / span * 2097151.0;
            t = std::min(std::max(t, 0.0), 2097151.0);
            code |= spread21((uint64_t)t) << c;
        }
        key[i] = {{inst ? inst[i] : 0, code}, i};
    }
    std::sort(key.begin(), key.end());
    std::vector<int> perm(n);
    for (int d = 0; d < n; ++d) perm[d] = key[d].second;
    return perm;
}

}  // namespace

struct mfcfd_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool have_params = false;
    bool have_geom = false;
    mfcfd_params params{};
    FlowParams fp{};

    int n = 0;
    // geometry, device order
    DevBuf<double4> pos;
    DevBuf<double> frame, coef, dgeo;   // dgeo: per direction list, D_inv's Cramer numerator and determinant
    DevBuf<int> status, nbhs, conn, dir_cnt, lst, alias, respos, perm, nwalk;
    DevBuf<long long> dir_off;
    DevBuf<unsigned> walk;
    DevBuf<int> idx_interior, idx_wall, idx_outer, idx_sup_in, idx_sup_out;
    DevBuf<int> idx_fused_interior, idx_fused_wall, idx_generic;
    TypeLists tl{};
    // state
    DevBuf<double> prim, U, q, qm, dq, delt, tsden, flux_res, delUp, delUn, dinv, temp, staging;
    DevBuf<float> q32, qm32, dq32;   // mixed precision (q32: q and an FP32 tsden per point)
    // residual
    DevBuf<double> partial;
    DevBuf<int> partial_pos;
    DevBuf<int> stencil_bad;   // bind-time singular-stencil flag
    int n_partial = 0;
    DevBuf<ResidualRecord> rec;
    DevBuf<double> hist;
    ResidualRecord* host_rec = nullptr;  // pinned
    // LU-SGS sweep plan of one sweep direction on the device (SweepPlan):
    // execution order, taken entries and their weights, snapshot and mark lists
    struct DevSweepDir {
        DevBuf<int> seq, old_pts, old_src, wait_src;
        DevBuf<long long> take_ptr;
        DevBuf<unsigned> take;
        DevBuf<unsigned char> late;
        DevBuf<double> wgt;
        std::vector<int> level_ptr;   // host copy of the level cuts of seq
        int count = 0;
        void release() {
            seq.release();
            old_pts.release();
            old_src.release();
            wait_src.release();
            take_ptr.release();
            take.release();
            late.release();
            wgt.release();
            level_ptr.clear();
            count = 0;
        }
    };
    DevSweepDir sw_fwd, sw_bwd;
    int plan_form = -1;    // iteration form the plan was built for (-1: none)
    // sweep values: published prim (fine-grained: polled while the sweep
    // runs), prim and delU before the sweep, published delU of the copy
    // sources (fine-grained) and each point's slot in it (-1: none)
    DevBuf<double> pub, pold, dold, dsrc;
    DevBuf<int> src_slot_f, src_slot_b;
    DevBuf<unsigned> sweep_ctr;
    // host inputs of plan building (device order), kept from bind
    std::vector<int> h_status, h_alias, h_nbhs, h_conn, h_cnt, h_lst;
    std::vector<long long> h_off;
    std::vector<double> h_x, h_y, h_z;
    std::vector<int> h_pwa;                              // point_with_alias as device numbers, -1 empty
    std::vector<int> h_wall, h_int, h_out, h_sin, h_sout;  // class lists, host order, device numbers
    // clouds of a batch (mfcfd_bind_batch): instance of each device point,
    // freestream (u1_inf, u2_inf) per instance, device-order ranges; a single
    // cloud is one instance with batch = false
    bool batch = false;
    int n_inst = 1;
    DevBuf<int> inst_id, inst_beg, inst_end, hist_ctr;
    DevBuf<double> inst_fs, inst_norm, ressq, inst_hist;
    DevBuf<int> chunk_beg, chunk_end, chunk_ptr;   // per-cloud residual chunks
    DevBuf<double> chunk_sum;
    int n_chunks = 0;
    std::vector<int> inst_count;
    // launch shapes chosen by mfcfd_autotune (defaults until then)
    LaunchShapes shapes;
    // one iteration captured as a HIP graph (mfcfd_run_graph)
    hipGraphExec_t graph_exec = nullptr;
    bool graph_stale = true;
    int sweep_grid = 0, sweep_grid_mixed = 0;
    // bookkeeping of fpi_solver.f90:38-43
    double res_old = 0.0, res_new = 0.0, residue = 0.0, max_res = 0.0, sum_res_sqr = 0.0;
    int max_res_point = 0, last_t = 0;
    int n_regular = 0;
    int n_fused_interior_pts = 0, n_fused_general_pts = 0;
    std::vector<int> host_class;   // per host point: 0 fused interior, 1 fused general, 2 list-by-list
    // HIP events around each fused interior flux launch (mfcfd_kernel_timing)
    int time_kernel = 0;   // MFCFD_TIME_* of mfcfd_kernel_timing (0: off)
    std::vector<hipEvent_t> tev;
    size_t tev_used = 0;
    double tev_ms = 0.0;
    long long tev_launches = 0;

    DevGeom geom() const {
        DevGeom g;
        g.n = n;
        g.pos = pos.p;
        g.frame = frame.p;
        g.status = status.p;
        g.nbhs = nbhs.p;
        g.conn = conn.p;
        g.dir_cnt = dir_cnt.p;
        g.dir_off = dir_off.p;
        g.lst = lst.p;
        g.walk = walk.p;
        g.nwalk = nwalk.p;
        g.coef = coef.p;
        g.dgeo = dgeo.p;
        g.alias = alias.p;
        g.respos = respos.p;
        g.perm = perm.p;
        g.inst = batch ? inst_id.p : nullptr;
        g.inst_fs = batch ? inst_fs.p : nullptr;
        return g;
    }
    DevState state() const {
        DevState s;
        s.prim = prim.p;
        s.U = U.p;
        s.q = q.p;
        s.qm = qm.p;
        s.dq = dq.p;
        s.delt = delt.p;
        s.tsden = tsden.p;
        s.flux_res = flux_res.p;
        s.delUp = delUp.p;
        s.delUn = delUn.p;
        s.dinv = dinv.p;
        s.q32 = q32.p;
        s.qm32 = qm32.p;
        s.dq32 = dq32.p;
        return s;
    }
    void release_all() {
        DevBuf<double>* dbl[] = {&frame, &coef, &dgeo, &prim, &U, &q, &qm, &dq, &delt, &tsden, &flux_res, &delUp, &delUn,
                                 &dinv, &temp, &staging, &partial, &hist};
        for (auto* b : dbl) b->release();
        DevBuf<int>* ints[] = {&status, &nbhs, &conn, &dir_cnt, &lst, &alias, &respos, &perm, &nwalk,
                               &idx_interior,
                               &idx_wall, &idx_outer, &idx_sup_in, &idx_sup_out, &idx_fused_interior,
                               &idx_fused_wall, &idx_generic, &partial_pos, &stencil_bad};
        for (auto* b : ints) b->release();
        sw_fwd.release();
        sw_bwd.release();
        plan_form = -1;
        pub.release();
        pold.release();
        dold.release();
        dsrc.release();
        src_slot_f.release();
        src_slot_b.release();
        q32.release();
        qm32.release();
        dq32.release();
        sweep_ctr.release();
        DevBuf<int>* ib[] = {&inst_id, &inst_beg, &inst_end, &hist_ctr, &chunk_beg, &chunk_end, &chunk_ptr};
        for (auto* b : ib) b->release();
        DevBuf<double>* db[] = {&inst_fs, &inst_norm, &ressq, &inst_hist, &chunk_sum};
        for (auto* b : db) b->release();
        dir_off.release();
        walk.release();
        pos.release();
    }
    ~mfcfd_ctx() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        for (hipEvent_t e : tev) (void)hipEventDestroy(e);
        release_all();
        rec.release();
        if (host_rec) (void)hipHostFree(host_rec);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace {

int check_ctx(mfcfd_ctx* c, bool need_geom) {
    if (!c) return fail(MFCFD_EINVAL, "null context");
    if (!c->have_params) return fail(MFCFD_ESTATE, "mfcfd_set_params has not been called");
    if (need_geom && !c->have_geom) return fail(MFCFD_ESTATE, "mfcfd_bind_geometry has not been called");
    if (hipSetDevice(c->device) != hipSuccess) return fail(MFCFD_EHIP, "hipSetDevice(%d) failed", c->device);
    return MFCFD_OK;
}

template <typename T>
int upload(DevBuf<T>& b, const T* host, size_t count, hipStream_t st) {
    if (b.alloc(count) != hipSuccess) return fail(MFCFD_ENOMEM, "hipMalloc of %zu bytes failed", count * sizeof(T));
    if (count) HIP_TRY(hipMemcpyAsync(b.p, host, count * sizeof(T), hipMemcpyHostToDevice, st));
    return MFCFD_OK;
}

struct ArrayDesc {
    DevBuf<double>* buf;
    int per_point;
};

// FP32 array the mixed form keeps for an array id (null: none)
float* mixed_copy_of(mfcfd_ctx* c, int id) {
    switch (id) {
        case MFCFD_ARR_Q: return c->q32.p;
        case MFCFD_ARR_QM: return c->qm32.p;
        case MFCFD_ARR_DQ: return c->dq32.p;
        default: return nullptr;
    }
}

ArrayDesc array_of(mfcfd_ctx* c, int id) {
    switch (id) {
        case MFCFD_ARR_PRIM: return {&c->prim, 5};
        case MFCFD_ARR_U: return {&c->U, 5};
        case MFCFD_ARR_Q: return {&c->q, 5};
        case MFCFD_ARR_QM: return {&c->qm, 10};
        case MFCFD_ARR_DQ: return {&c->dq, 15};
        case MFCFD_ARR_DELT: return {&c->delt, 1};
        case MFCFD_ARR_FLUX_RES: return {&c->flux_res, 5};
        case MFCFD_ARR_DELUP: return {&c->delUp, 5};
        case MFCFD_ARR_DELUN: return {&c->delUn, 5};
        default: return {nullptr, 0};
    }
}

int launch_error(const char* what) {
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(MFCFD_EHIP, "%s launch: %s", what, hipGetErrorString(e));
    return MFCFD_OK;
}

// index into c->tev of a pair of timing events (begin, end) for one more
// timed launch, or -1.  An index, not a pointer: a later call may grow (and
// move) the vector.
long long timing_pair(mfcfd_ctx* c) {
    if (c->tev_used + 2 > c->tev.size()) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess) return -1;
        if (hipEventCreate(&b) != hipSuccess) {
            (void)hipEventDestroy(a);
            return -1;
        }
        c->tev.push_back(a);
        c->tev.push_back(b);
    }
    const long long at = (long long)c->tev_used;
    c->tev_used += 2;
    return at;
}

void flux(mfcfd_ctx* c, int parts) {
    if (c->params.flux_order == MFCFD_FLUX_LISTS) {
        launch_flux_residual_reference(c->geom(), c->state(), c->fp, c->tl, c->stream);
        return;
    }
    const LaunchShapes& ls = c->shapes;
    if (c->time_kernel == MFCFD_TIME_FLUX_INTERIOR && (parts & 1)) {
        if (c->tev_used + 2 > c->tev.size()) {
            hipEvent_t a, b;
            if (hipEventCreate(&a) == hipSuccess) {
                if (hipEventCreate(&b) == hipSuccess) {
                    c->tev.push_back(a);
                    c->tev.push_back(b);
                } else {
                    (void)hipEventDestroy(a);
                }
            }
        }
        if (c->tev_used + 2 <= c->tev.size()) {
            hipEventRecord(c->tev[c->tev_used], c->stream);
            launch_flux_residual(c->geom(), c->state(), c->fp, c->tl, c->stream, 1, ls);
            hipEventRecord(c->tev[c->tev_used + 1], c->stream);
            c->tev_used += 2;
            parts &= ~1;
        }
    }
    if (parts) launch_flux_residual(c->geom(), c->state(), c->fp, c->tl, c->stream, parts, ls);
}

// after a stream synchronize: fold the recorded launch times into the totals
void collect_kernel_times(mfcfd_ctx* c) {
    for (size_t e = 0; e + 1 < c->tev_used; e += 2) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->tev[e], c->tev[e + 1]) == hipSuccess) {
            c->tev_ms += ms;
            c->tev_launches += 1;
        }
    }
    c->tev_used = 0;
}

// the implicit iteration form of the current parameters
int implicit_form(const mfcfd_ctx* c) { return c->params.mode == MFCFD_EXPLICIT ? MFCFD_IMPLICIT : c->params.mode; }

SweepAsync sweep_args(mfcfd_ctx* c, mfcfd_ctx::DevSweepDir& d) {
    SweepAsync a;
    a.seq = d.seq.p;
    a.count = d.count;
    a.take_ptr = d.take_ptr.p;
    a.take = d.take.p;
    a.late = d.late.p;
    a.wgt = d.wgt.p;
    a.old_pts = d.old_pts.p;
    a.n_old = (int)d.old_pts.n;
    a.old_src = d.old_src.p;
    a.n_old_src = (int)d.old_src.n;
    a.wait_src = d.wait_src.p;
    a.n_wait_src = (int)d.wait_src.n;
    a.pub = c->pub.p;
    a.pold = c->pold.p;
    a.dold = c->dold.p;
    a.dsrc = c->dsrc.p;
    a.src_slot = (&d == &c->sw_fwd) ? c->src_slot_f.p : c->src_slot_b.p;
    a.ctr = c->sweep_ctr.p;
    a.stalls = &c->rec.p->stalls;
    a.zero = 0;
    return a;
}

int do_sweeps(mfcfd_ctx* c) {
    const DevGeom g = c->geom();
    const DevState s = c->state();
    if (c->params.mode == MFCFD_EXPLICIT) {
        launch_explicit_sweep(g, s, c->stream);
        return launch_error("explicit sweep");
    }
    if (c->plan_form != c->params.mode)
        return fail(MFCFD_ESTATE, "no sweep plan for iteration form %d (internal: prepare_plan not called)",
                    c->params.mode);
    launch_dinv(g, s, c->fp, c->tl, c->stream);
    if (c->params.sweep_schedule == MFCFD_SWEEP_LEVELS) {
        SweepSchedule sch;
        sch.n_levels_fwd = (int)c->sw_fwd.level_ptr.size() - 1;
        sch.n_levels_bwd = (int)c->sw_bwd.level_ptr.size() - 1;
        sch.fwd_ptr = c->sw_fwd.level_ptr.data();
        sch.bwd_ptr = c->sw_bwd.level_ptr.data();
        sch.fwd_points = c->sw_fwd.seq.p;
        sch.bwd_points = c->sw_bwd.seq.p;
        sch.n_swept = c->sw_fwd.count;
        launch_sweeps_levels(g, s, c->fp, sch, c->stream);
        return launch_error("implicit sweeps");
    }
    const SweepAsync fwd = sweep_args(c, c->sw_fwd);
    const SweepAsync bwd = sweep_args(c, c->sw_bwd);
    const int grid = c->fp.mixed ? c->sweep_grid_mixed : c->sweep_grid;
    // events around each sweep kernel launch (mfcfd_kernel_timing)
    hipEvent_t ev[4];
    hipEvent_t* evp = nullptr;
    if (c->time_kernel == MFCFD_TIME_SWEEPS) {
        const long long a = timing_pair(c);
        const long long b = a >= 0 ? timing_pair(c) : -1;
        if (a >= 0 && b >= 0) {
            for (int e = 0; e < 2; ++e) {
                ev[e] = c->tev[a + e];
                ev[2 + e] = c->tev[b + e];
            }
            evp = ev;
        } else if (a >= 0) {
            c->tev_used -= 2;
        }
    }
    launch_sweeps_async(g, s, c->fp, fwd, bwd, std::max(1, grid / std::max(1, c->shapes.sweep_div)), c->stream,
                        evp);
    return launch_error("implicit sweeps");
}

// Builds the LU-SGS plan of the given implicit form (SweepPlan) from the
// host copies kept at bind, uploads it and computes every entry's weight.
// Synchronous; never called while a graph is being captured.
int build_plan(mfcfd_ctx* c, int form) {
    const int n = c->n;
    const bool backup = form == MFCFD_IMPLICIT_BACKUP;
    SweepPlanInputs in;
    in.n = n;
    in.status = c->h_status.data();
    in.alias = c->h_alias.data();
    in.nbhs = c->h_nbhs.data();
    in.conn = c->h_conn.data();
    in.cnt = c->h_cnt.data();
    in.off = c->h_off.data();
    in.lst = c->h_lst.data();
    in.x = c->h_x.data();
    in.y = c->h_y.data();
    in.z = c->h_z.data();
    in.backup = backup;
    if (!backup) {
        // implicit_forward_and_reverse_sweeps.f90:26-85: point_with_alias
        // order; wall, interior and supersonic points are visited, outer
        // points and empty entries are not
        for (int p : c->h_pwa) {
            if (p < 0) continue;
            const int st = c->h_status[p];
            if (st == 0 || st == 1 || st == 5 || st == 6) in.visit_fwd.push_back(p);
        }
        in.visit_bwd.assign(in.visit_fwd.rbegin(), in.visit_fwd.rend());
    } else {
        // implicit_forward_and_reverse_sweeps_backup.f90:25-139: each class
        // in index order, and in reverse for the backward sweep
        const std::vector<int>* cls[5] = {&c->h_wall, &c->h_int, &c->h_out, &c->h_sin, &c->h_sout};
        for (auto* v : cls) {
            in.visit_fwd.insert(in.visit_fwd.end(), v->begin(), v->end());
            in.visit_bwd.insert(in.visit_bwd.end(), v->rbegin(), v->rend());
        }
    }
    SweepPlan plan;
    if (plan.build(in)) return fail(MFCFD_EINVAL, "sweep plan: a dependency does not precede its point");
    hipStream_t st = c->stream;
    c->sw_fwd.release();
    c->sw_bwd.release();
    c->plan_form = -1;
    int rc = 0;
    auto put = [&](mfcfd_ctx::DevSweepDir& d, const SweepDirPlan& pd) -> int {
        int r = 0;
        d.count = (int)pd.seq.size();
        d.level_ptr = pd.level_ptr;
        const int none = 0;
        const unsigned none_u = 0u;
        if ((r = upload(d.seq, pd.seq.empty() ? &none : pd.seq.data(), std::max<size_t>(pd.seq.size(), 1), st)))
            return r;
        if ((r = upload(d.take_ptr, pd.take_ptr.data(), pd.take_ptr.size(), st))) return r;
        if ((r = upload(d.take, pd.take.empty() ? &none_u : pd.take.data(), std::max<size_t>(pd.take.size(), 1), st)))
            return r;
        const unsigned char none_c = 0;
        if ((r = upload(d.late, pd.late.empty() ? &none_c : pd.late.data(), std::max<size_t>(pd.late.size(), 1), st)))
            return r;
        if (d.wgt.alloc(std::max<size_t>(pd.take.size(), 1)) != hipSuccess)
            return fail(MFCFD_ENOMEM, "sweep weights (%zu entries)", pd.take.size());
        if ((r = upload(d.old_pts, pd.old_pts.data(), pd.old_pts.size(), st))) return r;
        if ((r = upload(d.old_src, pd.old_src.data(), pd.old_src.size(), st))) return r;
        if ((r = upload(d.wait_src, pd.wait_src.data(), pd.wait_src.size(), st))) return r;
        return MFCFD_OK;
    };
    if ((rc = put(c->sw_fwd, plan.fwd))) return rc;
    if ((rc = put(c->sw_bwd, plan.bwd))) return rc;
    if (c->pub.n == 0 && c->pub.alloc_fine(5ULL * n) != hipSuccess)
        return fail(MFCFD_ENOMEM, "sweep values (fine-grained)");
    // copy sources read as this sweep's delU: a slot each in dsrc
    const size_t n_src = std::max(plan.fwd.wait_src.size(), plan.bwd.wait_src.size());
    if (n_src > 0) {
        if (c->dsrc.n < 5 * n_src && c->dsrc.alloc_fine(5 * n_src) != hipSuccess)
            return fail(MFCFD_ENOMEM, "sweep values (fine-grained)");
        for (int dir = 0; dir < 2; ++dir) {
            const std::vector<int>& ws = dir == 0 ? plan.fwd.wait_src : plan.bwd.wait_src;
            std::vector<int> slot(n, -1);
            for (size_t j = 0; j < ws.size(); ++j) slot[ws[j]] = (int)j;
            if ((rc = upload(dir == 0 ? c->src_slot_f : c->src_slot_b, slot.data(), slot.size(), st))) return rc;
            HIP_TRY(hipStreamSynchronize(st));   // slot ends here
        }
    }
    if (c->pold.n == 0 && c->pold.alloc(5ULL * n) != hipSuccess) return fail(MFCFD_ENOMEM, "sweep values");
    const bool copies = !plan.fwd.old_src.empty() || !plan.bwd.old_src.empty();
    if (copies && c->dold.n == 0 && c->dold.alloc(5ULL * n) != hipSuccess)
        return fail(MFCFD_ENOMEM, "sweep values");
    if (c->sweep_ctr.n == 0) {
        if (c->sweep_ctr.alloc(kSweepCounters) != hipSuccess) return fail(MFCFD_ENOMEM, "sweep counters");
        HIP_TRY(hipMemsetAsync(c->sweep_ctr.p, 0, kSweepCounters * sizeof(unsigned), st));
    }
    launch_sweep_weights(c->geom(), sweep_args(c, c->sw_fwd), c->fp.power, st);
    launch_sweep_weights(c->geom(), sweep_args(c, c->sw_bwd), c->fp.power, st);
    if ((rc = launch_error("sweep weights"))) return rc;
    HIP_TRY(hipStreamSynchronize(st));   // the host plan ends here
    c->sweep_grid = sweep_async_grid(c->sw_fwd.count, c->params.power == 0.0, false);
    c->sweep_grid_mixed = sweep_async_grid(c->sw_fwd.count, c->params.power == 0.0, true);
    c->plan_form = form;
    c->graph_stale = true;
    return MFCFD_OK;
}

// every entry point that may run the implicit sweeps calls this first (and
// outside any graph capture): the plan of the current iteration form
int prepare_plan(mfcfd_ctx* c) {
    if (c->params.mode == MFCFD_EXPLICIT || c->plan_form == c->params.mode) return MFCFD_OK;
    return build_plan(c, c->params.mode);
}

int do_state_update(mfcfd_ctx* c, double* hist, int hist_index, const InstanceResidual* ir = nullptr,
                    bool with_explicit_sweep = false) {
    launch_state_update(c->geom(), c->state(), c->fp, c->tl, c->partial.p, c->partial_pos.p, c->n_partial,
                        c->rec.p, hist, hist_index, c->stream, ir, with_explicit_sweep);
    return launch_error("state update");
}

// the sweeps and the state update of one iteration; the explicit form's
// sweep runs inside the state update's pass
int do_sweeps_and_update(mfcfd_ctx* c, double* hist, int hist_index, const InstanceResidual* ir = nullptr) {
    if (c->params.mode == MFCFD_EXPLICIT) return do_state_update(c, hist, hist_index, ir, true);
    int rc = do_sweeps(c);
    if (rc) return rc;
    return do_state_update(c, hist, hist_index, ir);
}

int do_iteration(mfcfd_ctx* c, double* hist, int hist_index, const InstanceResidual* ir = nullptr) {
    const DevGeom g = c->geom();
    const DevState s = c->state();
    launch_conserved_and_q(g, s, c->fp, c->stream);
    if (c->fp.inner_iterations > 0) {
        launch_q_derivatives(g, s, c->fp, c->temp.p, c->stream);
        launch_timestep(g, s, c->fp, c->stream);
    } else {
        launch_qderiv_timestep(g, s, c->fp, c->stream, c->shapes.qderiv);
    }
    flux(c, kFluxAll);
    int rc = launch_error("iteration");
    if (rc) return rc;
    return do_sweeps_and_update(c, hist, hist_index, ir);
}

// fpi_solver.f90:38-43
void residue_bookkeeping(mfcfd_ctx* c, int t, double res_new) {
    c->res_new = res_new;
    if (t <= 2) {
        c->res_old = res_new;
        c->residue = 0.0;
    } else {
        c->residue = std::log10(res_new / c->res_old);
    }
    c->last_t = t;
}

int fetch_record(mfcfd_ctx* c) {
    HIP_TRY(hipMemcpyAsync(c->host_rec, c->rec.p, sizeof(ResidualRecord), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    collect_kernel_times(c);
    c->res_new = c->host_rec->res_new;
    c->sum_res_sqr = c->host_rec->sum_res_sqr;
    c->max_res = c->host_rec->max_res;
    if (c->host_rec->max_res_point > 0) c->max_res_point = c->host_rec->max_res_point;
    if (c->host_rec->stalls) {
        // reported once: clear the count so later sweeps wait normally again
        // (this runs outside any graph capture)
        const int stalls = c->host_rec->stalls;
        HIP_TRY(hipMemsetAsync(&c->rec.p->stalls, 0, sizeof(int), c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        return fail(MFCFD_EHIP, "implicit sweep: %d dependency waits exceeded the spin limit", stalls);
    }
    return MFCFD_OK;
}

// The list direction d of host point r walks, as the reference reads it:
// split lists through the raw (22,N) offset, wall zneg and outer zneg
// replaced by what the reference walks instead (conn / nothing).
struct HostList {
    const int* p;
    int count;
};

HostList host_list(const mfcfd_geometry* gm, int r, int d) {
    const int st = gm->status[r];
    if (st == 1 && d == ZNEG) return {gm->conn + 27LL * r, gm->nbhs[r]};  // wall_flux_dGzneg.f90:61-65
    if (st == 1 && d == ZPOS) return {nullptr, 0};                          // not walked for wall points
    if (st == 2 && d == ZNEG) return {nullptr, 0};                          // not walked for outer points
    if (st != 0 && st != 1 && st != 2) return {nullptr, 0};
    return {gm->split_conn[d] + 22LL * r, gm->split_nbhs[d][r]};
}

// The per-iteration residual history of mfcfd_run_noalloc: one double per
// iteration, written by the state update's last workgroup, in a file-scope
// device array so that the call allocates nothing.  Every device holds its
// own copy for the life of the process, and host code reaches it only
// through the address hipGetSymbolAddress returned with that device current
// (the array's name in host code is not a device address).  Calls on one
// device share the array: two contexts that run mfcfd_run_noalloc at the same
// time on one device race.  Every entry a call reads is written by the same
// call first, so nothing is cleared.
static __device__ __align__(256) double scratch[MFCFD_NOALLOC_MAX_ITERS];

constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

// once per device, with that device current (mfcfd_create)
int scratch_address(int device) {
    if (device >= kMaxDevices) return fail(MFCFD_EINVAL, "device %d beyond the %d supported", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch[device]) return MFCFD_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(MFCFD_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    g_scratch[device] = static_cast<double*>(p);
    return MFCFD_OK;
}

}  // namespace

// ---------------------------------------------------------------------------
extern "C" {

const char* mfcfd_last_error(void) { return g_last_error; }
int mfcfd_abi_version(void) { return MFCFD_ABI_VERSION; }

int mfcfd_create(int device, mfcfd_ctx** out) {
    if (!out) return fail(MFCFD_EINVAL, "null output pointer");
    *out = nullptr;
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(MFCFD_EINVAL, "device %d not present (%d devices)", device, count);
    HIP_TRY(hipSetDevice(device));
    if (int rc = scratch_address(device)) return rc;
    mfcfd_ctx* c = new (std::nothrow) mfcfd_ctx();
    if (!c) return fail(MFCFD_ENOMEM, "context allocation failed");
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return fail(MFCFD_EHIP, "hipStreamCreate failed");
    }
    if (c->rec.alloc(1) != hipSuccess ||
        hipHostMalloc((void**)&c->host_rec, sizeof(ResidualRecord), hipHostMallocDefault) != hipSuccess) {
        delete c;
        return fail(MFCFD_ENOMEM, "residual buffers");
    }
    *out = c;
    return MFCFD_OK;
}

int mfcfd_destroy(mfcfd_ctx* ctx) {
    if (!ctx) return MFCFD_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    delete ctx;
    return MFCFD_OK;
}

int mfcfd_set_params(mfcfd_ctx* c, const mfcfd_params* p) {
    if (!c || !p) return fail(MFCFD_EINVAL, "null argument");
    if (p->mode != MFCFD_EXPLICIT && p->mode != MFCFD_IMPLICIT && p->mode != MFCFD_IMPLICIT_BACKUP)
        return fail(MFCFD_EINVAL, "mode %d", p->mode);
    if (p->precision != MFCFD_FP64 && p->precision != MFCFD_MIXED)
        return fail(MFCFD_EINVAL, "precision %d", p->precision);
    if (p->precision == MFCFD_MIXED) {
        if (p->flux_order != MFCFD_FLUX_FUSED)
            return fail(MFCFD_EUNSUPPORTED, "mixed precision runs with the fused flux order only");
        if (p->sweep_schedule != MFCFD_SWEEP_ASYNC)
            return fail(MFCFD_EUNSUPPORTED, "mixed precision runs with the sync-free sweeps only");
        if (p->inner_iterations > 0)
            return fail(MFCFD_EUNSUPPORTED, "mixed precision has no second-order inner iterations");
    }
    if (p->flux_order != MFCFD_FLUX_FUSED && p->flux_order != MFCFD_FLUX_LISTS)
        return fail(MFCFD_EINVAL, "flux_order %d", p->flux_order);
    if (p->sweep_schedule != MFCFD_SWEEP_ASYNC && p->sweep_schedule != MFCFD_SWEEP_LEVELS)
        return fail(MFCFD_EINVAL, "sweep_schedule %d", p->sweep_schedule);
    if (!(p->cfl > 0.0)) return fail(MFCFD_EINVAL, "CFL must be positive");
    if (p->inner_iterations < 0) return fail(MFCFD_EINVAL, "inner_iterations < 0");
    if (c->have_geom && p->power != c->params.power)
        return fail(MFCFD_EINVAL, "power changes the stencil weights; set it before mfcfd_bind_geometry");
    c->params = *p;
    // parameter.f90:13-33
    FlowParams& f = c->fp;
    f.cfl = p->cfl;
    f.vl_const = p->vl_const;
    f.power = p->power;
    f.pi = 4.0 * std::atan(1.0);
    const double theta = p->aoa * f.pi / 180.0;
    f.rho_inf = 1.0;
    f.u1_inf = p->mach * std::cos(theta);
    f.u2_inf = p->mach * std::sin(theta);
    f.u3_inf = 0.0;
    f.pr_inf = 1.0 / 1.4;
    f.inner_iterations = p->inner_iterations;
    f.mixed = p->precision == MFCFD_MIXED ? 1 : 0;
    f.backup = p->mode == MFCFD_IMPLICIT_BACKUP ? 1 : 0;
    c->have_params = true;
    c->graph_stale = true;
    if (c->have_geom && p->inner_iterations > 0 && c->temp.n == 0) {
        if (c->temp.alloc(15ULL * c->n) != hipSuccess) return fail(MFCFD_ENOMEM, "temp(3,5,N)");
    }
    return MFCFD_OK;
}

int mfcfd_generate_split_stencils(int n, const double* x, const double* y, const double* z,
                                  const double* tan1, const double* tan2, const double* nor,
                                  const int* status, const int* nbhs, const int* conn, int* split_nbhs[6],
                                  int* split_conn[6]) {
    if (n <= 0 || !x || !y || !z || !tan1 || !tan2 || !nor || !status || !nbhs || !conn)
        return fail(MFCFD_EINVAL, "null or empty input");
    for (int d = 0; d < 6; ++d)
        if (!split_nbhs[d] || !split_conn[d]) return fail(MFCFD_EINVAL, "null split array %d", d);
    for (int i = 0; i < n; ++i) {
        if (nbhs[i] < 0 || nbhs[i] > kConnCols) return fail(MFCFD_EINVAL, "point %d: nbhs=%d", i + 1, nbhs[i]);
        for (int j = 0; j < nbhs[i]; ++j)
            if (conn[27LL * i + j] < 1 || conn[27LL * i + j] > n)
                return fail(MFCFD_EINVAL, "point %d: conn(%d)=%d out of range", i + 1, j + 1, conn[27LL * i + j]);
    }
    generate_split_stencils(n, x, y, z, tan1, tan2, nor, status, nbhs, conn, split_nbhs, split_conn);
    return MFCFD_OK;
}

int mfcfd_aliasing(int n, const int* nbhs, const int* conn, int* alias, int* point_with_alias) {
    if (n <= 0 || !nbhs || !conn || !alias || !point_with_alias) return fail(MFCFD_EINVAL, "null or empty input");
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < nbhs[i]; ++j)
            if (conn[27LL * i + j] < 1 || conn[27LL * i + j] > n)
                return fail(MFCFD_EINVAL, "point %d: conn(%d) out of range", i + 1, j + 1);
    aliasing(n, nbhs, conn, alias, point_with_alias);
    return MFCFD_OK;
}

namespace {

// inst_of (host order) and fs (u1_inf, u2_inf per instance) describe a
// batch; both null for a single cloud
int bind_impl(mfcfd_ctx* c, const mfcfd_geometry* gm, const int* inst_of, int n_inst, const double* fs) {
    int rc = 0;
    if (!gm || gm->n <= 0) return fail(MFCFD_EINVAL, "empty geometry");
    // walk entries pack a point number in 26 bits (and gather kernels run
    // 4 lanes per point on 32-bit thread indices)
    if (gm->n > kMaxPoints) return fail(MFCFD_EINVAL, "%d points: at most %d per context", gm->n, kMaxPoints);
    const int n = gm->n;
    const double* dp[] = {gm->x, gm->y, gm->z, gm->tan1, gm->tan2, gm->nor, gm->min_dist};
    for (auto* p : dp)
        if (!p) return fail(MFCFD_EINVAL, "null geometry array");
    if (!gm->status || !gm->nbhs || !gm->conn || !gm->alias || !gm->point_with_alias)
        return fail(MFCFD_EINVAL, "null index array");
    for (int d = 0; d < 6; ++d)
        if (!gm->split_nbhs[d] || !gm->split_conn[d]) return fail(MFCFD_EINVAL, "null split array %d", d);

    // ---- validation --------------------------------------------------
    for (int i = 0; i < n; ++i) {
        const int st = gm->status[i];
        if (st != 0 && st != 1 && st != 2 && st != 5 && st != 6)
            return fail(MFCFD_EINVAL, "point %d: unknown status %d", i + 1, st);
        if (gm->nbhs[i] < 0 || gm->nbhs[i] > kConnCols)
            return fail(MFCFD_EINVAL, "point %d: nbhs=%d outside 0..27", i + 1, gm->nbhs[i]);
        for (int j = 0; j < gm->nbhs[i]; ++j) {
            const int k = gm->conn[27LL * i + j];
            if (k < 1 || k > n) return fail(MFCFD_EINVAL, "point %d: conn(%d)=%d out of range", i + 1, j + 1, k);
        }
    }
    for (int d = 0; d < 6; ++d)
        for (int i = 0; i < n; ++i) {
            const HostList hl = host_list(gm, i, d);
            if (!hl.p) continue;
            if (hl.count < 0 || hl.count > kConnCols)
                return fail(MFCFD_EINVAL, "point %d: list %d has %d entries", i + 1, d, hl.count);
            if (hl.count == 0)
                return fail(MFCFD_EINVAL,
                            "point %d: list %d, which its flux walks, is empty; the reference would divide "
                            "by the zero determinant of its least-squares matrix",
                            i + 1, d);
            if (hl.p != gm->conn + 27LL * i && 22LL * i + hl.count > 22LL * n)
                return fail(MFCFD_EINVAL,
                            "point %d: split list %d has %d entries and runs past the (22,N) array "
                            "(the reference writes and reads it out of bounds)",
                            i + 1, d, hl.count);
            for (int j = 0; j < hl.count; ++j)
                if (hl.p[j] < 1 || hl.p[j] > n)
                    return fail(MFCFD_EINVAL, "point %d: list %d entry %d = %d out of range", i + 1, d, j + 1,
                                hl.p[j]);
        }
    for (int i = 0; i < n; ++i) {
        const int p = gm->point_with_alias[i];
        if (p < 0 || p > n) return fail(MFCFD_EINVAL, "point_with_alias(%d)=%d out of range", i + 1, p);
    }

    // ---- device order ------------------------------------------------
    std::vector<int> perm = morton_order(n, gm->x, gm->y, gm->z, inst_of);  // device -> host
    std::vector<int> inv(n);
    for (int d = 0; d < n; ++d) inv[perm[d]] = d;

    std::vector<double4> pos4(n);
    std::vector<double> fr(9ULL * n);
    std::vector<int> status(n), nbhs(n), conn0(27ULL * n, -1), alias(n);
    for (int d = 0; d < n; ++d) {
        const int r = perm[d];
        pos4[d] = make_double4(gm->x[r], gm->y[r], gm->z[r], gm->min_dist[r]);
        for (int cc = 0; cc < 3; ++cc) {
            fr[9ULL * d + cc] = gm->tan1[3ULL * r + cc];
            fr[9ULL * d + 3 + cc] = gm->tan2[3ULL * r + cc];
            fr[9ULL * d + 6 + cc] = gm->nor[3ULL * r + cc];
        }
        status[d] = gm->status[r];
        nbhs[d] = gm->nbhs[r];
        for (int j = 0; j < nbhs[d]; ++j) conn0[27ULL * d + j] = inv[gm->conn[27LL * r + j] - 1];
        alias[d] = gm->alias[r];
    }

    // direction lists, masks and regularity
    std::vector<int> cnt(6ULL * n, 0);
    std::vector<long long> off(6ULL * n, 0);
    std::vector<int> lst;
    lst.reserve((size_t)n * 76);
    // The fused flux kernel visits each point's walk once: conn, with a bit
    // per direction list an entry belongs to, then every list entry that does
    // not continue an in-order match against conn (the entries the 22-column
    // spill carries over from the neighbouring point), each with its own
    // list's bit.  Every list is then exactly the walk entries carrying its
    // bit, duplicates included.  A point whose walk would exceed kWalkCols
    // entries takes the list-by-list kernel.  Interior points whose walk is
    // conn with each entry in exactly one list per axis take the fused
    // kernel's interior class; the others, with the wall and outer points,
    // its general class.
    std::vector<unsigned> walk((size_t)kWalkCols * n, 0);
    std::vector<int> nwalk(n, 0);
    std::vector<unsigned char> mask((size_t)kWalkCols * n, 0);
    std::vector<unsigned char> walkable(n, 0);
    c->n_regular = 0;
    for (int d = 0; d < n; ++d) {
        const int r = perm[d];
        const int st = status[d];
        bool fits = (st == 0 || st == 1 || st == 2);
        bool reg = true;
        const int nb = nbhs[d];
        const int* cr = gm->conn + 27LL * r;
        unsigned* wk = walk.data() + (size_t)kWalkCols * d;
        unsigned char* mk = mask.data() + (size_t)kWalkCols * d;
        for (int j = 0; j < nb; ++j) wk[j] = (unsigned)conn0[27ULL * d + j];
        int nw = nb;
        for (int dir = 0; dir < 6; ++dir) {
            const HostList hl = host_list(gm, r, dir);
            off[(size_t)dir * n + d] = (long long)lst.size();
            cnt[(size_t)dir * n + d] = hl.p ? hl.count : 0;
            if (!hl.p) continue;
            int j = 0;
            for (int e = 0; e < hl.count; ++e) {
                const int k1 = hl.p[e];
                lst.push_back(inv[k1 - 1]);
                while (j < nb && cr[j] != k1) ++j;
                if (j < nb) {
                    mk[j] |= (unsigned char)(1u << dir);
                    ++j;
                } else if (nw < kWalkCols) {
                    reg = false;
                    wk[nw] = (unsigned)inv[k1 - 1];
                    mk[nw] = (unsigned char)(1u << dir);
                    ++nw;
                } else {
                    reg = false;
                    fits = false;
                }
            }
        }
        bool simple = st == 0 && reg;
        for (int j = 0; j < nb && simple; ++j)
            for (int a = 0; a < 3; ++a) {
                const unsigned ab = (mk[j] >> (2 * a)) & 3u;
                if (!(ab == 1u || ab == 2u)) simple = false;
            }
        // entries packed, then sorted by neighbour number: the four lanes of a
        // point and the points of a wave then step through their neighbours
        // in the same index order (measured: 3-5% faster gathers than conn
        // order; only the order of the point's terms changes)
        for (int j = 0; j < nw; ++j) wk[j] = (wk[j] << kWalkMaskBits) | mk[j];
        std::sort(wk, wk + nw);
        nwalk[d] = nw;
        walkable[d] = fits ? (simple ? 2 : 1) : 0;
        c->n_regular += reg ? 1 : 0;
    }

    // class lists (device numbers, device order) and the residual order
    std::vector<int> li, lw, lo, lsi, lso, lfi, lfw, lgen;
    for (int d = 0; d < n; ++d) {
        switch (status[d]) {
            case 0: li.push_back(d); (walkable[d] == 2 ? lfi : walkable[d] ? lfw : lgen).push_back(d); break;
            case 1: lw.push_back(d); (walkable[d] ? lfw : lgen).push_back(d); break;
            case 2: lo.push_back(d); (walkable[d] ? lfw : lgen).push_back(d); break;
            case 5: lsi.push_back(d); break;
            case 6: lso.push_back(d); break;
        }
    }
    // the general class grouped by status -- wall, outer, then the interior
    // points the spill altered -- within windows of kGroupWindow points, each
    // group in device order: a wave's points then take the same flux path
    // (octant or half-range lists) instead of both, while every XCD's
    // contiguous share of the list keeps the same mix of classes (grouping
    // the whole list gave the octant-heavy wall points to a few XCDs)
    {
        constexpr size_t kGroupWindow = 4096;
        const auto rank = [&](int d) { return status[d] == 1 ? 0 : status[d] == 2 ? 1 : 2; };
        for (size_t w0 = 0; w0 < lfw.size(); w0 += kGroupWindow) {
            const size_t w1 = std::min(lfw.size(), w0 + kGroupWindow);
            std::stable_sort(lfw.begin() + w0, lfw.begin() + w1, [&](int a, int b) { return rank(a) < rank(b); });
        }
    }
    c->n_fused_interior_pts = (int)lfi.size();
    c->n_fused_general_pts = (int)lfw.size();
    // the reference visits wall points, then interior points, each in host order
    std::vector<int> respos(n, 0x7fffffff);
    {
        int pos_w = 0, pos_i = 0, nw = 0;
        for (int r = 0; r < n; ++r) nw += gm->status[r] == 1;
        for (int r = 0; r < n; ++r) {
            if (gm->status[r] == 1) respos[inv[r]] = pos_w++;
            else if (gm->status[r] == 0) respos[inv[r]] = nw + pos_i++;
        }
        // supersonic points are walked in host order
        std::sort(lsi.begin(), lsi.end(), [&](int a, int b) { return perm[a] < perm[b]; });
        std::sort(lso.begin(), lso.end(), [&](int a, int b) { return perm[a] < perm[b]; });
    }

    c->release_all();
    c->have_geom = false;
    c->n = n;
    hipStream_t st = c->stream;
    if ((rc = upload(c->pos, pos4.data(), pos4.size(), st))) return rc;
    if ((rc = upload(c->frame, fr.data(), fr.size(), st))) return rc;
    if ((rc = upload(c->status, status.data(), status.size(), st))) return rc;
    if ((rc = upload(c->nbhs, nbhs.data(), nbhs.size(), st))) return rc;
    if ((rc = upload(c->conn, conn0.data(), conn0.size(), st))) return rc;
    if ((rc = upload(c->dir_cnt, cnt.data(), cnt.size(), st))) return rc;
    if ((rc = upload(c->dir_off, off.data(), off.size(), st))) return rc;
    if ((rc = upload(c->lst, lst.data(), lst.size(), st))) return rc;
    if ((rc = upload(c->walk, walk.data(), walk.size(), st))) return rc;
    if ((rc = upload(c->nwalk, nwalk.data(), nwalk.size(), st))) return rc;
    if ((rc = upload(c->alias, alias.data(), alias.size(), st))) return rc;
    if ((rc = upload(c->respos, respos.data(), respos.size(), st))) return rc;
    if ((rc = upload(c->perm, perm.data(), perm.size(), st))) return rc;
    if ((rc = upload(c->idx_interior, li.data(), li.size(), st))) return rc;
    if ((rc = upload(c->idx_wall, lw.data(), lw.size(), st))) return rc;
    if ((rc = upload(c->idx_outer, lo.data(), lo.size(), st))) return rc;
    if ((rc = upload(c->idx_sup_in, lsi.data(), lsi.size(), st))) return rc;
    if ((rc = upload(c->idx_sup_out, lso.data(), lso.size(), st))) return rc;
    if ((rc = upload(c->idx_fused_interior, lfi.data(), lfi.size(), st))) return rc;
    if ((rc = upload(c->idx_fused_wall, lfw.data(), lfw.size(), st))) return rc;
    if ((rc = upload(c->idx_generic, lgen.data(), lgen.size(), st))) return rc;
    c->tl.n_interior = (int)li.size();
    c->tl.n_wall = (int)lw.size();
    c->tl.n_outer = (int)lo.size();
    c->tl.n_sup_in = (int)lsi.size();
    c->tl.n_sup_out = (int)lso.size();
    c->tl.interior = c->idx_interior.p;
    c->tl.wall = c->idx_wall.p;
    c->tl.outer = c->idx_outer.p;
    c->tl.sup_in = c->idx_sup_in.p;
    c->tl.sup_out = c->idx_sup_out.p;
    c->host_class.assign(n, 2);
    for (int d : lfi) c->host_class[perm[d]] = 0;
    for (int d : lfw) c->host_class[perm[d]] = 1;
    c->tl.n_fused_interior = (int)lfi.size();
    c->tl.n_fused_wall = (int)lfw.size();
    c->tl.n_generic = (int)lgen.size();
    c->tl.fused_interior = c->idx_fused_interior.p;
    c->tl.fused_wall = c->idx_fused_wall.p;
    c->tl.generic = c->idx_generic.p;

    if (c->coef.alloc(18ULL * n) != hipSuccess) return fail(MFCFD_ENOMEM, "stencil coefficients");
    launch_stencil_coefficients(c->geom(), c->coef.p, c->fp.power, st);
    if (c->dgeo.alloc(12ULL * n) != hipSuccess) return fail(MFCFD_ENOMEM, "D_inv geometry");
    launch_dinv_geometry(c->geom(), c->dgeo.p, c->fp.power, st);
    if ((rc = launch_error("stencil coefficients"))) return rc;
    {
        // singular least-squares stencils: the reference divides by their
        // zero determinant; refused here instead
        if (c->stencil_bad.alloc(1) != hipSuccess) return fail(MFCFD_ENOMEM, "stencil check");
        const int none = 0x7fffffff;
        int bad = none;
        HIP_TRY(hipMemcpyAsync(c->stencil_bad.p, &none, sizeof(int), hipMemcpyHostToDevice, st));
        launch_stencil_check(c->geom(), c->fp.power, c->stencil_bad.p, st);
        if ((rc = launch_error("stencil check"))) return rc;
        HIP_TRY(hipMemcpyAsync(&bad, c->stencil_bad.p, sizeof(int), hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        if (bad != none) {
            static const char* names[7] = {"xpos", "xneg", "ypos", "yneg", "zpos", "zneg", "conn"};
            const int pt = perm[bad / 8] + 1;
            c->release_all();
            return fail(MFCFD_EINVAL,
                        "point %d: the least-squares matrix of its %s stencil is singular (coplanar or collinear "
                        "neighbours); the reference would divide by its zero determinant",
                        pt, names[bad % 8]);
        }
    }

    // flow state (zero until uploaded; initial_conditions.f90:23-24 zero delUp/delUn)
    DevBuf<double>* st5[] = {&c->prim, &c->U, &c->q, &c->flux_res, &c->delUp, &c->delUn};
    for (auto* b : st5) {
        if (b->alloc(5ULL * n) != hipSuccess) return fail(MFCFD_ENOMEM, "state arrays");
        HIP_TRY(hipMemsetAsync(b->p, 0, 5ULL * n * sizeof(double), st));
    }
    if (c->qm.alloc(10ULL * n) != hipSuccess || c->dq.alloc(15ULL * n) != hipSuccess ||
        c->delt.alloc((size_t)n) != hipSuccess || c->tsden.alloc((size_t)n) != hipSuccess ||
        c->dinv.alloc((size_t)n) != hipSuccess ||
        c->staging.alloc(15ULL * n) != hipSuccess)
        return fail(MFCFD_ENOMEM, "state arrays");
    HIP_TRY(hipMemsetAsync(c->qm.p, 0, 10ULL * n * sizeof(double), st));
    HIP_TRY(hipMemsetAsync(c->dq.p, 0, 15ULL * n * sizeof(double), st));
    if (c->q32.alloc((size_t)kQ32Stride * n) != hipSuccess || c->qm32.alloc(10ULL * n) != hipSuccess ||
        c->dq32.alloc(15ULL * n) != hipSuccess)
        return fail(MFCFD_ENOMEM, "FP32 q arrays");
    HIP_TRY(hipMemsetAsync(c->q32.p, 0, (size_t)kQ32Stride * n * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(c->qm32.p, 0, 10ULL * n * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(c->dq32.p, 0, 15ULL * n * sizeof(float), st));
    HIP_TRY(hipMemsetAsync(c->delt.p, 0, (size_t)n * sizeof(double), st));
    HIP_TRY(hipMemsetAsync(c->dinv.p, 0, (size_t)n * sizeof(double), st));
    if (c->fp.inner_iterations > 0 && c->temp.alloc(15ULL * n) != hipSuccess) return fail(MFCFD_ENOMEM, "temp");
    c->n_partial = state_update_blocks(n);
    if (c->partial.alloc(2ULL * c->n_partial) != hipSuccess || c->partial_pos.alloc(2ULL * c->n_partial) != hipSuccess)
        return fail(MFCFD_ENOMEM, "residual partials");

    // host inputs of the LU-SGS sweep plans (device order); a plan is built
    // for the iteration form in use when the first implicit sweep needs it
    c->h_status = status;
    c->h_alias = alias;
    c->h_nbhs = nbhs;
    c->h_conn = conn0;
    c->h_cnt = cnt;
    c->h_off = off;
    c->h_lst = lst;
    c->h_x.resize(n);
    c->h_y.resize(n);
    c->h_z.resize(n);
    for (int d = 0; d < n; ++d) {
        c->h_x[d] = pos4[d].x;
        c->h_y[d] = pos4[d].y;
        c->h_z[d] = pos4[d].z;
    }
    c->h_pwa.assign(n, -1);
    for (int i = 0; i < n; ++i) {
        const int p = gm->point_with_alias[i];
        if (p > 0) c->h_pwa[i] = inv[p - 1];
    }
    {
        // class lists in host index order (point_preprocessor.f90:41-95)
        c->h_wall.clear();
        c->h_int.clear();
        c->h_out.clear();
        c->h_sin.clear();
        c->h_sout.clear();
        for (int r = 0; r < n; ++r) {
            switch (gm->status[r]) {
                case 0: c->h_int.push_back(inv[r]); break;
                case 1: c->h_wall.push_back(inv[r]); break;
                case 2: c->h_out.push_back(inv[r]); break;
                case 5: c->h_sin.push_back(inv[r]); break;
                case 6: c->h_sout.push_back(inv[r]); break;
            }
        }
    }
    HIP_TRY(hipMemsetAsync(c->rec.p, 0, sizeof(ResidualRecord), st));
    // instances: device-order ranges (one for a single cloud), per-point
    // instance and freestream table for a batch
    {
        c->batch = inst_of != nullptr;
        c->n_inst = c->batch ? n_inst : 1;
        std::vector<int> beg(c->n_inst, n), end(c->n_inst, 0), id(n, 0);
        for (int d = 0; d < n; ++d) {
            const int b = c->batch ? inst_of[perm[d]] : 0;
            id[d] = b;
            beg[b] = std::min(beg[b], d);
            end[b] = std::max(end[b], d + 1);
        }
        c->inst_count.assign(c->n_inst, 0);
        std::vector<double> norm(c->n_inst);
        for (int b = 0; b < c->n_inst; ++b) {
            if (end[b] <= beg[b]) beg[b] = end[b] = 0;
            c->inst_count[b] = end[b] - beg[b];
            norm[b] = (double)c->inst_count[b];
        }
        if ((rc = upload(c->inst_beg, beg.data(), beg.size(), st))) return rc;
        if ((rc = upload(c->inst_end, end.data(), end.size(), st))) return rc;
        if ((rc = upload(c->inst_norm, norm.data(), norm.size(), st))) return rc;
        // residual chunks: each cloud's range cut into kResidualChunk points
        std::vector<int> cb, ce, cp(c->n_inst + 1, 0);
        for (int b = 0; b < c->n_inst; ++b) {
            for (int a0 = beg[b]; a0 < end[b]; a0 += kResidualChunk) {
                cb.push_back(a0);
                ce.push_back(std::min(end[b], a0 + kResidualChunk));
            }
            cp[b + 1] = (int)cb.size();
        }
        c->n_chunks = (int)cb.size();
        if (cb.empty()) {   // keep the buffers allocated
            cb.push_back(0);
            ce.push_back(0);
        }
        if ((rc = upload(c->chunk_beg, cb.data(), cb.size(), st))) return rc;
        if ((rc = upload(c->chunk_end, ce.data(), ce.size(), st))) return rc;
        if ((rc = upload(c->chunk_ptr, cp.data(), cp.size(), st))) return rc;
        if (c->chunk_sum.alloc(cb.size()) != hipSuccess) return fail(MFCFD_ENOMEM, "residual chunk sums");
        if (c->batch) {
            if ((rc = upload(c->inst_id, id.data(), id.size(), st))) return rc;
            if ((rc = upload(c->inst_fs, fs, 2ULL * n_inst, st))) return rc;
        }
        if (c->ressq.alloc((size_t)n) != hipSuccess || c->hist_ctr.alloc(1) != hipSuccess)
            return fail(MFCFD_ENOMEM, "instance residual buffers");
        HIP_TRY(hipMemsetAsync(c->ressq.p, 0, (size_t)n * sizeof(double), st));
        HIP_TRY(hipMemsetAsync(c->hist_ctr.p, 0, sizeof(int), st));
        // the pageable host vectors of this block end here: their copies must
        // have completed
        HIP_TRY(hipStreamSynchronize(st));
    }
    HIP_TRY(hipStreamSynchronize(st));
    c->have_geom = true;
    c->graph_stale = true;
    c->res_old = c->res_new = c->residue = 0.0;
    c->last_t = 0;
    return MFCFD_OK;
}

}  // namespace

int mfcfd_bind_geometry(mfcfd_ctx* c, const mfcfd_geometry* gm) {
    int rc = check_ctx(c, false);
    if (rc) return rc;
    return bind_impl(c, gm, nullptr, 1, nullptr);
}

int mfcfd_bind_batch(mfcfd_ctx* c, int n_inst, const mfcfd_geometry* geoms, const double* mach,
                     const double* aoa) {
    int rc = check_ctx(c, false);
    if (rc) return rc;
    if (n_inst <= 0 || !geoms || !mach || !aoa) return fail(MFCFD_EINVAL, "empty batch");
    long long total = 0;
    for (int b = 0; b < n_inst; ++b) {
        const mfcfd_geometry* g = geoms + b;
        if (g->n <= 0 || !g->x || !g->y || !g->z || !g->tan1 || !g->tan2 || !g->nor || !g->min_dist ||
            !g->status || !g->nbhs || !g->conn || !g->alias || !g->point_with_alias)
            return fail(MFCFD_EINVAL, "instance %d: null or empty geometry", b);
        for (int d = 0; d < 6; ++d)
            if (!g->split_nbhs[d] || !g->split_conn[d]) return fail(MFCFD_EINVAL, "instance %d: null split", b);
        for (int i = 0; i < g->n; ++i) {
            if (g->status[i] == 5 || g->status[i] == 6)
                return fail(MFCFD_EUNSUPPORTED, "instance %d: supersonic points in a batch", b);
            // a list that spills past the instance's last point would land in
            // the next instance's columns
            for (int d = 0; d < 6; ++d) {
                const HostList hl = host_list(g, i, d);
                if (hl.p && hl.p != g->conn + 27LL * i && 22LL * i + hl.count > 22LL * g->n)
                    return fail(MFCFD_EINVAL, "instance %d point %d: split list %d runs past the instance", b,
                                i + 1, d);
            }
        }
        total += g->n;
    }
    if (total > 0x7fffffffLL / 27) return fail(MFCFD_EINVAL, "batch too large");
    const int n = (int)total;
    std::vector<double> x(n), y(n), z(n), t1(3ULL * n), t2(3ULL * n), nr(3ULL * n), md(n);
    std::vector<int> status(n), nbhs(n), conn(27ULL * n, 0), alias(n), pwa(n), inst(n);
    std::vector<int> scnt[6], sconn[6];
    for (int d = 0; d < 6; ++d) {
        scnt[d].assign(n, 0);
        sconn[d].assign(22ULL * n, 0);
    }
    std::vector<double> fs(2ULL * n_inst);
    const double pi = 4.0 * std::atan(1.0);
    long long off = 0;
    for (int b = 0; b < n_inst; ++b) {
        const mfcfd_geometry* g = geoms + b;
        const int m = g->n;
        const int o = (int)off;
        auto shift = [o](int v) { return v > 0 ? v + o : 0; };
        for (int i = 0; i < m; ++i) {
            x[o + i] = g->x[i];
            y[o + i] = g->y[i];
            z[o + i] = g->z[i];
            md[o + i] = g->min_dist[i];
            for (int cc = 0; cc < 3; ++cc) {
                t1[3LL * (o + i) + cc] = g->tan1[3LL * i + cc];
                t2[3LL * (o + i) + cc] = g->tan2[3LL * i + cc];
                nr[3LL * (o + i) + cc] = g->nor[3LL * i + cc];
            }
            status[o + i] = g->status[i];
            nbhs[o + i] = g->nbhs[i];
            for (int j = 0; j < 27; ++j) conn[27LL * (o + i) + j] = shift(g->conn[27LL * i + j]);
            alias[o + i] = g->alias[i] > 0 ? g->alias[i] + o : 0;
            pwa[o + i] = shift(g->point_with_alias[i]);
            inst[o + i] = b;
            for (int d = 0; d < 6; ++d) {
                scnt[d][o + i] = g->split_nbhs[d][i];
                for (int j = 0; j < 22; ++j) sconn[d][22LL * (o + i) + j] = shift(g->split_conn[d][22LL * i + j]);
            }
        }
        // parameter.f90:29-31 with the instance's Mach and angle of attack
        const double theta = aoa[b] * pi / 180.0;
        fs[2 * b] = mach[b] * std::cos(theta);
        fs[2 * b + 1] = mach[b] * std::sin(theta);
        off += m;
    }
    mfcfd_geometry mg;
    mg.n = n;
    mg.x = x.data();
    mg.y = y.data();
    mg.z = z.data();
    mg.tan1 = t1.data();
    mg.tan2 = t2.data();
    mg.nor = nr.data();
    mg.min_dist = md.data();
    mg.status = status.data();
    mg.nbhs = nbhs.data();
    mg.conn = conn.data();
    for (int d = 0; d < 6; ++d) {
        mg.split_nbhs[d] = scnt[d].data();
        mg.split_conn[d] = sconn[d].data();
    }
    mg.alias = alias.data();
    mg.point_with_alias = pwa.data();
    return bind_impl(c, &mg, inst.data(), n_inst, fs.data());
}

int mfcfd_batch_info(mfcfd_ctx* c, int* n_inst, int* counts) {
    if (!c || !c->have_geom) return fail(MFCFD_ESTATE, "no geometry bound");
    if (n_inst) *n_inst = c->n_inst;
    if (counts)
        for (int b = 0; b < c->n_inst; ++b) counts[b] = c->inst_count[b];
    return MFCFD_OK;
}

int mfcfd_bind_geometry_arrays(mfcfd_ctx* c, int n, const double* x, const double* y, const double* z,
                               const double* tan1, const double* tan2, const double* nor,
                               const double* min_dist, const int* status, const int* nbhs, const int* conn,
                               const int* xpos_nbhs, const int* xneg_nbhs, const int* ypos_nbhs,
                               const int* yneg_nbhs, const int* zpos_nbhs, const int* zneg_nbhs,
                               const int* xpos_conn, const int* xneg_conn, const int* ypos_conn,
                               const int* yneg_conn, const int* zpos_conn, const int* zneg_conn,
                               const int* alias, const int* point_with_alias) {
    mfcfd_geometry g;
    g.n = n;
    g.x = x;
    g.y = y;
    g.z = z;
    g.tan1 = tan1;
    g.tan2 = tan2;
    g.nor = nor;
    g.min_dist = min_dist;
    g.status = status;
    g.nbhs = nbhs;
    g.conn = conn;
    const int* cnts[6] = {xpos_nbhs, xneg_nbhs, ypos_nbhs, yneg_nbhs, zpos_nbhs, zneg_nbhs};
    const int* lists[6] = {xpos_conn, xneg_conn, ypos_conn, yneg_conn, zpos_conn, zneg_conn};
    for (int d = 0; d < 6; ++d) {
        g.split_nbhs[d] = cnts[d];
        g.split_conn[d] = lists[d];
    }
    g.alias = alias;
    g.point_with_alias = point_with_alias;
    return mfcfd_bind_geometry(c, &g);
}

int mfcfd_upload(mfcfd_ctx* c, int id, const double* host) {
    int rc = check_ctx(c, true);
    if (rc) return rc;
    ArrayDesc a = array_of(c, id);
    if (!a.buf || !host) return fail(MFCFD_EINVAL, "bad array id %d or null pointer", id);
    const size_t bytes = (size_t)a.per_point * c->n * sizeof(double);
    if (id == MFCFD_ARR_PRIM) {
        // a state the iteration cannot start from: q_variables.f90 takes the
        // log of density and pressure, and NaN spreads through every stencil
        for (long long r = 0; r < c->n; ++r) {
            const double* p = host + 5 * r;
            bool ok = p[0] > 0.0 && p[4] > 0.0;
            for (int v = 0; v < 5; ++v) ok = ok && std::isfinite(p[v]);
            if (!ok)
                return fail(MFCFD_EINVAL,
                            "prim of point %lld = (%g, %g, %g, %g, %g): not finite with positive density and "
                            "pressure",
                            r + 1, p[0], p[1], p[2], p[3], p[4]);
        }
    }
    HIP_TRY(hipMemcpyAsync(c->staging.p, host, bytes, hipMemcpyHostToDevice, c->stream));
    launch_gather(a.buf->p, c->staging.p, c->perm.p, c->n, a.per_point, c->stream);
    // mixed form: the double array is staging for the FP32 copy the kernels read
    if (float* f32 = c->fp.mixed ? mixed_copy_of(c, id) : nullptr)
        launch_to_float(f32, id == MFCFD_ARR_Q ? kQ32Stride : a.per_point, a.buf->p, a.per_point, c->n, c->stream);
    if ((rc = launch_error("upload permutation"))) return rc;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MFCFD_OK;
}

int mfcfd_download(mfcfd_ctx* c, int id, double* host) {
    int rc = check_ctx(c, true);
    if (rc) return rc;
    ArrayDesc a = array_of(c, id);
    if (!a.buf || !host) return fail(MFCFD_EINVAL, "bad array id %d or null pointer", id);
    const size_t bytes = (size_t)a.per_point * c->n * sizeof(double);
    if (float* f32 = c->fp.mixed ? mixed_copy_of(c, id) : nullptr)
        launch_to_double(a.buf->p, f32, id == MFCFD_ARR_Q ? kQ32Stride : a.per_point, a.per_point, c->n, c->stream);
    launch_scatter(c->staging.p, a.buf->p, c->perm.p, c->n, a.per_point, c->stream);
    if ((rc = launch_error("download permutation"))) return rc;
    HIP_TRY(hipMemcpyAsync(host, c->staging.p, bytes, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MFCFD_OK;
}

int mfcfd_compute_conserved_vector(mfcfd_ctx* c) {
    int rc = check_ctx(c, true);
    if (rc) return rc;
    launch_conserved_vector(c->geom(), c->state(), c->stream);
    return launch_error("compute_conserved_vector");
}

int mfcfd_eval_q_variables(mfcfd_ctx* c) {
    int rc = check_ctx(c, true);
    if (rc) return rc;
    launch_q_variables(c->geom(), c->state(), c->fp, c->stream);
    return launch_error("eval_q_variables");
}

int mfcfd_eval_q_derivatives(mfcfd_ctx* c) {
    int rc = check_ctx(c, true);
    if (rc) return rc;
    launch_q_derivatives(c->geom(), c->state(), c->fp, c->temp.p, c->stream);
    return launch_error("eval_q_derivatives");
}

int mfcfd_timestep_delt(mfcfd_ctx* c) {
    int rc = check_ctx(c, true);
    if (rc) return rc;
    launch_timestep(c->geom(), c->state(), c->fp, c->stream);
    return launch_error("timestep_delt");
}

int mfcfd_eval_flux_residual(mfcfd_ctx* c) {
    int rc = check_ctx(c, true);
    if (rc) return rc;
    flux(c, kFluxAll);
    return launch_error("eval_flux_residual");
}

int mfcfd_implicit_forward_backward_sweeps(mfcfd_ctx* c) {
    int rc = check_ctx(c, true);
    if (rc) return rc;
    if ((rc = prepare_plan(c))) return rc;
    return do_sweeps(c);
}

int mfcfd_state_update(mfcfd_ctx* c) {
    int rc = check_ctx(c, true);
    if (rc) return rc;
    if ((rc = do_state_update(c, nullptr, 0))) return rc;
    return fetch_record(c);
}

int mfcfd_fpi_solver(mfcfd_ctx* c, int t, double* res_new, double* residue) {
    int rc = check_ctx(c, true);
    if (rc) return rc;
    if ((rc = prepare_plan(c))) return rc;
    if ((rc = do_iteration(c, nullptr, 0))) return rc;
    if ((rc = fetch_record(c))) return rc;
    residue_bookkeeping(c, t, c->res_new);
    if (res_new) *res_new = c->res_new;
    if (residue) *residue = c->residue;
    return MFCFD_OK;
}

int mfcfd_run(mfcfd_ctx* c, int t0, int n_iters, double* res_new_hist, double* residue_hist) {
    int rc = check_ctx(c, true);
    if (rc) return rc;
    if (n_iters < 0) return fail(MFCFD_EINVAL, "n_iters < 0");
    if (n_iters == 0) return MFCFD_OK;
    if ((rc = prepare_plan(c))) return rc;
    if (c->hist.n < (size_t)n_iters) {
        if (c->hist.alloc((size_t)n_iters) != hipSuccess) return fail(MFCFD_ENOMEM, "history buffer");
        c->graph_stale = true;
    }
    for (int it = 0; it < n_iters; ++it)
        if ((rc = do_iteration(c, c->hist.p, it))) return rc;
    std::vector<double> h(n_iters);
    HIP_TRY(hipMemcpyAsync(h.data(), c->hist.p, sizeof(double) * n_iters, hipMemcpyDeviceToHost, c->stream));
    if ((rc = fetch_record(c))) return rc;
    for (int it = 0; it < n_iters; ++it) {
        residue_bookkeeping(c, t0 + it, h[it]);
        if (res_new_hist) res_new_hist[it] = c->res_new;
        if (residue_hist) residue_hist[it] = c->residue;
    }
    return MFCFD_OK;
}

int mfcfd_run_noalloc(mfcfd_ctx* c, int t0, int n_iters, double* res_new_hist, double* residue_hist) {
    int rc = check_ctx(c, true);
    if (rc) return rc;
    if (n_iters < 0) return fail(MFCFD_EINVAL, "n_iters < 0");
    if (n_iters > MFCFD_NOALLOC_MAX_ITERS)
        return fail(MFCFD_EINVAL, "n_iters %d exceeds the no-allocation history of %d iterations", n_iters,
                    MFCFD_NOALLOC_MAX_ITERS);
    if (n_iters == 0) return MFCFD_OK;
    if (c->time_kernel) return fail(MFCFD_ESTATE, "kernel timing creates events; disable it before mfcfd_run_noalloc");
    if (c->params.mode != MFCFD_EXPLICIT && c->plan_form != c->params.mode)
        return fail(MFCFD_ESTATE, "the sweep plan of this iteration form is not built (building it allocates): "
                                  "call mfcfd_sweep_info, mfcfd_fpi_solver or mfcfd_run once first");
    double* hist = g_scratch[c->device];
    if (!hist) return fail(MFCFD_ESTATE, "no scratch address for device %d", c->device);
    for (int it = 0; it < n_iters; ++it)
        if ((rc = do_iteration(c, hist, it))) return rc;
    if ((rc = fetch_record(c))) return rc;
    // the history comes back through a fixed host buffer, a block at a time
    double h[256];
    for (int off = 0; off < n_iters; off += 256) {
        const int m = std::min(256, n_iters - off);
        HIP_TRY(hipMemcpyAsync(h, hist + off, sizeof(double) * m, hipMemcpyDeviceToHost, c->stream));
        HIP_TRY(hipStreamSynchronize(c->stream));
        for (int j = 0; j < m; ++j) {
            residue_bookkeeping(c, t0 + off + j, h[j]);
            if (res_new_hist) res_new_hist[off + j] = c->res_new;
            if (residue_hist) residue_hist[off + j] = c->residue;
        }
    }
    return MFCFD_OK;
}

int mfcfd_run_graph(mfcfd_ctx* c, int t0, int n_iters, double* res_new_hist, double* inst_res_hist) {
    int rc = check_ctx(c, true);
    if (rc) return rc;
    if (n_iters < 0) return fail(MFCFD_EINVAL, "n_iters < 0");
    if (n_iters == 0) return MFCFD_OK;
    if (c->time_kernel) return fail(MFCFD_ESTATE, "kernel timing records events; disable it before graph runs");
    // plan building allocates and copies: never inside the capture
    if ((rc = prepare_plan(c))) return rc;
    // the history buffers' addresses are captured: grow them before capture
    const size_t need = (size_t)std::max(n_iters, 64);
    if (c->hist.n < need) {
        if (c->hist.alloc(2 * need) != hipSuccess) return fail(MFCFD_ENOMEM, "history buffer");
        c->graph_stale = true;
    }
    if (c->inst_hist.n < c->hist.n * c->n_inst) {
        if (c->inst_hist.alloc(c->hist.n * c->n_inst) != hipSuccess)
            return fail(MFCFD_ENOMEM, "instance history buffer");
        c->graph_stale = true;
    }
    InstanceResidual ir;
    ir.ressq = c->ressq.p;
    ir.beg = c->inst_beg.p;
    ir.end = c->inst_end.p;
    ir.norm = c->inst_norm.p;
    ir.n_inst = c->n_inst;
    ir.inst_hist = c->inst_hist.p;
    ir.counter = c->hist_ctr.p;
    ir.chunk_beg = c->chunk_beg.p;
    ir.chunk_end = c->chunk_end.p;
    ir.chunk_ptr = c->chunk_ptr.p;
    ir.n_chunks = c->n_chunks;
    ir.chunk_sum = c->chunk_sum.p;
    if (c->graph_stale || !c->graph_exec) {
        if (c->graph_exec) {
            (void)hipGraphExecDestroy(c->graph_exec);
            c->graph_exec = nullptr;
        }
        hipGraph_t graph = nullptr;
        HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = do_iteration(c, c->hist.p, 0, &ir);
        const hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return fail(MFCFD_EHIP, "graph capture: %s", hipGetErrorString(e));
        const hipError_t ie = hipGraphInstantiate(&c->graph_exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ie != hipSuccess) return fail(MFCFD_EHIP, "graph instantiate: %s", hipGetErrorString(ie));
        c->graph_stale = false;
    }
    // history rows restart at 0 for this call (outside the graph)
    HIP_TRY(hipMemsetAsync(c->hist_ctr.p, 0, sizeof(int), c->stream));
    for (int it = 0; it < n_iters; ++it) HIP_TRY(hipGraphLaunch(c->graph_exec, c->stream));
    std::vector<double> h(n_iters);
    HIP_TRY(hipMemcpyAsync(h.data(), c->hist.p, sizeof(double) * n_iters, hipMemcpyDeviceToHost, c->stream));
    if (inst_res_hist)
        HIP_TRY(hipMemcpyAsync(inst_res_hist, c->inst_hist.p, sizeof(double) * n_iters * c->n_inst,
                               hipMemcpyDeviceToHost, c->stream));
    if ((rc = fetch_record(c))) return rc;
    for (int it = 0; it < n_iters; ++it) {
        residue_bookkeeping(c, t0 + it, h[it]);
        if (res_new_hist) res_new_hist[it] = c->res_new;
    }
    return MFCFD_OK;
}

}  // extern "C"

namespace {

// mean milliseconds of fn over `repeats` launches (after one untimed one),
// each preceded by pre(): the kernel that runs before it in an iteration,
// so every candidate meets the caches as the iteration leaves them
template <typename P, typename F>
int time_launches(mfcfd_ctx* c, int repeats, P pre, F fn, double* ms_out) {
    std::vector<hipEvent_t> ev(2 * (size_t)repeats, nullptr);
    int rc = 0;
    for (auto& e : ev)
        if (hipEventCreate(&e) != hipSuccess) {
            e = nullptr;
            rc = fail(MFCFD_EHIP, "hipEventCreate");
            break;
        }
    if (!rc) {
        pre();
        fn();
        for (int r = 0; r < repeats; ++r) {
            pre();
            hipEventRecord(ev[2 * r], c->stream);
            fn();
            hipEventRecord(ev[2 * r + 1], c->stream);
        }
        rc = launch_error("autotune");
        if (!rc && hipEventSynchronize(ev.back()) != hipSuccess) rc = fail(MFCFD_EHIP, "event synchronize");
    }
    double total = 0.0;
    for (int r = 0; r < repeats && !rc; ++r) {
        float ms = 0.f;
        hipEventElapsedTime(&ms, ev[2 * r], ev[2 * r + 1]);
        total += ms;
    }
    for (auto e : ev)
        if (e) (void)hipEventDestroy(e);
    *ms_out = total / repeats;
    return rc;
}

}  // namespace

extern "C" {

int mfcfd_autotune(mfcfd_ctx* c, int repeats, mfcfd_tuning* out) {
    int rc = check_ctx(c, true);
    if (rc) return rc;
    if (repeats <= 0) repeats = 3;
    const DevGeom g = c->geom();
    const DevState s = c->state();
    const int blocks[3] = {64, 128, 256};
    // LDS reservations that cap a kernel at the given waves per CU: fewer
    // resident points per XCD keep their neighbours' data in that XCD's L2
    int lds_cu = 0, lds_blk = 0;
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, c->device) != hipSuccess ||
        hipDeviceGetAttribute(&lds_blk, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device) != hipSuccess)
        lds_cu = lds_blk = 0;
    auto caps = [&](int block, std::initializer_list<int> waves_per_cu) {
        std::vector<int> out;
        for (int wcu : waves_per_cu) {
            const int per_cu = wcu * 64 / block;   // blocks per CU
            if (per_cu < 1 || per_cu * block != wcu * 64 || lds_cu <= 0) continue;
            const int lds = (lds_cu / per_cu) & ~255;
            if (lds <= 0 || lds > lds_blk || (per_cu + 1) * lds <= lds_cu) continue;
            out.push_back(lds);
        }
        return out;
    };
    mfcfd_tuning t{};
    if ((rc = prepare_plan(c))) return rc;
    // inputs of the timed kernels: the first steps of an iteration from prim
    launch_conserved_and_q(g, s, c->fp, c->stream);
    // predecessors in the iteration: the q-variables before the q-derivatives,
    // the q-derivatives (with the shape chosen for them) before the fluxes
    auto pre_q = [&] { launch_q_variables(g, s, c->fp, c->stream); };
    LaunchShapes best_for_pre;
    auto pre_f = [&] {
        if (c->fp.inner_iterations == 0) launch_qderiv_timestep(g, s, c->fp, c->stream, best_for_pre.qderiv);
    };
    LaunchShapes best = c->shapes;
    if (c->fp.inner_iterations == 0) {
        const int waves[3] = {2, 4, 6};
        double best_ms = 1e30;
        for (int bl : blocks)
            for (int w : waves) {
                KernelShape ks;
                ks.block = bl;
                ks.waves = w;
                double ms = 0.0;
                if ((rc = time_launches(c, repeats, pre_q, [&] { launch_qderiv_timestep(g, s, c->fp, c->stream, ks); }, &ms)))
                    return rc;
                if (bl == 256 && w == 4) t.qderiv_default_ms = ms;
                if (ms < best_ms) {
                    best_ms = ms;
                    best.qderiv = ks;
                }
            }
        for (int lds : caps(best.qderiv.block, {8, 12})) {
            KernelShape ks = best.qderiv;
            ks.lds = lds;
            double ms = 0.0;
            if ((rc = time_launches(c, repeats, pre_q, [&] { launch_qderiv_timestep(g, s, c->fp, c->stream, ks); }, &ms)))
                return rc;
            if (ms < best_ms) {
                best_ms = ms;
                best.qderiv = ks;
            }
        }
        t.qderiv_ms = best_ms;
        best_for_pre = best;
        launch_qderiv_timestep(g, s, c->fp, c->stream, best.qderiv);
    } else {
        launch_q_derivatives(g, s, c->fp, c->temp.p, c->stream);
        launch_timestep(g, s, c->fp, c->stream);
    }
    if (c->params.flux_order == MFCFD_FLUX_FUSED) {
        // FP32 variants also come capped at 5 and 6 waves per SIMD (a few
        // spilled registers for more resident waves)
        const std::vector<int> waves = c->fp.mixed ? std::vector<int>{1, 2, 3, 4, 5, 6} : std::vector<int>{1, 2, 3, 4};
        for (int part = 1; part <= 2; ++part) {
            const int count = part == 1 ? c->tl.n_fused_interior : c->tl.n_fused_wall;
            if (count == 0) continue;
            double best_ms = 1e30;
            KernelShape pick;
            for (int bl : blocks)
                for (int w : waves) {
                    LaunchShapes ls = best;
                    (part == 1 ? ls.flux_interior : ls.flux_wall) = KernelShape{bl, w};
                    double ms = 0.0;
                    if ((rc = time_launches(c, repeats, pre_f,
                                            [&] { launch_flux_residual(g, s, c->fp, c->tl, c->stream, part, ls); },
                                            &ms)))
                        return rc;
                    if (bl == 256 && w == 2) (part == 1 ? t.flux_interior_default_ms : t.flux_wall_default_ms) = ms;
                    if (ms < best_ms) {
                        best_ms = ms;
                        pick = KernelShape{bl, w};
                    }
                }
            for (int lds : caps(pick.block, {4, 6, 8, 12})) {
                LaunchShapes ls = best;
                KernelShape ks = pick;
                ks.lds = lds;
                (part == 1 ? ls.flux_interior : ls.flux_wall) = ks;
                double ms = 0.0;
                if ((rc = time_launches(c, repeats, pre_f,
                                        [&] { launch_flux_residual(g, s, c->fp, c->tl, c->stream, part, ls); }, &ms)))
                    return rc;
                if (ms < best_ms) {
                    best_ms = ms;
                    pick = ks;
                }
            }
            (part == 1 ? best.flux_interior : best.flux_wall) = pick;
            (part == 1 ? t.flux_interior_ms : t.flux_wall_ms) = best_ms;
        }
    }
    flux(c, kFluxAll);
    // the sweeps are a function of flux_res, prim, U and delt only (each
    // point reads its dependencies' values of the same sweep), so repeating
    // them leaves the state as one sweep would
    // (the alternative driver's old entries read values of the previous
    // repeat, so its sweeps are not tuned)
    if (c->params.mode == MFCFD_IMPLICIT && c->params.sweep_schedule == MFCFD_SWEEP_ASYNC) {
        const int divs[4] = {1, 2, 4, 8};
        double best_ms = 1e30;
        for (int dv : divs) {
            c->shapes.sweep_div = dv;
            double ms = 0.0;
            if ((rc = time_launches(c, repeats, [] {}, [&] { (void)do_sweeps(c); }, &ms))) return rc;
            if (dv == 1) t.sweeps_default_ms = ms;
            if (ms < best_ms) {
                best_ms = ms;
                best.sweep_div = dv;
            }
        }
        t.sweeps_ms = best_ms;
        if ((rc = fetch_record(c))) return rc;  // surfaces a sweep stall
    }
    c->shapes = best;
    c->graph_stale = true;
    HIP_TRY(hipStreamSynchronize(c->stream));
    t.flux_interior_block = best.flux_interior.block;
    t.flux_interior_waves = best.flux_interior.waves;
    t.flux_wall_block = best.flux_wall.block;
    t.flux_wall_waves = best.flux_wall.waves;
    t.qderiv_block = best.qderiv.block;
    t.qderiv_waves = best.qderiv.waves;
    t.sweep_div = best.sweep_div;
    t.flux_interior_lds = best.flux_interior.lds;
    t.flux_wall_lds = best.flux_wall.lds;
    t.qderiv_lds = best.qderiv.lds;
    if (out) *out = t;
    return MFCFD_OK;
}

int mfcfd_set_tuning(mfcfd_ctx* c, const mfcfd_tuning* in) {
    if (!c || !in) return fail(MFCFD_EINVAL, "null argument");
    auto ok_block = [](int b) { return b == 0 || b == 64 || b == 128 || b == 256; };
    if (!ok_block(in->flux_interior_block) || !ok_block(in->flux_wall_block) || !ok_block(in->qderiv_block))
        return fail(MFCFD_EINVAL, "block sizes are 64, 128 or 256 (0: default)");
    if (in->sweep_div < 0 || in->sweep_div > 64) return fail(MFCFD_EINVAL, "sweep_div %d", in->sweep_div);
    if (in->flux_interior_lds < 0 || in->flux_wall_lds < 0 || in->qderiv_lds < 0)
        return fail(MFCFD_EINVAL, "negative LDS reservation");
    c->shapes.flux_interior = KernelShape{in->flux_interior_block, in->flux_interior_waves, in->flux_interior_lds};
    c->shapes.flux_wall = KernelShape{in->flux_wall_block, in->flux_wall_waves, in->flux_wall_lds};
    c->shapes.qderiv = KernelShape{in->qderiv_block, in->qderiv_waves, in->qderiv_lds};
    c->shapes.sweep_div = in->sweep_div > 0 ? in->sweep_div : 1;
    c->graph_stale = true;
    return MFCFD_OK;
}

int mfcfd_get_residual(mfcfd_ctx* c, mfcfd_residual* out) {
    if (!c || !out) return fail(MFCFD_EINVAL, "null argument");
    out->res_old = c->res_old;
    out->res_new = c->res_new;
    out->residue = c->residue;
    out->max_res = c->max_res;
    out->sum_res_sqr = c->sum_res_sqr;
    out->max_res_point = c->max_res_point;
    out->iteration = c->last_t;
    return MFCFD_OK;
}

int mfcfd_synchronize(mfcfd_ctx* c) {
    if (!c) return fail(MFCFD_EINVAL, "null context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    return MFCFD_OK;
}

int mfcfd_time_phases(mfcfd_ctx* c, int t0, int n_iters, double* out_ms) {
    int rc = check_ctx(c, true);
    if (rc) return rc;
    if (n_iters <= 0 || !out_ms) return fail(MFCFD_EINVAL, "n_iters must be positive and out_ms non-null");
    if ((rc = prepare_plan(c))) return rc;
    hipEvent_t ev[MFCFD_PHASE_COUNT + 1];
    for (auto& e : ev) HIP_TRY(hipEventCreate(&e));
    double acc[MFCFD_PHASE_COUNT] = {0, 0, 0, 0, 0, 0};
    const DevGeom g = c->geom();
    const DevState s = c->state();
    for (int it = 0; it < n_iters && rc == 0; ++it) {
        hipEventRecord(ev[0], c->stream);
        launch_conserved_and_q(g, s, c->fp, c->stream);
        hipEventRecord(ev[1], c->stream);
        if (c->fp.inner_iterations > 0) {
            launch_q_derivatives(g, s, c->fp, c->temp.p, c->stream);
            launch_timestep(g, s, c->fp, c->stream);
        } else {
            launch_qderiv_timestep(g, s, c->fp, c->stream, c->shapes.qderiv);
        }
        hipEventRecord(ev[2], c->stream);
        if (c->params.flux_order == MFCFD_FLUX_LISTS) {
            launch_flux_residual_reference(g, s, c->fp, c->tl, c->stream);
            hipEventRecord(ev[3], c->stream);
        } else {
            flux(c, 1);
            hipEventRecord(ev[3], c->stream);
            flux(c, 6);
        }
        hipEventRecord(ev[4], c->stream);
        if ((rc = launch_error("phases"))) break;
        // explicit form: the sweep is part of the state update's pass
        const bool fused = c->params.mode == MFCFD_EXPLICIT;
        if (!fused && (rc = do_sweeps(c))) break;
        hipEventRecord(ev[5], c->stream);
        if ((rc = do_state_update(c, nullptr, 0, nullptr, fused))) break;
        hipEventRecord(ev[6], c->stream);
        if (hipEventSynchronize(ev[6]) != hipSuccess) {
            rc = fail(MFCFD_EHIP, "event synchronize");
            break;
        }
        for (int ph = 0; ph < MFCFD_PHASE_COUNT; ++ph) {
            float ms = 0.f;
            hipEventElapsedTime(&ms, ev[ph], ev[ph + 1]);
            acc[ph] += ms;
        }
        if ((rc = fetch_record(c))) break;
        residue_bookkeeping(c, t0 + it, c->res_new);
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    if (rc) return rc;
    for (int ph = 0; ph < MFCFD_PHASE_COUNT; ++ph) out_ms[ph] = acc[ph] / n_iters;
    return MFCFD_OK;
}

int mfcfd_sweep_info(mfcfd_ctx* c, int* levels_fwd, int* levels_bwd, int* n_swept) {
    int rc = check_ctx(c, true);
    if (rc) return rc;
    // the plan of the implicit form in use (the canonical one in explicit mode)
    const int form = implicit_form(c);
    if (c->plan_form != form && (rc = build_plan(c, form))) return rc;
    if (levels_fwd) *levels_fwd = (int)c->sw_fwd.level_ptr.size() - 1;
    if (levels_bwd) *levels_bwd = (int)c->sw_bwd.level_ptr.size() - 1;
    if (n_swept) *n_swept = c->sw_fwd.count;
    return MFCFD_OK;
}

int mfcfd_point_class_map(mfcfd_ctx* c, int* out) {
    if (!c || !c->have_geom) return fail(MFCFD_ESTATE, "no geometry bound");
    if (!out) return fail(MFCFD_EINVAL, "null argument");
    std::memcpy(out, c->host_class.data(), sizeof(int) * c->host_class.size());
    return MFCFD_OK;
}

int mfcfd_kernel_timing(mfcfd_ctx* c, int enable) {
    if (!c) return fail(MFCFD_EINVAL, "null context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    if (enable < 0 || enable > MFCFD_TIME_SWEEPS) return fail(MFCFD_EINVAL, "kernel timing selector %d", enable);
    c->time_kernel = enable;
    c->tev_used = 0;
    c->tev_ms = 0.0;
    c->tev_launches = 0;
    return MFCFD_OK;
}

int mfcfd_kernel_time(mfcfd_ctx* c, double* total_ms, long long* launches) {
    if (!c) return fail(MFCFD_EINVAL, "null context");
    HIP_TRY(hipStreamSynchronize(c->stream));
    collect_kernel_times(c);
    if (total_ms) *total_ms = c->tev_ms;
    if (launches) *launches = c->tev_launches;
    return MFCFD_OK;
}

int mfcfd_point_classes(mfcfd_ctx* c, int* n_fused_interior, int* n_fused_wall, int* n_listwise) {
    if (!c || !c->have_geom) return fail(MFCFD_ESTATE, "no geometry bound");
    if (n_fused_interior) *n_fused_interior = c->n_fused_interior_pts;
    if (n_fused_wall) *n_fused_wall = c->n_fused_general_pts;
    if (n_listwise) *n_listwise = c->tl.n_generic;
    return MFCFD_OK;
}

}  // extern "C"
