// kernels_sweep.hip -- LU-SGS sweeps without level barriers.
//
// The forward sweep of implicit_forward_and_reverse_sweeps.f90:26-57 visits
// the swept points in alias order; point i reads delUp of every stencil
// neighbour with a smaller alias, already updated earlier in the same sweep
// (implicit_LdelU_forward_sweep_interior.f90:51-117).  The backward sweep
// (implicit_forward_and_reverse_sweeps.f90:58-85) mirrors it with delUn and
// larger aliases.  The alternative driver (implicit_forward_and_reverse_
// sweeps_backup.f90:25-139) visits wall, interior, outer and supersonic
// points in type order with the same alias tests, so a neighbour it visits
// later contributes its value from before the sweep.
//
// The host plan (preprocess.h, SweepPlan) lists, per execution position,
// the entries the point takes, each packed as k << 4 | wait << 3 | d:
// direction list d (kTakeCopy for a supersonic point's copy of its nearest
// neighbour), and whether the entry reads this sweep's value of k ("wait":
// k executes earlier) or the value from before the sweep ("old").  Every
// entry's least-squares weight w (1/dist**power times the Cramer cofactors of
// its list applied to the entry's offset) is computed once, at bind
// (k_sweep_weights).  A position's sweep term is sum_e (Gnew_e - Gold_e)*w_e,
// which equals the reference's Cramer solve of the summed products up to
// rounding.
//
// Per sweep:
//   k_sweep_snap   the values old entries read: prim(U_k + delU_k) of every
//                  point an entry reads as old (pold), and delU of every copy
//                  source read as old (dold), before anything is overwritten;
//   k_sweep_mark   marks pub (the published primitive state) of every
//                  position, and delU of every copy source read as this
//                  sweep's value, with a signalling-NaN pattern that no
//                  arithmetic produces (kPending);
//   k_sweep_async  a persistent grid of workgroups of kSweepWaves waves; each
//                  round thread 0 takes kSweepWaves consecutive positions
//                  with one atomic and wave w works on position base + w.
//                  Lane e of the wave takes entry e (e + 64, ... for longer
//                  lists): Gold from the neighbour's state before the sweep
//                  first (it does not depend on the sweep, so its loads and
//                  arithmetic overlap the wait), then the neighbour's
//                  published primitive state -- polled until it is no
//                  longer pending for a wait entry, read from pold for an
//                  old one -- and Gnew from it.  The wave's sum is a
//                  fixed-order DPP reduction that leaves the sums in lane
//                  kTailLane.  The entries of the dependency that executes
//                  last (the plan puts them at the end, late[t] of them) are
//                  held back: the wave sums the others first, then waits for
//                  that one point on lanes 60..63 and adds their few terms
//                  with two row shifts, so the full reduction is off the
//                  dependency chain.  Lane kTailLane publishes the point's state
//                  prim(U_i + delU_i), computed once here instead of once
//                  per entry that reads it, then stores its delU.  The
//                  mixed form publishes that state as five FP32 values in
//                  three 64-bit words: its consumers evaluate their fluxes
//                  in FP32, so three words per poll instead of five.
//
// Memory model.  Inside k_sweep_async the memory other waves read while it
// runs is pub (and dsrc, the published delU of the copy sources), each 64-bit
// value written once.  These arrays are fine-grained device memory: hipMalloc's
// coarse-grained memory is coherent only at kernel boundaries, and on MI355X
// a value written there by one wave was measured still reading as pending,
// through agent-scope atomic loads on another CU, a second after its writer
// had finished.  A position publishes each value with a device-scope atomic
// swap; a waiting lane reads with agent-scope atomic loads and, should a
// value still read pending after kSpinAtomic polls, re-reads it with
// device-scope atomic adds of zero, which are performed at the device's
// coherence point (the zero comes from a kernel argument, so the compiler
// cannot turn the add into a load).  A non-pending value is the final one:
// the marks are written before the kernel starts and each location is
// written once while it runs.  Each value is checked on its own, so no
// ordering between locations is needed.  Everything else the kernel reads
// is not written while it runs, and delU, which the later kernels read, is
// ordered by the kernel boundary.
//
// Progress.  Every wait entry reads a point that executes earlier, so it was
// taken earlier (or in the same round) by a wave that is running.  The lowest
// unfinished position never waits on an unfinished one, and the barrier
// between a workgroup's rounds only delays its next positions, which are
// higher than every position taken, so every wait ends.  A wait that
// exceeds the spin limit is counted in the residual record and the host
// reports an error; the kernel still runs to completion.
//
// T is the arithmetic of the split fluxes and of the published state (float
// in the mixed form); the weights, the sums, D_inv and the new delU stay FP64.
// The FP64 flux evaluations take the sweep form of qlskum_math.h (flux_eval
// with SWEEP), and this unit is compiled with -ffp-contract=on (a*b + c
// fused within an expression): every Gold and Gnew of both sweeps is
// evaluated the same way, so an unchanged neighbour contributes exactly zero.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "device_state.h"
#include "kernels.h"
#include "preprocess.h"
#include "qlskum_math.h"

namespace mfcfd {

namespace {

// waves per workgroup: one point per wave, and one ticket counter atomic per
// workgroup and round for kSweepWaves consecutive positions.  Same-address
// device-scope atomics serialise (~12 ns each), so a ticket per 4 positions
// capped the sweep near 330 positions per microsecond, close to what the 8M
// sweeps need (220 FP64, 270 mixed): with 8 per ticket the FP64 sweep took
// 35.3 -> 33.6 ms and the mixed one 29.2 -> 26.3 ms.  8 waves of a workgroup
// still fit the FP64 kernel's 2 waves per SIMD.
constexpr int kSweepWaves = 8;
constexpr int kSweepBlock = 64 * kSweepWaves;
// the late entries sit on lanes 64 - kMaxLate .. 63, which two row shifts sum
static_assert(kMaxLate >= 1 && kMaxLate <= 4, "late lanes must fit the last four lanes of a row");
// about a second of polling; legitimate waits end within milliseconds
constexpr int kSpinLimit = 1 << 20;
// polls by atomic load before the atomic-add fallback
constexpr int kSpinAtomic = 4096;

// a signalling NaN with a payload: arithmetic only ever produces quiet NaNs
constexpr unsigned long long kPending = 0x7FF4DEAD5EEDF00DULL;

// publish one value: a device-scope atomic swap, performed at the coherence
// point every XCD's atomics read
__device__ __forceinline__ void store_value(double* p, double v) {
    __hip_atomic_exchange(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v),
                          __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// read NW 64-bit words written in this sweep, waiting until none is pending:
// agent-scope atomic loads of all of them at every poll (polling one value and
// re-reading the others once it arrives costs a second memory round trip on
// the dependency chain: measured 21% slower sweeps), device-scope atomic adds
// of zero after kSpinAtomic polls.  Gives up after the spin limit, or at once
// when another wait of this sweep already gave up, so a broken schedule ends
// the kernel instead of stalling it point by point.
template <int NW>
__device__ __forceinline__ void wait_words(double* v, unsigned long long out[NW], int* stalls,
                                           unsigned long long zero) {
    unsigned long long* w = reinterpret_cast<unsigned long long*>(v);
    bool ok = true;
#pragma unroll
    for (int r = 0; r < NW; ++r) {
        out[r] = __hip_atomic_load(w + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        ok = ok && out[r] != kPending;
    }
    if (ok) return;
    int spins = 0;
#pragma clang loop unroll(disable)
    for (;;) {
        ++spins;
        if ((spins & 255) == 0 && __hip_atomic_load(stalls, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
            return;
        if (spins > kSpinLimit) {
            __hip_atomic_fetch_add(stalls, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        __builtin_amdgcn_s_sleep(1);
        ok = true;
#pragma unroll
        for (int r = 0; r < NW; ++r) {
            out[r] = spins < kSpinAtomic
                         ? __hip_atomic_load(w + r, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                         : __hip_atomic_fetch_add(w + r, zero, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            ok = ok && out[r] != kPending;
        }
        if (ok) return;
    }
}

__device__ __forceinline__ void wait_read(double* v5, double out[5], int* stalls, unsigned long long zero) {
    unsigned long long b[5];
    wait_words<5>(v5, b, stalls, zero);
#pragma unroll
    for (int r = 0; r < 5; ++r) out[r] = __longlong_as_double((long long)b[r]);
}

// The mixed form publishes prim(U + delU) as five FP32 values in three 64-bit
// words (its consumers evaluate their fluxes in FP32 anyway): lo/hi halves
// p0/p1, p2/p3, p4/0.  A word with a zero high half never matches kPending;
// the other two would need a high half of 0x7FF4DEAD, a NaN payload that
// arithmetic does not produce (it yields the default NaN, or passes on the
// payload of a NaN operand, and no operand here carries that one).
__device__ __forceinline__ unsigned long long pack2(float lo, float hi) {
    return (unsigned long long)__float_as_uint(lo) | ((unsigned long long)__float_as_uint(hi) << 32);
}

__device__ __forceinline__ void wait_read_packed(double* v, float out[5], int* stalls, unsigned long long zero) {
    unsigned long long b[3];
    wait_words<3>(v, b, stalls, zero);
    out[0] = __uint_as_float((unsigned)b[0]);
    out[1] = __uint_as_float((unsigned)(b[0] >> 32));
    out[2] = __uint_as_float((unsigned)b[1]);
    out[3] = __uint_as_float((unsigned)(b[1] >> 32));
    out[4] = __uint_as_float((unsigned)b[2]);
}

// a 64-bit value every lane of the wave holds alike, in scalar registers
__device__ __forceinline__ long long uniform64(long long v) {
    const int lo = __builtin_amdgcn_readfirstlane((int)v);
    const int hi = __builtin_amdgcn_readfirstlane((int)(v >> 32));
    return ((long long)hi << 32) | (long long)(unsigned)lo;
}

// one DPP move of a double (both 32-bit halves through v_mov_b32_dpp): a
// lane whose source lane lies outside its row, or whose row is masked off,
// reads 0
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double v) {
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, true);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

// The lane that finishes a position.  The reduction below leaves the wave's
// sums in lane 63, so the point's own inputs are loaded into that lane too,
// and nothing crosses lanes between the reduction and the publication.
constexpr int kTailLane = 63;

// Sums over the 64 lanes of a wave (all active), in a fixed order, for five
// values at once, left in lane kTailLane.  Prefix sums within each row of 16
// lanes (row_shr 1, 2, 4, 8: lane 15 of a row holds the row's sum), row 0's
// sum into row 1 and row 2's into row 3 (row_bcast:15), then lane 31's into
// rows 2 and 3 (row_bcast:31): lane 63 ends with the total.  Each step is one
// VALU move per half, where a shuffle is a round trip through the LDS
// crossbar.  The five values take each step together, so five independent
// chains are in flight instead of one: the reduction sits on the sweep's
// dependency chain.
__device__ __forceinline__ void wave_sum5(double v[5]) {
#pragma unroll
    for (int r = 0; r < 5; ++r) v[r] = v[r] + dpp_move<0x111, 0xf>(v[r]);
#pragma unroll
    for (int r = 0; r < 5; ++r) v[r] = v[r] + dpp_move<0x112, 0xf>(v[r]);
#pragma unroll
    for (int r = 0; r < 5; ++r) v[r] = v[r] + dpp_move<0x114, 0xf>(v[r]);
#pragma unroll
    for (int r = 0; r < 5; ++r) v[r] = v[r] + dpp_move<0x118, 0xf>(v[r]);
#pragma unroll
    for (int r = 0; r < 5; ++r) v[r] = v[r] + dpp_move<0x142, 0xa>(v[r]);
#pragma unroll
    for (int r = 0; r < 5; ++r) v[r] = v[r] + dpp_move<0x143, 0xc>(v[r]);
}

// The end of a position, on lane kTailLane: U_i + delU and its primitive
// state, published (FP64, or FP32 packed in three words in the mixed form),
// delU published too when the point is a copy source read in this sweep,
// then delU stored for the later kernels.  Called at the end of both kinds
// of position, so that no values merge between them before it.
template <typename T>
__device__ __forceinline__ void finish_position(const SweepAsync& a, double* dU, int i, const Frame<double>& f,
                                                const double res[5], const double Ui[5], int slot) {
    double U[5];
#pragma unroll
    for (int r = 0; r < 5; ++r) U[r] = res[r] + Ui[r];
    if constexpr (sizeof(T) == 4) {
        const Frame<float> f32 = frame_as<float>(f);
        float U32[5], p32[5];
#pragma unroll
        for (int r = 0; r < 5; ++r) U32[r] = float(U[r]);
        conserved_to_prim(U32, f32, p32);
        unsigned long long* w = reinterpret_cast<unsigned long long*>(a.pub + 5 * (long long)i);
        __hip_atomic_exchange(w, pack2(p32[0], p32[1]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_exchange(w + 1, pack2(p32[2], p32[3]), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        __hip_atomic_exchange(w + 2, pack2(p32[4], 0.0f), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
        double pp[5];
        conserved_to_prim(U, f, pp);
#pragma unroll
        for (int r = 0; r < 5; ++r) store_value(a.pub + 5 * (long long)i + r, pp[r]);
    }
    if (slot >= 0) {
#pragma unroll
        for (int r = 0; r < 5; ++r) store_value(a.dsrc + 5 * (long long)slot + r, res[r]);
    }
#pragma unroll
    for (int r = 0; r < 5; ++r) dU[5 * (long long)i + r] = res[r];
}

// the list a sweep walks for point i and direction d (SweepPlan's
// sweep_list): the flux stencil's list, except an outer point's z term,
// which walks conn (implicit_LdelU_forward_sweep_outer.f90:346-347)
__device__ __forceinline__ void sweep_list_dev(const DevGeom& g, int cls, int d, int i, const int*& list,
                                               int& count) {
    if (cls == 2 && d == ZPOS) {
        list = g.conn + kConnCols * (long long)i;
        count = g.nbhs[i];
        return;
    }
    list = g.lst + g.dir_off[(long long)d * g.n + i];
    count = g.dir_cnt[(long long)d * g.n + i];
}

// setup: the weight of every taken entry, one thread per position.  Entries
// come grouped by direction; each direction's Cramer cofactors are those of
// the whole list the reference walks (list_cofactors, the same arithmetic as
// the flux stencils' weights).
template <bool UW>
__global__ void __launch_bounds__(kBlock) k_sweep_weights(DevGeom g, SweepAsync a, double power) {
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= a.count) return;
    const int i = a.seq[u];
    const int cls = g.status[i];
    const Frame<double> f = load_frame(g.frame, i);
    const double4 pi4 = g.pos[i];
    double c[3] = {0.0, 0.0, 0.0};
    int cur = -1;
    for (long long e = a.take_ptr[u]; e < a.take_ptr[u + 1]; ++e) {
        const unsigned packed = a.take[e];
        const int d = (int)(packed & 7u);
        if (d == (int)kTakeCopy) {
            a.wgt[e] = 0.0;
            continue;
        }
        if (d != cur) {
            const int* list;
            int count;
            sweep_list_dev(g, cls, d, i, list, count);
            list_cofactors<UW>(g, i, f, list, count, d >> 1, power, c);
            cur = d;
        }
        const double4 pk = g.pos[packed >> kTakeShift];
        const double delx = pk.x - pi4.x, dely = pk.y - pi4.y, delz = pk.z - pi4.z;
        const double dels = delx * f.t1[0] + dely * f.t1[1] + delz * f.t1[2];
        const double delt = delx * f.t2[0] + dely * f.t2[1] + delz * f.t2[2];
        const double deln = delx * f.n[0] + dely * f.n[1] + delz * f.n[2];
        a.wgt[e] = wls_weight<UW>(dels, delt, deln, power) * (c[0] * dels + c[1] * delt + c[2] * deln);
    }
}

// the values old entries read, taken before the sweep overwrites anything:
// pold = prim(U + delU) (get_primitive_vector.f90:10-41) of the points
// regular entries read as old, dold = delU of the copy sources read as old
__global__ void __launch_bounds__(kBlock) k_sweep_snap(DevGeom g, DevState s, const double* dU, SweepAsync a) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e < a.n_old) {
        const int k = a.old_pts[e];
        double U[5], prim[5];
#pragma unroll
        for (int r = 0; r < 5; ++r) U[r] = dU[5 * (long long)k + r] + s.U[5 * (long long)k + r];
        conserved_to_prim(U, load_frame(g.frame, k), prim);
#pragma unroll
        for (int r = 0; r < 5; ++r) a.pold[5 * (long long)k + r] = prim[r];
    } else if (e < (long long)a.n_old + a.n_old_src) {
        const int q = a.old_src[e - a.n_old];
#pragma unroll
        for (int r = 0; r < 5; ++r) a.dold[5 * (long long)q + r] = dU[5 * (long long)q + r];
    }
}

// pending marks: pub of every position and every slot of dsrc
__global__ void __launch_bounds__(kBlock) k_sweep_mark(SweepAsync a) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long n1 = 5LL * a.count;
    unsigned long long* dst;
    if (e < n1) {
        const long long t = e / 5;
        dst = reinterpret_cast<unsigned long long*>(a.pub) + 5LL * a.seq[t] + (e - 5 * t);
    } else if (e < n1 + 5LL * a.n_wait_src) {
        dst = reinterpret_cast<unsigned long long*>(a.dsrc) + (e - n1);
    } else {
        return;
    }
    __hip_atomic_store(dst, kPending, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

template <bool FORWARD, bool UW, typename T>
__global__ void __launch_bounds__(kSweepBlock) k_sweep_async(DevGeom g, DevState s, FlowParams fp,
                                                             SweepAsync a) {
    // double-buffered: round r writes slot r & 1 while a wave may still read
    // the other; a slot is rewritten two rounds later, after a barrier that
    // every wave reaches only after reading it
    __shared__ unsigned sh_base[2];
    const int lane = threadIdx.x & 63;
    const unsigned wave = threadIdx.x >> 6;
    double* dU = FORWARD ? s.delUp : s.delUn;
    for (unsigned round = 0;; ++round) {
        // a position's work up to its waits is off the dependency chain
        __builtin_amdgcn_s_setprio(0);
        if (threadIdx.x == 0)
            sh_base[round & 1] = __hip_atomic_fetch_add(a.ctr, (unsigned)kSweepWaves, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        // wave-uniform values are moved to scalar registers (readfirstlane):
        // the position, the point and its entry range
        const unsigned base = __builtin_amdgcn_readfirstlane(sh_base[round & 1]);
        if (base >= (unsigned)a.count) break;   // the same for the whole workgroup
        const unsigned t = __builtin_amdgcn_readfirstlane(base + wave);
        if (t >= (unsigned)a.count) continue;   // the last, partial round
        const int i = __builtin_amdgcn_readfirstlane(a.seq[t]);
        const long long tb = uniform64(a.take_ptr[t]), te = uniform64(a.take_ptr[t + 1]);
        const Frame<double> f = load_frame(g.frame, i);
        // the point's conserved state for the final update, loaded now so
        // that its latency stays off the path after the reduction
        double Ui[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        int slot = -1;   // the point's slot as a copy source read in this sweep
        if (lane == kTailLane) {
#pragma unroll
            for (int r = 0; r < 5; ++r) Ui[r] = s.U[5 * (long long)i + r];
            if (a.src_slot) slot = a.src_slot[i];
        }
        const unsigned first = te > tb ? __builtin_amdgcn_readfirstlane(a.take[tb]) : 0u;
        if (te > tb && (first & 7u) == kTakeCopy) {
            // supersonic point: the nearest neighbour's delU
            // (implicit_forward_and_reverse_sweeps.f90:38-52, 69-83)
            if (lane == kTailLane) {
                const long long q = (long long)(first >> kTakeShift);
                double res[5];
                if (first & kTakeWait) {
                    wait_read(a.dsrc + 5 * (long long)a.src_slot[q], res, a.stalls, a.zero);
                } else {
#pragma unroll
                    for (int r = 0; r < 5; ++r) res[r] = a.dold[5 * q + r];
                }
                finish_position<T>(a, dU, i, f, res, Ui, slot);
            }
        } else {
            const int cls = __builtin_amdgcn_readfirstlane(g.status[i]);
            const Frame<T> ft = frame_as<T>(f);
            // the point's own inputs of the final update, read before the wait
            double D_inv = 0.0, own[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            if (lane == kTailLane) {
                D_inv = s.dinv[i];
#pragma unroll
                for (int r = 0; r < 5; ++r)
                    own[r] = FORWARD ? s.flux_res[5 * (long long)i + r] : s.delUp[5 * (long long)i + r];
            }
            // The entries of the wait dependency that executes last (it
            // usually publishes last, well after the others) sit at the end of
            // the list, at most kMaxLate of them, one per lane from lane
            // 64 - n_late.  Their Gold is computed now, before any wait; the
            // other entries are summed over the wave first, and these then
            // only add their own terms.  The mixed form only: in FP64 the
            // second flux evaluation this costs every position (a whole
            // wave's worth of instructions for a few lanes) loads the SIMDs
            // more than the shorter chain gains (measured: 35.9 -> 38.1 ms
            // per FP64 sweep, 30.7 -> 29.2 ms per mixed one).
            const int n_late = sizeof(T) == 4 ? __builtin_amdgcn_readfirstlane((int)a.late[t]) : 0;
            const long long tl = te - n_late;
            const int lj = lane - (64 - n_late);   // >= 0 on a late lane
            T Gold_l[5] = {T(0), T(0), T(0), T(0), T(0)};
            double w_l = 0.0;
            long long k_l = 0;
            int d_l = 0;
            if (lj >= 0) {
                const unsigned packed = a.take[tl + lj];
                k_l = (long long)(packed >> kTakeShift);
                d_l = (int)(packed & 7u);
                w_l = a.wgt[tl + lj];
                T prim_old[5];
#pragma unroll
                for (int r = 0; r < 5; ++r) prim_old[r] = T(s.prim[5 * k_l + r]);
                flux_eval<T, true>(dir_family(cls, d_l), d_l >> 1, (d_l & 1) ? -1 : 1, ft, prim_old, T(fp.pi),
                                   Gold_l);
            }
            double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
            for (long long e = tb + lane; e < tl; e += 64) {
                const unsigned packed = a.take[e];
                const long long k = (long long)(packed >> kTakeShift);
                const int d = (int)(packed & 7u);
                const double wgt = a.wgt[e];
                // the flux of the neighbour's state before the sweep first:
                // it does not depend on the sweep, so its loads and
                // arithmetic overlap the wait for the state of this sweep,
                // whose flux follows
                T prim_old[5], Gold[5];
#pragma unroll
                for (int r = 0; r < 5; ++r) prim_old[r] = T(s.prim[5 * k + r]);
                flux_eval<T, true>(dir_family(cls, d), d >> 1, (d & 1) ? -1 : 1, ft, prim_old, T(fp.pi), Gold);
                T pt[5], Gnew[5];
                if (packed & kTakeWait) {
                    if constexpr (sizeof(T) == 4) {
                        wait_read_packed(a.pub + 5 * k, pt, a.stalls, a.zero);
                    } else {
                        double pk[5];
                        wait_read(a.pub + 5 * k, pk, a.stalls, a.zero);
#pragma unroll
                        for (int r = 0; r < 5; ++r) pt[r] = T(pk[r]);
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < 5; ++r) pt[r] = T(a.pold[5 * k + r]);
                }
                // with no late entries set apart, the wave is on the dependency
                // chain from here to the publication: its SIMD issues for it
                // before a wave still working ahead of its waits
                if (n_late == 0) __builtin_amdgcn_s_setprio(3);
                flux_eval<T, true>(dir_family(cls, d), d >> 1, (d & 1) ? -1 : 1, ft, pt, T(fp.pi), Gnew);
#pragma unroll
                for (int r = 0; r < 5; ++r) acc[r] = acc[r] + (double)(Gnew[r] - Gold[r]) * wgt;
            }
            wave_sum5(acc);
            if (n_late > 0) {
                // the late entries: their terms, summed into lane kTailLane by
                // two row shifts over lanes 60..63, then added to the others'
                double c[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
                if (lj >= 0) {
                    T pt[5], Gnew[5];
                    if constexpr (sizeof(T) == 4) {
                        wait_read_packed(a.pub + 5 * k_l, pt, a.stalls, a.zero);
                    } else {
                        double pk[5];
                        wait_read(a.pub + 5 * k_l, pk, a.stalls, a.zero);
#pragma unroll
                        for (int r = 0; r < 5; ++r) pt[r] = T(pk[r]);
                    }
                    // on the dependency chain from here to the publication
                    __builtin_amdgcn_s_setprio(3);
                    flux_eval<T, true>(dir_family(cls, d_l), d_l >> 1, (d_l & 1) ? -1 : 1, ft, pt, T(fp.pi),
                                       Gnew);
#pragma unroll
                    for (int r = 0; r < 5; ++r) c[r] = (double)(Gnew[r] - Gold_l[r]) * w_l;
                }
#pragma unroll
                for (int r = 0; r < 5; ++r) c[r] = c[r] + dpp_move<0x111, 0xf>(c[r]);
#pragma unroll
                for (int r = 0; r < 5; ++r) c[r] = c[r] + dpp_move<0x112, 0xf>(c[r]);
#pragma unroll
                for (int r = 0; r < 5; ++r) acc[r] = acc[r] + c[r];
            }
            if (lane == kTailLane) {
                // wall terms are doubled (implicit_LdelU_forward_sweep_wall.f90:405);
                // the product with 1 or 2 is exact
                const double twice = cls == 1 ? 2.0 : 1.0;
                double res[5];
#pragma unroll
                for (int r = 0; r < 5; ++r) {
                    const double term = twice * acc[r];
                    res[r] = FORWARD ? -(D_inv * (own[r] + term)) : own[r] - D_inv * term;
                }
                finish_position<T>(a, dU, i, f, res, Ui, slot);
            }
        }
    }
    // the last workgroup out resets the counters for the next sweep
    if (threadIdx.x == 0) {
        const unsigned done = __hip_atomic_fetch_add(a.ctr + 1, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        if (done == gridDim.x - 1) {
            __hip_atomic_store(a.ctr, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            __hip_atomic_store(a.ctr + 1, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
}

__global__ void k_copy_sweep(double* dst, const double* src, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

template <bool FORWARD, bool UW, typename T>
int resident_grid(int count) {
    int dev = 0, per_cu = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_sweep_async<FORWARD, UW, T>,
                                                     kSweepBlock, 0) != hipSuccess)
        return 0;
    if (per_cu < 1) per_cu = 1;
    const long long grid = (long long)per_cu * prop.multiProcessorCount;
    return (int)(grid < count ? grid : count);
}

template <bool UW, typename T>
int pair_grid(int count) {
    const int a = resident_grid<true, UW, T>(count);
    const int b = resident_grid<false, UW, T>(count);
    return a < b ? a : b;
}

template <bool FORWARD, bool UW, typename T>
void launch_one(const DevGeom& g, const DevState& s, const FlowParams& p, const SweepAsync& a, int grid,
                hipStream_t st, const hipEvent_t* ev) {
    double* dU = FORWARD ? s.delUp : s.delUn;
    const long long n_snap = (long long)a.n_old + a.n_old_src;
    if (n_snap > 0)
        hipLaunchKernelGGL(k_sweep_snap, dim3(grid_for(n_snap, kBlock)), dim3(kBlock), 0, st, g, s,
                           (const double*)dU, a);
    const long long n_mark = 5LL * ((long long)a.count + a.n_wait_src);
    hipLaunchKernelGGL(k_sweep_mark, dim3(grid_for(n_mark, kBlock)), dim3(kBlock), 0, st, a);
    if (ev) (void)hipEventRecord(ev[0], st);
    hipLaunchKernelGGL((k_sweep_async<FORWARD, UW, T>), dim3(grid), dim3(kSweepBlock), 0, st, g, s, p, a);
    if (ev) (void)hipEventRecord(ev[1], st);
}

template <bool UW, typename T>
void launch_pair(const DevGeom& g, const DevState& s, const FlowParams& p, const SweepAsync& fwd,
                 const SweepAsync& bwd, int grid, hipStream_t st, const hipEvent_t* ev) {
    launch_one<true, UW, T>(g, s, p, fwd, grid, st, ev);
    launch_one<false, UW, T>(g, s, p, bwd, grid, st, ev ? ev + 2 : nullptr);
}

}  // namespace

int sweep_async_grid(int count, bool unit_weight, bool mixed) {
    if (count <= 0) return 0;
    if (mixed) return unit_weight ? pair_grid<true, float>(count) : pair_grid<false, float>(count);
    return unit_weight ? pair_grid<true, double>(count) : pair_grid<false, double>(count);
}

void launch_sweep_weights(const DevGeom& g, const SweepAsync& a, double power, hipStream_t st) {
    if (a.count <= 0) return;
    if (power == 0.0)
        hipLaunchKernelGGL(k_sweep_weights<true>, dim3(grid_for(a.count, kBlock)), dim3(kBlock), 0, st, g, a, power);
    else
        hipLaunchKernelGGL(k_sweep_weights<false>, dim3(grid_for(a.count, kBlock)), dim3(kBlock), 0, st, g, a, power);
}

void launch_sweeps_async(const DevGeom& g, const DevState& s, const FlowParams& p, const SweepAsync& fwd,
                         const SweepAsync& bwd, int grid, hipStream_t st, const hipEvent_t* ev) {
    if (fwd.count > 0 && grid > 0) {
        if (p.mixed) {
            if (p.power == 0.0)
                launch_pair<true, float>(g, s, p, fwd, bwd, grid, st, ev);
            else
                launch_pair<false, float>(g, s, p, fwd, bwd, grid, st, ev);
        } else {
            if (p.power == 0.0)
                launch_pair<true, double>(g, s, p, fwd, bwd, grid, st, ev);
            else
                launch_pair<false, double>(g, s, p, fwd, bwd, grid, st, ev);
        }
    }
    // implicit_forward_and_reverse_sweeps.f90:86-88; the alternative driver
    // leaves it to state_update_implicit.f90:98-100
    if (!p.backup)
        hipLaunchKernelGGL(k_copy_sweep, dim3(grid_for(5LL * g.n, kBlock)), dim3(kBlock), 0, st, s.delUp,
                           (const double*)s.delUn, 5LL * g.n);
}

}  // namespace mfcfd
