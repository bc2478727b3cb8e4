// kernels_sweep.hip -- LU-SGS sweeps without level barriers.
//
// The forward sweep of implicit_forward_and_reverse_sweeps.f90:26-57 visits
// the swept points in alias order; point i reads delUp of every stencil
// neighbour with a smaller alias, already updated earlier in the same sweep
// (implicit_LdelU_forward_sweep_interior.f90:51-117).  The backward sweep
// (implicit_forward_and_reverse_sweeps.f90:58-85) mirrors it with delUn and
// larger aliases.  The result is a Gauss-Seidel recurrence whose dependency
// graph is fixed by the alias order.
//
// Here a persistent grid of workgroups takes tickets from a counter: each
// round a workgroup of kSweepWaves waves takes kSweepWaves consecutive
// positions with one atomic, and wave w processes position base + w of the
// execution order.  Every point a wave waits on has a smaller position, so
// it was taken earlier (or in the same round) by a wave that is running.
// The lowest unfinished position never waits, and the barrier between a
// workgroup's rounds only delays its next positions, which are higher than
// every position taken, so every wait ends.
//
// The values themselves are the completion flags.  Before a sweep, a kernel
// marks the five outputs of every swept point with a signalling-NaN
// pattern that no arithmetic produces (kPending).  A point writes its five
// results once, with agent-scope relaxed atomic stores; a wave first
// computes everything that does not depend on this sweep's values
// (geometry, Cramer weights, the old fluxes Gold), then polls its
// dependencies' values with agent-scope relaxed atomic loads until none is
// pending, and uses the values so read.  Each 64-bit value is written once
// and checked on its own, so no ordering between locations -- and no cache
// write-back or invalidation -- is needed.  A wait that exceeds the spin
// limit is counted in the residual record and the host reports an error;
// the kernel still runs to completion.
//
// One wave of 64 lanes handles one point.  The host lists, per execution
// position and sweep direction, the entries of the point's direction lists
// that this sweep takes (the neighbours with a smaller alias in the forward
// sweep, a larger one in the backward sweep; a wall point's z term walks
// conn, so its ZPOS list is skipped), each packed as k << 3 | direction.
// Lane e takes entry e (e + 64, ... for longer lists), and the sum over all
// entries (the sweep term is the sum of the directions' dG, times 2 at wall
// points) is a fixed-order DPP reduction within the wave (wave_sum).  Each entry's
// contribution is (Gnew - Gold) * w (c0*dels + c1*delt + c2*deln) with the
// Cramer cofactors c of the list (k_stencil_coef), which equals the
// reference's Cramer solve on the summed products up to rounding.  T is the
// arithmetic of the split fluxes (float in the mixed form); the weights, the
// sums, D_inv and the new delU stay FP64.
#include <hip/hip_runtime.h>

#include "device_common.h"
#include "device_state.h"
#include "kernels.h"
#include "qlskum_math.h"

namespace mfcfd {

namespace {

// waves per workgroup: one point per wave, and one ticket counter atomic per
// workgroup and round for kSweepWaves consecutive positions (same-address
// device-scope atomics serialise: ~12 ns each, which caps single-ticket
// rounds at ~80 points per microsecond)
constexpr int kSweepWaves = 4;
constexpr int kSweepBlock = 64 * kSweepWaves;
// about a second of polling; legitimate waits end within milliseconds
constexpr int kSpinLimit = 1 << 20;

// a signalling NaN with a payload: arithmetic only ever produces quiet NaNs
constexpr unsigned long long kPending = 0x7FF4DEAD5EEDF00DULL;

__device__ __forceinline__ void store_value(double* p, double v) {
    __hip_atomic_store(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v),
                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// read the five values of a neighbour, waiting until none is pending: a
// neighbour swept earlier in this sweep has then written them, any other
// holds a value from before the sweep, which is never pending.  Gives up
// after the spin limit, or at once when another wait of this sweep already
// gave up, so a broken schedule ends the kernel instead of stalling it point
// by point.
__device__ __forceinline__ void wait_read(const double* v5, double out[5], int* stalls) {
    int spins = 0;
    for (;;) {
        bool ok = true;
#pragma unroll
        for (int r = 0; r < 5; ++r) {
            const unsigned long long b = __hip_atomic_load(reinterpret_cast<const unsigned long long*>(v5 + r),
                                                           __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            out[r] = __longlong_as_double((long long)b);
            ok = ok && b != kPending;
        }
        if (ok) return;
        ++spins;
        if ((spins & 255) == 0 && __hip_atomic_load(stalls, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0)
            return;
        if (spins > kSpinLimit) {
            __hip_atomic_fetch_add(stalls, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return;
        }
        __builtin_amdgcn_s_sleep(1);
    }
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

// Sum over the 64 lanes of a wave (all active), in a fixed order, returned
// to every lane.  Prefix sums within each row of 16 lanes (row_shr 1, 2, 4,
// 8: lane 15 of a row holds the row's sum), row 0's sum into row 1 and row
// 2's into row 3 (row_bcast:15), then lane 31's into rows 2 and 3
// (row_bcast:31): lane 63 ends with the total, read back with readlane.
// Each step is one VALU move per half, where a shuffle is a round trip
// through the LDS crossbar.
__device__ __forceinline__ double wave_sum(double v) {
    v = v + dpp_move<0x111, 0xf>(v);
    v = v + dpp_move<0x112, 0xf>(v);
    v = v + dpp_move<0x114, 0xf>(v);
    v = v + dpp_move<0x118, 0xf>(v);
    v = v + dpp_move<0x142, 0xa>(v);
    v = v + dpp_move<0x143, 0xc>(v);
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)b, 63);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
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
    const double* dU = FORWARD ? s.delUp : s.delUn;
    for (unsigned round = 0;; ++round) {
        if (threadIdx.x == 0)
            sh_base[round & 1] = __hip_atomic_fetch_add(a.ctr, (unsigned)kSweepWaves, __ATOMIC_RELAXED,
                                                         __HIP_MEMORY_SCOPE_AGENT);
        __syncthreads();
        const unsigned base = sh_base[round & 1];
        if (base >= (unsigned)a.count) break;   // the same for the whole workgroup
        const unsigned t = base + wave;
        if (t >= (unsigned)a.count) continue;   // the last, partial round
        const int i = a.seq[t];
        const int cls = g.status[i];
        const Frame<double> f = load_frame(g.frame, i);
        const Frame<T> ft = frame_as<T>(f);
        const double4 pi4 = g.pos[i];
        // the point's own inputs of the final update, read before the wait
        double D_inv = 0.0, own[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        if (lane == 0) {
            D_inv = s.dinv[i];
#pragma unroll
            for (int r = 0; r < 5; ++r) own[r] = FORWARD ? s.flux_res[5 * (long long)i + r] : s.delUp[5 * (long long)i + r];
        }
        double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
        const long long tb = a.take_ptr[t], te = a.take_ptr[t + 1];
        for (long long e = tb + lane; e < te; e += 64) {
            const unsigned packed = a.take[e];
            const int k = (int)(packed >> 3);
            const int d = (int)(packed & 7u);
            // everything that does not depend on this sweep's values first
            const double4 pk = g.pos[k];
            const double delx = pk.x - pi4.x, dely = pk.y - pi4.y, delz = pk.z - pi4.z;
            const double dels = delx * f.t1[0] + dely * f.t1[1] + delz * f.t1[2];
            const double delt = delx * f.t2[0] + dely * f.t2[1] + delz * f.t2[2];
            const double deln = delx * f.n[0] + dely * f.n[1] + delz * f.n[2];
            const double* c = g.coef + 18 * (long long)i + 3 * d;
            const double wgt = wls_weight<UW>(dels, delt, deln, fp.power) * (c[0] * dels + c[1] * delt + c[2] * deln);
            T prim_old[5], Gold[5];
#pragma unroll
            for (int r = 0; r < 5; ++r) prim_old[r] = T(s.prim[5 * (long long)k + r]);
            flux_eval<T>(dir_family(cls, d), d >> 1, (d & 1) ? -1 : 1, ft, prim_old, T(fp.pi), Gold);
            double Uk[5];
#pragma unroll
            for (int r = 0; r < 5; ++r) Uk[r] = s.U[5 * (long long)k + r];
            const Frame<double> fk = load_frame(g.frame, k);
            // then this neighbour's values of the current sweep
            double dk[5], U[5], prim[5];
            wait_read(dU + 5 * (long long)k, dk, a.stalls);
#pragma unroll
            for (int r = 0; r < 5; ++r) U[r] = dk[r] + Uk[r];
            conserved_to_prim(U, fk, prim);
            T pt[5], Gnew[5];
#pragma unroll
            for (int r = 0; r < 5; ++r) pt[r] = T(prim[r]);
            flux_eval<T>(dir_family(cls, d), d >> 1, (d & 1) ? -1 : 1, ft, pt, T(fp.pi), Gnew);
#pragma unroll
            for (int r = 0; r < 5; ++r) acc[r] = acc[r] + (double)(Gnew[r] - Gold[r]) * wgt;
        }
#pragma unroll
        for (int r = 0; r < 5; ++r) acc[r] = wave_sum(acc[r]);
        if (lane == 0) {
            if (FORWARD) {
#pragma unroll
                for (int r = 0; r < 5; ++r)
                    store_value(s.delUp + 5 * (long long)i + r, -(D_inv * (own[r] + (cls == 1 ? 2.0 * acc[r] : acc[r]))));
            } else {
#pragma unroll
                for (int r = 0; r < 5; ++r)
                    store_value(s.delUn + 5 * (long long)i + r, own[r] - D_inv * (cls == 1 ? 2.0 * acc[r] : acc[r]));
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

// mark the outputs of the swept points as not yet written
__global__ void k_mark_pending(double* out, const int* seq, int count) {
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 5LL * count) return;
    const long long t = e / 5;
    reinterpret_cast<unsigned long long*>(out)[5LL * seq[t] + (e - 5 * t)] = kPending;
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

template <bool UW, typename T>
void launch_pair(const DevGeom& g, const DevState& s, const FlowParams& p, const SweepAsync& fwd,
                 const SweepAsync& bwd, int grid, hipStream_t st) {
    const dim3 mg(grid_for(5LL * fwd.count, kBlock));
    hipLaunchKernelGGL(k_mark_pending, mg, dim3(kBlock), 0, st, s.delUp, fwd.seq, fwd.count);
    hipLaunchKernelGGL((k_sweep_async<true, UW, T>), dim3(grid), dim3(kSweepBlock), 0, st, g, s, p, fwd);
    hipLaunchKernelGGL(k_mark_pending, mg, dim3(kBlock), 0, st, s.delUn, bwd.seq, bwd.count);
    hipLaunchKernelGGL((k_sweep_async<false, UW, T>), dim3(grid), dim3(kSweepBlock), 0, st, g, s, p, bwd);
}

}  // namespace

int sweep_async_grid(int count, bool unit_weight, bool mixed) {
    if (count <= 0) return 0;
    if (mixed) return unit_weight ? pair_grid<true, float>(count) : pair_grid<false, float>(count);
    return unit_weight ? pair_grid<true, double>(count) : pair_grid<false, double>(count);
}

void launch_sweeps_async(const DevGeom& g, const DevState& s, const FlowParams& p, const SweepAsync& fwd,
                         const SweepAsync& bwd, int grid, hipStream_t st) {
    if (fwd.count > 0 && grid > 0) {
        if (p.mixed) {
            if (p.power == 0.0)
                launch_pair<true, float>(g, s, p, fwd, bwd, grid, st);
            else
                launch_pair<false, float>(g, s, p, fwd, bwd, grid, st);
        } else {
            if (p.power == 0.0)
                launch_pair<true, double>(g, s, p, fwd, bwd, grid, st);
            else
                launch_pair<false, double>(g, s, p, fwd, bwd, grid, st);
        }
    }
    // implicit_forward_and_reverse_sweeps.f90:86-88
    hipLaunchKernelGGL(k_copy_sweep, dim3(grid_for(5LL * g.n, kBlock)), dim3(kBlock), 0, st, s.delUp,
                       (const double*)s.delUn, 5LL * g.n);
}

}  // namespace mfcfd
