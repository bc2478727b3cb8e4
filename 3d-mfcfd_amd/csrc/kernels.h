// kernels.h -- launch wrappers of the HIP kernels (internal to the library).
#pragma once

#include <hip/hip_runtime.h>

#include "device_state.h"

namespace mfcfd {

// device point numbers of each class
struct TypeLists {
    int n_interior, n_wall, n_outer, n_sup_in, n_sup_out;
    const int* interior;
    const int* wall;
    const int* outer;
    const int* sup_in;
    const int* sup_out;
    // flux residual work split: regular interior / regular wall points take
    // the fused stencil kernel, everything else the list-by-list kernel
    int n_fused_interior, n_fused_wall, n_generic;
    const int* fused_interior;
    const int* fused_wall;
    const int* generic;
};

// launch shape of a gather kernel: block size and the minimum waves per
// SIMD its compiled variant targets (its register cap); 0 selects the default
struct KernelShape {
    int block = 0;
    int waves = 0;
    int lds = 0;   // dynamic LDS bytes reserved per block: caps the blocks resident per CU
};
// shapes the autotuner chooses (defaults until it runs)
struct LaunchShapes {
    KernelShape flux_interior, flux_wall, qderiv;
    int sweep_div = 1;   // sync-free sweeps use the resident grid / sweep_div
};

// level-set schedule of both sweeps: the plan's execution orders cut at
// their dependency levels (one launch per level)
struct SweepSchedule {
    int n_levels_fwd, n_levels_bwd;
    const int* fwd_ptr;      // host copy: level l covers fwd_points[fwd_ptr[l] .. fwd_ptr[l+1])
    const int* bwd_ptr;
    const int* fwd_points;   // device
    const int* bwd_points;   // device
    int n_swept;
};

// counters of the sync-free sweeps: [0] ticket, [1] finished workgroups
// (the rest spare)
constexpr int kSweepCounters = 8;

// sync-free sweeps (kernels_sweep.hip); one per sweep direction, from the
// host plan (SweepPlan in preprocess.h)
struct SweepAsync {
    const int* seq;             // visited points (device numbers) in execution order
    int count;
    const long long* take_ptr;  // [count + 1]: position t takes take[take_ptr[t] .. take_ptr[t+1])
    const unsigned* take;       // k << 4 | wait << 3 | direction (7: supersonic copy) of every taken entry
    const unsigned char* late;  // [count]: position t's last late[t] entries read the wait dependency
                                // that executes last (at most 4; 0: none set apart)
    double* wgt;                // [entries] least-squares weight of each entry (set at bind)
    const int* old_pts;         // points some entry reads as old: pold is taken for them
    int n_old;
    const int* old_src;         // copy sources read as old: dold is taken for them
    int n_old_src;
    const int* wait_src;        // copy sources read as this sweep's delU (slot j of dsrc: wait_src[j])
    int n_wait_src;
    double* pub;                // [5n] published prim(U + delU) of the positions (completion marks;
                                // fine-grained memory)
    double* pold;               // [5n] prim(U + delU) before the sweep, for old entries
    double* dold;               // [5n] delU before the sweep, for copies from old sources (may be null)
    double* dsrc;               // [5 n_wait_src] published delU of the copy sources read as this
                                // sweep's value (fine-grained; may be null)
    const int* src_slot;        // [n] slot of each point in dsrc, -1 if none (may be null)
    unsigned* ctr;              // [kSweepCounters] ticket counter, finished workgroups (zero between sweeps)
    int* stalls;                // waits that hit the spin limit
    unsigned long long zero;    // 0: the operand of the fallback polling atomic adds
};

// setup: Cramer weights of every direction list (geometry only)
void launch_stencil_coefficients(const DevGeom& g, double* coef, double power, hipStream_t st);
// setup: *first_bad = min(8*point + list) over singular least-squares stencils
// (list 0..5 the direction lists, 6 conn); untouched when there are none
void launch_stencil_check(const DevGeom& g, double power, int* first_bad, hipStream_t st);
// boundary permutations: dst[k*m + c] = src[perm[k]*m + c] and the inverse
void launch_gather(double* dst, const double* src, const int* perm, int n, int m, hipStream_t st);
// precision conversions of a flat array (mixed form uploads and downloads)
// (m values per point; the FP32 array holds stride32 per point)
void launch_to_float(float* dst, int stride32, const double* src, int m, long long n, hipStream_t st);
void launch_to_double(double* dst, const float* src, int stride32, int m, long long n, hipStream_t st);
void launch_scatter(double* dst, const double* src, const int* perm, int n, int m, hipStream_t st);

// compute_conserved_vector.f90 / q_variables.f90
void launch_conserved_vector(const DevGeom& g, const DevState& s, hipStream_t st);
void launch_q_variables(const DevGeom& g, const DevState& s, const FlowParams& p, hipStream_t st);
void launch_conserved_and_q(const DevGeom& g, const DevState& s, const FlowParams& p, hipStream_t st);
// q_derivatives.f90 / timestep_delt.f90
void launch_q_derivatives(const DevGeom& g, const DevState& s, const FlowParams& p, double* temp,
                          hipStream_t st);
void launch_timestep(const DevGeom& g, const DevState& s, const FlowParams& p, hipStream_t st);
void launch_qderiv_timestep(const DevGeom& g, const DevState& s, const FlowParams& p, hipStream_t st,
                            const KernelShape& ks = KernelShape());
// flux_residual.f90; part bit 0: fused interior, bit 1: fused wall, bit 2: list-by-list
constexpr int kFluxAll = 7;
void launch_flux_residual(const DevGeom& g, const DevState& s, const FlowParams& p,
                          const TypeLists& tl, hipStream_t st, int parts = kFluxAll,
                          const LaunchShapes& ls = LaunchShapes());
// reference-order flux residual of every point of a class (parity path)
void launch_flux_residual_reference(const DevGeom& g, const DevState& s, const FlowParams& p,
                                    const TypeLists& tl, hipStream_t st);
// explicit form of the sweep (delUn = -delt*flux_res, delUp = delUn)
void launch_explicit_sweep(const DevGeom& g, const DevState& s, hipStream_t st);
// implicit_forward_and_reverse_sweeps.f90
// D_inv of the wall and interior points, and of the outer points when the
// alternative driver sweeps them (p.backup)
void launch_dinv(const DevGeom& g, const DevState& s, const FlowParams& p, const TypeLists& tl,
                 hipStream_t st);
// setup: the geometric part of every direction's D_inv term (12 doubles per point)
void launch_dinv_geometry(const DevGeom& g, double* dgeo, double power, hipStream_t st);
void launch_sweeps_levels(const DevGeom& g, const DevState& s, const FlowParams& p,
                          const SweepSchedule& sch, hipStream_t st);
// forward and backward sweeps without level barriers, then delUp = delUn
// (ev: null, or begin/end events recorded around the forward and the
// backward sweep kernel launches)
void launch_sweeps_async(const DevGeom& g, const DevState& s, const FlowParams& p, const SweepAsync& fwd,
                         const SweepAsync& bwd, int grid, hipStream_t st, const hipEvent_t* ev = nullptr);
// resident workgroup count used for the sync-free sweeps
int sweep_async_grid(int count, bool unit_weight, bool mixed);
// setup: the least-squares weight of every entry of a sweep's plan
void launch_sweep_weights(const DevGeom& g, const SweepAsync& a, double power, hipStream_t st);
// state_update.f90 (+ residual reduction into *rec; hist may be null)
// per-cloud residuals of a batch, recorded in row *counter of inst_hist
// (n_inst per row); the counter advances once per state update
// Each cloud's range is cut into chunks of at most kResidualChunk points;
// chunk c covers [chunk_beg[c], chunk_end[c]), and cloud b owns chunks
// chunk_ptr[b] .. chunk_ptr[b+1]
constexpr int kResidualChunk = 2048;
struct InstanceResidual {
    double* ressq;       // [n] per-point squared density change
    const int* beg;      // [n_inst] device-order point range of each cloud
    const int* end;
    const double* norm;  // [n_inst] the cloud's point count
    int n_inst;
    double* inst_hist;
    int* counter;
    const int* chunk_beg;
    const int* chunk_end;
    const int* chunk_ptr;   // [n_inst + 1]
    int n_chunks;
    double* chunk_sum;      // [n_chunks] scratch
};
// hist[hist_index] (or hist[*ir->counter] when ir is given) receives res_new;
// with_explicit_sweep: do the explicit form's sweep (launch_explicit_sweep)
// in the same pass
void launch_state_update(const DevGeom& g, const DevState& s, const FlowParams& p,
                         const TypeLists& tl, double* partial, int* partial_pos, int n_partial,
                         ResidualRecord* rec, double* hist, int hist_index, hipStream_t st,
                         const InstanceResidual* ir = nullptr, bool with_explicit_sweep = false);
int state_update_blocks(int n);

}  // namespace mfcfd
