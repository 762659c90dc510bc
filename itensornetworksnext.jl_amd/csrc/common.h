// common.h -- shared device/host definitions of the MI355X BP engine (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define BP_MAXR 12    // max site-tensor rank (physical index + ports)
#define BP_BLOCK 256  // threads per workgroup of the generic kernels

// Flat topology as the kernels read it (device pointers).
struct DevTopo {
    const int *port_off;        // nv + 1
    const int *port_nbr;        // nports
    const int *port_edge;       // nports
    const int *port_rev;        // nports
    const int *edge_dim;        // ne
    const int *phys_dim;        // nv
    const int *owner;           // nports: vertex owning each port
    const int *vnet;            // nv: network of each vertex
    const long long *site_off;  // nv + 1
    const long long *msg_off;   // nports + 1
    int nv;
    int nports;
    int plain;
};

// Per-network convergence state read by the sweep kernels (a network whose `done` flag
// is set is skipped: its messages are carried over unchanged) and updated once per sweep
// by k_reduce_nets (StopWhenConverged / StopAfterIteration,
// AlgorithmsInterfaceExtensions.jl:84-109, beliefpropagation.jl:27-55).  Folding the
// reduction into the sweep kernels with device-scope atomics and fences was measured
// slower on the 8-XCD part than this separate launch.
struct NetCtl {
    int *done;     // per network: 1 once stopped
    int enabled;   // 0: plain sweep, every network updated
};

// Stopping parameters of a run, read from device memory so that a captured sweep sequence
// serves every (maxiter, tol, criterion).
struct RunParams {
    int maxiter;
    int crit;  // 0: fidelity, 1: max-abs (BP_CRIT_*)
    double tol;
};

// The per-network convergence bookkeeping, folded into a sweep kernel that covers every
// message of every network (acc == nullptr: off).  Each work item folds its largest residual
// under the run's criterion into one of its network's FOLD_SLOTS accumulators with one
// atomic max (no fence: the kernel's end orders it before the next launch; the slots keep
// the atomics of one network from serialising on one address).  The stopping rule on those
// maxima is applied by the next sweep kernel as it starts (prev set: every work item of a
// network reads the same maxima and reaches the same decision, and the network's lead item
// records it), or by a one-thread-per-network kernel after the last sweep of a sequence.
// Accumulators rotate through FOLD_BUFS buffers: sweep k folds into buffer k % FOLD_BUFS,
// reads buffer (k - 1) % FOLD_BUFS and clears buffer (k + 1) % FOLD_BUFS, which sweep k - 2
// read last and sweep k + 1 folds into.
#define FOLD_SLOTS 32  // accumulators per network (work item u folds into slot u % 32)
#define FOLD_BUFS 4    // rotating accumulator buffers (a divisor of the captured sweep count)
struct SweepFold {
    unsigned long long *acc;         // this sweep's: per network, FOLD_SLOTS residual bits
    const RunParams *rp;
    const unsigned long long *prev;  // the previous sweep's, its rule not yet applied (or null)
    unsigned long long *clear;       // the next sweep's, cleared by each network's lead item
    const int *lead;                 // per network: its lead work item
    int *iter;                       // per network run state, written by the lead item
    double *delta;
    int *at;
};

// max that keeps a NaN and ignores values below zero (a fidelity residual can round to
// -1e-16): the rule of the per-slot reduction.  Values >= 0 or NaN order like their bits.
__device__ __forceinline__ double fold_max(double m, double r) {
    return (r > m || r != r) ? r : m;
}

// Ordering of LDS traffic within one wave: lane A's store is seen by lane B's later load
// once the wave's LDS operations have completed (s_waitcnt lgkmcnt(0)); the empty asm keeps
// the compiler from moving LDS accesses across.  Used instead of __syncthreads where the
// buffer belongs to the wave alone, so the waves of a workgroup do not wait for each other.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0); vmcnt, expcnt unrestricted
    __builtin_amdgcn_wave_barrier();
    asm volatile("" ::: "memory");
}

// wave64 reductions (DPP/permute through the compiler's __shfl_xor lowering)
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

template <typename T>
__device__ __forceinline__ T wave_max(T v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        T w = __shfl_xor(v, o, 64);
        v = (w > v || w != w) ? w : v;  // NaN-propagating max
    }
    return v;
}

// Exchanges inside 16-lane rows through DPP (a VALU operand modifier: no LDS round trip).
// row_partner<s>(v) returns v of a lane whose index differs from this lane's in bit s and
// possibly in lower bits only: l ^ 1 and l ^ 2 (quad permutations), l ^ 7 (row_half_mirror),
// l ^ 15 (row_mirror).  A halving exchange that takes its steps in the order s = 8, 4, 2, 1
// over such partners still sums every lane of the row into every entry (each step pairs
// lanes that hold the same entries), and lane l ends with entry l % 16.
template <int S>
__device__ __forceinline__ float row_partner(float v) {
    static_assert(S == 1 || S == 2 || S == 4 || S == 8, "row partner distance");
    constexpr int ctrl = S == 1 ? 0xB1 : S == 2 ? 0x4E : S == 4 ? 0x141 : 0x140;
    // every partner is inside the row, so the bound control never applies; with it set the
    // compiler can fold the move into the consuming add (v_add_f32 with a DPP source)
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), ctrl, 0xF, 0xF, true));
}

template <int S>
__device__ __forceinline__ double row_partner(double v) {
    constexpr int ctrl = S == 1 ? 0xB1 : S == 2 ? 0x4E : S == 4 ? 0x141 : 0x140;
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), ctrl, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), ctrl, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Both members of a pair of 16-lane rows (rows 0 and 1, rows 2 and 3) or of the two 32-lane
// halves, through v_permlane16_swap / v_permlane32_swap (VALU): with both operands v, the
// swap returns v of the pair's first member in [0] and of its second in [1], on every lane.
struct DPair { double first, second; };

__device__ __forceinline__ DPair rows16_pair(double v) {
    const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
    const auto a = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
    return {__hiloint2double((int)b[0], (int)a[0]), __hiloint2double((int)b[1], (int)a[1])};
}

__device__ __forceinline__ DPair halves32_pair(double v) {
    const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
    const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    return {__hiloint2double((int)b[0], (int)a[0]), __hiloint2double((int)b[1], (int)a[1])};
}

// wave64 sum and NaN-propagating max of a double on the VALU alone (DPP inside rows, then
// the row and half swaps); every lane gets the same result.
__device__ __forceinline__ double wave_sum_v(double v) {
    v += row_partner<1>(v);
    v += row_partner<2>(v);
    v += row_partner<4>(v);
    v += row_partner<8>(v);
    const DPair r = rows16_pair(v);
    v = r.first + r.second;
    const DPair h = halves32_pair(v);
    return h.first + h.second;
}

__device__ __forceinline__ double vmax_nan(double v, double w) { return (w > v || w != w) ? w : v; }

__device__ __forceinline__ double wave_max_v(double v) {
    v = vmax_nan(v, row_partner<1>(v));
    v = vmax_nan(v, row_partner<2>(v));
    v = vmax_nan(v, row_partner<4>(v));
    v = vmax_nan(v, row_partner<8>(v));
    const DPair r = rows16_pair(v);
    v = vmax_nan(r.first, r.second);
    const DPair h = halves32_pair(v);
    return vmax_nan(h.first, h.second);
}

// Neumaier-compensated running sum of doubles (the log-norm accumulations).
struct NeuSum {
    double s = 0.0, c = 0.0;
    __device__ void add(double x) {
        const double t = s + x;
        if (fabs(s) >= fabs(x)) c += (s - t) + x; else c += (x - t) + s;
        s = t;
    }
};

// block-wide sum of a double over BP_BLOCK threads; `red` is LDS of >= BP_BLOCK/64 doubles.
__device__ __forceinline__ double block_sum(double v, double *red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 0; i < nw; ++i) s += red[i];
    __syncthreads();
    return s;
}

__device__ __forceinline__ double block_max(double v, double *red) {
    v = wave_max(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[w] = v;
    __syncthreads();
    double s = red[0];
    const int nw = (blockDim.x + 63) >> 6;
    for (int i = 1; i < nw; ++i) s = (red[i] > s || red[i] != red[i]) ? red[i] : s;
    __syncthreads();
    return s;
}
