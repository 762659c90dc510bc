// bp_engine.hip -- C-ABI implementation of the MI355X belief-propagation engine.
//
// Host side of libbpmi355x.so: context, topology/tensor upload, sweep scheduling with
// per-network convergence on the device, hipGraph replay, scalars and log-norm, FP64
// polish, adjoint.  Kernels live in generic.h (runtime shapes), vertex.h (the
// vertex-centric MFMA path) and adjoint.h.  Entry points are declared in
// include/bp_mi355x.h.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/bp_mi355x.h"
#include "common.h"
#include "generic.h"
#include "complex.h"
#include "vertex.h"
#include "vertex16.h"
#include "vertex64.h"
#include "vertex64l.h"
#include "vertex4.h"
#include "adjoint.h"
#include "vertex_adjoint.h"
#include "contract.h"
#include "gate.h"

// Process-wide plan epoch: moves on every device (re)allocation and every change of a
// kernel choice, so that a cached sweep graph never replays stale pointers.
static std::atomic<unsigned> g_epoch{1};
static void bump_epoch() { g_epoch.fetch_add(1); }

// ------------------------------------------------------------------ error state

static thread_local std::string g_err;

static int set_err(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(expr)                                                                  \
    do {                                                                              \
        hipError_t e_ = (expr);                                                       \
        if (e_ != hipSuccess)                                                         \
            return set_err(BP_ERR_HIP, "%s failed: %s (%s:%d)", #expr,                \
                           hipGetErrorString(e_), __FILE__, __LINE__);                \
    } while (0)

#define RETCHK(expr)            \
    do {                        \
        int rc_ = (expr);       \
        if (rc_ != BP_OK) return rc_; \
    } while (0)

// ------------------------------------------------------------------ no-allocation scratch
//
// bp_iterate_diff_noalloc stages its input and keeps its per-slot partial results in this
// file-scope device array instead of memory it allocates (1 Mi doubles = 8 MiB, held by
// every GPU the library runs on until the process ends).  In host code `scratch` names a
// host-side stand-in, not a device address: the device address comes from
// hipGetSymbolAddress, taken once per device when the first context is created there
// (bp_create) and kept in g_scratch_ptr.  Calls on one GPU share the array, so calls that
// overlap on one GPU race.
#define BP_SCRATCH_DOUBLES (1 << 20)
#define BP_SCRATCH_DEVICES 64
static __device__ __align__(256) double scratch[BP_SCRATCH_DOUBLES];
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[BP_SCRATCH_DEVICES] = {};

// ------------------------------------------------------------------ small kernels

template <typename T>
__global__ void k_init_messages(DevTopo tp, T *m, int kind) {
    const int q = blockIdx.x;
    if (q >= tp.nports) return;
    const long long off = tp.msg_off[q];
    const int len = (int)(tp.msg_off[q + 1] - off);
    const int D = tp.edge_dim[tp.port_edge[q]];
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
        T v = T(1);
        if (kind == BP_INIT_IDENTITY && !tp.plain) v = (i / D == i % D) ? T(1) : T(0);
        m[off + i] = v;
    }
}

template <typename Tin, typename Tout>
__global__ void k_convert(const Tin *in, Tout *out, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = (Tout)in[i];
}

__global__ void k_round_bf16(const double *in, float *out, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = bf16r((float)in[i]);
}

// Start of a run: clear every network's state and the folded accumulators, and set the
// run's stopping parameters (passed by value, so no host copy precedes the sweeps).
__global__ void k_reset_nets(int nnet, int *done, int *iter, double *delta, int *at,
                             unsigned long long *fold_acc, int fold_n, RunParams *rp,
                             RunParams params) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n == 0 && rp) *rp = params;
    if (n >= nnet) return;
    done[n] = 0;
    iter[n] = 0;
    delta[n] = INFINITY;
    at[n] = -1;
    if (fold_acc && n < fold_n)  // every rotating buffer of the folded bookkeeping
        for (int k = 0; k < FOLD_BUFS * FOLD_SLOTS; ++k)
            fold_acc[(long long)(k / FOLD_SLOTS) * fold_n * FOLD_SLOTS +
                     (long long)n * FOLD_SLOTS + k % FOLD_SLOTS] = 0ull;
}

// Per-network convergence bookkeeping after one Jacobi sweep (one workgroup per
// network): StopWhenConverged on `res` (fidelity or max-abs per slot) and
// StopAfterIteration(maxiter) (AlgorithmsInterfaceExtensions.jl:84-109,
// beliefpropagation.jl:27-55).

__global__ __launch_bounds__(1024) void k_reduce_nets(const int *pbeg, const double *res_fid,
                                                      const double *res_abs,
                                                      const RunParams *rp, int *done, int *iter,
                                                      double *delta, int *at) {
    __shared__ double red[16];
    const int n = blockIdx.x;
    if (done[n]) return;  // uniform over the workgroup
    const int maxiter = rp->maxiter;
    const double tol = rp->tol;
    const double *res = rp->crit == BP_CRIT_FIDELITY ? res_fid : res_abs;
    double m = 0.0;
    for (int q = pbeg[n] + threadIdx.x; q < pbeg[n + 1]; q += blockDim.x) {
        const double r = res[q];
        m = (r > m || r != r) ? r : m;
    }
    m = block_max(m, red);
    if (threadIdx.x == 0) {
        const int it = iter[n] + 1;
        iter[n] = it;
        delta[n] = m;
        if (tol >= 0.0 && m < tol) {
            done[n] = 1;
            at[n] = it;
        } else if (maxiter > 0 && it >= maxiter) {
            done[n] = 1;
        }
    }
}

// Per-slot distance between two message buffers (one wave per slot): the reference's
// iterate_diff term 1 - |<a^, b^>|^2 (beliefpropagation.jl:261-267) and max |a - b|.
template <typename Ta, typename Tb>
__global__ void k_diff_slots(DevTopo tp, const Ta *a, const Tb *b, double *res_fid,
                             double *res_abs) {
    const int q = blockIdx.x;
    if (q >= tp.nports) return;
    const long long mo = tp.msg_off[q];
    const int len = (int)(tp.msg_off[q + 1] - mo);
    double ab = 0.0, aa = 0.0, bb = 0.0, mx = 0.0;
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
        const double x = (double)a[mo + i], y = (double)b[mo + i];
        ab += x * y;
        aa += x * x;
        bb += y * y;
        const double d = fabs(x - y);
        mx = (d > mx || d != d) ? d : mx;
    }
    ab = wave_sum(ab);
    aa = wave_sum(aa);
    bb = wave_sum(bb);
    mx = wave_max(mx);
    if (threadIdx.x == 0) {
        const double c = ab / sqrt(aa * bb);
        res_fid[q] = 1.0 - c * c;
        res_abs[q] = mx;
    }
}

// BP log-norm per network (one workgroup per network): Neumaier-compensated (NeuSum)
// sum log|Z_v| - sum log|Z_e| with the reference's complex promotion and -Inf rule
// (messagecache.jl:185-201).

__global__ __launch_bounds__(BP_BLOCK) void k_lognorm_nets(DevTopo tp, const int *vbeg,
                                                            const int *pbeg, const double *zv,
                                                            const double *ze, double *out_re,
                                                            double *out_im) {
    __shared__ double ss[BP_BLOCK], cc[BP_BLOCK];
    __shared__ int cnt[4];
    const int n = blockIdx.x;
    if (threadIdx.x < 4) cnt[threadIdx.x] = 0;
    __syncthreads();
    NeuSum acc;
    int negv = 0, nege = 0, zv0 = 0, ze0 = 0;
    for (int u = vbeg[n] + threadIdx.x; u < vbeg[n + 1]; u += blockDim.x) {
        const double x = zv[u];
        if (x == 0.0) { zv0 = 1; continue; }
        if (x < 0.0) ++negv;
        acc.add(log(fabs(x)));
    }
    for (int q = pbeg[n] + threadIdx.x; q < pbeg[n + 1]; q += blockDim.x) {
        if (q > tp.port_rev[q]) continue;
        const double x = ze[tp.port_edge[q]];
        if (x == 0.0) { ze0 = 1; continue; }
        if (x < 0.0) ++nege;
        acc.add(-log(fabs(x)));
    }
    ss[threadIdx.x] = acc.s;
    cc[threadIdx.x] = acc.c;
    atomicAdd(&cnt[0], negv);
    atomicAdd(&cnt[1], nege);
    atomicOr(&cnt[2], zv0);
    atomicOr(&cnt[3], ze0);
    __syncthreads();
    if (threadIdx.x == 0) {
        NeuSum t;
        for (int i = 0; i < (int)blockDim.x; ++i) t.add(ss[i]);
        double c = 0.0;
        for (int i = 0; i < (int)blockDim.x; ++i) c += cc[i];
        double re = t.s + (t.c + c);
        double im = M_PI * (double)(cnt[0] - cnt[1]);
        if (cnt[3]) { re = -INFINITY; im = 0.0; }
        else if (cnt[2]) re = -INFINITY;
        out_re[n] = re;
        out_im[n] = im;
    }
}

// ------------------------------------------------------------------ context

// Message slots run by the generic kernel, grouped by the shape of their owning vertex
// (physical dimension d, largest bond dimension D; k is the largest degree in the bucket).
// Each bucket is one launch with its own LDS / workspace sizing, workgroup size `bs` and
// storage choice `lds`.  Degrees share a bucket: splitting them measured slower (more,
// smaller launches per sweep).
struct GenBucket {
    int d = 0, k = 0, maxD = 0;
    long long max_site = 0;
    int max_mats = 0;
    std::vector<int> slots;
    int *d_slots = nullptr;
    int bs = BP_BLOCK;  // 128, 256 or 512 threads
    int lds = 1;        // intermediates: 0 global workspace, 1 LDS up to 64 KiB (more
                        // workgroups per CU), 2 LDS up to 160 KiB
    std::string key;
};

// One autotuner decision.  Generic buckets: workgroup size and storage.  Vertex mask
// entries: whether networks of largest bond dimension D run on the vertex kernels.
struct TuneEntry {
    int block = BP_BLOCK, lds = 1, vertex = 1;
    double us = 0.0, us_alt = 0.0;
};

struct bp_context {
    int device = 0;
    int dtype = BP_F64;
    int path = BP_PATH_AUTO;
    hipStream_t own_stream = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    double last_ms = 0.0;

    // topology (host)
    int nv = 0, ne = 0, nports = 0, nnet = 0, plain = 0;
    std::vector<int> port_off, port_nbr, port_edge, port_rev, edge_dim, phys_dim, owner, vnet;
    std::vector<int> vbeg, pbeg;
    std::vector<long long> site_off, msg_off;
    long long nsite = 0, nmsg = 0;
    int maxD = 0;
    long long max_site = 0;
    int max_mats = 0;
    bool have_topo = false, have_sites = false, have_msgs = false;

    // topology (device)
    int *d_port_off = nullptr, *d_port_nbr = nullptr, *d_port_edge = nullptr,
        *d_port_rev = nullptr, *d_edge_dim = nullptr, *d_phys_dim = nullptr, *d_owner = nullptr,
        *d_vnet = nullptr, *d_vbeg = nullptr, *d_pbeg = nullptr;
    long long *d_site_off = nullptr, *d_msg_off = nullptr;

    // data
    double *d_sites64 = nullptr;  // FP64 master copy
    void *d_sites = nullptr;      // working dtype copy (FP32 for F32/BF16)
    float *d_sites_bf = nullptr;  // BF16-rounded FP32 copy (BF16 contexts, generic kernel)
    void *d_msg[2] = {nullptr, nullptr};
    int cur = 0;
    double *d_msg64[2] = {nullptr, nullptr};  // FP64 polish buffers
    int cur64 = 0;
    bool fp64_active = false;  // messages currently live in d_msg64[cur64]
    double *d_res_fid = nullptr, *d_res_abs = nullptr;
    double *d_zv = nullptr, *d_ze = nullptr, *d_lre = nullptr, *d_lim = nullptr;
    int *d_done = nullptr, *d_iter = nullptr, *d_at = nullptr;
    double *d_delta = nullptr;
    unsigned long long *d_fold_acc = nullptr;  // per network (folded residual reduction)
    char *h_state = nullptr;                    // pinned copy of the run state (bp_run)
    size_t h_state_bytes = 0;
    int fold_n = 0;                             // networks d_fold_acc holds
    int fold_on = 1;                            // option "fold_bookkeeping"
    int normalize = 1;                          // option "normalize" (0: SimpleMessageUpdate(normalize = false))
    void *d_ws = nullptr;
    size_t ws_bytes = 0;

    // vertex path plan
    VertexPlan vplan;
    bool vplan_ok = false;
    unsigned vx_mask = 0x1FFFEu;  // bit D: networks of largest bond dimension D may use it
    int vx_variant = -1;        // D <= 8 vertex sweep kernel (0: two LDS buffers per wave,
                                // 1: one buffer per wave, two waves per SIMD; -1: the faster
                                // one at 32x32 D=8: 1 in FP32 with d >= 2, else 0)
    int v16_waves = 16;         // waves per slice workgroup of the bond-dimension-16 kernel
    int vx_small = 1;           // option "small_bond": FP32 networks of bond dimension <= 4 on
                                // the small-bond kernels (vertex4.h), not padded to 8

    // FP64 D <= 8 vertex kernel (vertex64.h): the sweeps of FP64 contexts and the FP64
    // polish of every context.  Built on first use for the current topology and mask.
    Vertex64Plan v64;
    Vertex64LPlan v64l;         // bond dimensions 9..16 (vertex64l.h), when the D <= 8 plan
    int v64_kind = 0;           // declines: 1 the D <= 8 kernel, 2 the large-bond one
    bool v64_tried = false;     // plan attempted (v64_plan_ok tells the outcome)
    bool v64_plan_ok = false;
    bool v64_sites = false;     // padded FP64 slices match the current sites
    bool v64_ok = false;        // plan and slices ready

    // cached sweep graphs (16 sweeps + bookkeeping), one per starting buffer parity; valid
    // while graph_epoch equals the process-wide plan epoch (which moves whenever a captured
    // pointer or kernel choice may have changed)
    hipGraphExec_t gexec[2] = {nullptr, nullptr};
    // a run's last, shorter chunk (maxiter % 16 sweeps), captured by bp_prepare_run only
    hipGraphExec_t gtail[2] = {nullptr, nullptr};
    int gtail_n[2] = {0, 0};
    unsigned graph_epoch = 0;
    RunParams *d_runp = nullptr;

    // generic-kernel buckets of the resolved path (rebuilt when the path or plan changes)
    std::vector<GenBucket> gb;
    int gb_path = -1;
    std::map<std::string, TuneEntry> tune;  // autotuner choices (cache)
    void *d_vsites = nullptr;  // site tensors re-laid out for the vertex kernel

    // adjoint scratch
    AdjointScratch adj;

    size_t elt() const {
        return dtype == BP_C128 ? sizeof(cplx) : dtype == BP_F64 ? sizeof(double) : sizeof(float);
    }
    // doubles per element in host buffers, FP64 site copies and scalars (2: complex)
    int hw() const { return dtype == BP_C128 ? 2 : 1; }
    bool is_cplx() const { return dtype == BP_C128; }

    DevTopo topo() const {
        DevTopo t;
        t.port_off = d_port_off;
        t.port_nbr = d_port_nbr;
        t.port_edge = d_port_edge;
        t.port_rev = d_port_rev;
        t.edge_dim = d_edge_dim;
        t.phys_dim = d_phys_dim;
        t.owner = d_owner;
        t.vnet = d_vnet;
        t.site_off = d_site_off;
        t.msg_off = d_msg_off;
        t.nv = nv;
        t.nports = nports;
        t.plain = plain;
        return t;
    }
};

template <typename T>
static int upload(T **dst, const std::vector<T> &v) {
    if (*dst) HIPCHK(hipFree(*dst));
    *dst = nullptr;
    if (v.empty()) return BP_OK;
    HIPCHK(hipMalloc((void **)dst, sizeof(T) * v.size()));
    HIPCHK(hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return BP_OK;
}

static int dev_alloc(void **p, size_t bytes) {
    bump_epoch();
    if (*p) HIPCHK(hipFree(*p));
    *p = nullptr;
    if (bytes == 0) bytes = 16;
    HIPCHK(hipMalloc(p, bytes));
    return BP_OK;
}

// Measured at 32x32 D=8 (100 back-to-back sweeps, us per sweep), FP32 variants 0 / 1 / 4 / 5:
// d=1 23.5 / 22.4 / 23.0 / 22.7, d=2 40.8 / 34.4 / 32.8 / 31.8, d=3 73.9 / 54.1 / 54.1 / 51.2,
// d=4 75.7 / 60.5 / 57.8 / 57.0; BF16 d=2 variants 0 / 1: 33.2 / 30.5.
static int vx_variant_of(const bp_context *c) {
    if (c->vx_variant >= 0) return c->vx_variant;
    if (!c->vplan_ok) return -1;  // no vertex plan: no D <= 8 sweep kernel is chosen yet
    if (c->dtype == BP_F32) return c->vplan.d >= 2 ? 5 : 1;
    return 1;
}

static size_t run_state_bytes(int nnet) { return (sizeof(double) + 3 * sizeof(int)) * nnet; }

static void graphs_release(bp_context *c) {
    for (auto &g : c->gexec) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    for (int i = 0; i < 2; ++i) {
        if (c->gtail[i]) (void)hipGraphExecDestroy(c->gtail[i]);
        c->gtail[i] = nullptr;
        c->gtail_n[i] = 0;
    }
    c->graph_epoch = 0;
}

static void free_all(bp_context *c) {
    graphs_release(c);
    void *ptrs[] = {c->d_port_off, c->d_port_nbr, c->d_port_edge, c->d_port_rev, c->d_edge_dim,
                    c->d_phys_dim, c->d_owner, c->d_vnet, c->d_vbeg, c->d_pbeg, c->d_site_off,
                    c->d_msg_off, c->d_sites64, c->d_sites, c->d_msg[0], c->d_msg[1],
                    c->d_msg64[0], c->d_msg64[1], c->d_res_fid, c->d_res_abs, c->d_zv, c->d_ze,
                    c->d_lre, c->d_lim, c->d_delta, c->d_ws,
                    c->d_vsites, c->d_sites_bf, c->d_runp, c->d_fold_acc};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    c->adj.release();
    vertex_plan_release(&c->vplan);
    vertex64_release(&c->v64);
    vertex64l_release(&c->v64l);
    c->v64_kind = 0;
    c->v64_tried = c->v64_plan_ok = c->v64_sites = c->v64_ok = false;
    for (auto &b : c->gb)
        if (b.d_slots) (void)hipFree(b.d_slots);
    c->gb.clear();
    c->gb_path = -1;
}

// ------------------------------------------------------------------ generic launches

struct GenLaunch {
    int mats_elems, out_elems, red_elems;
    long long buf_elems;
    int use_lds;
    size_t lds_bytes;
    int grid;
    long long ws_stride;
};

static GenLaunch gen_launch(bp_context *c, int nitems, size_t esz) {
    GenLaunch g;
    g.mats_elems = std::max(c->max_mats, 1);
    g.out_elems = std::max(c->maxD * c->maxD, 1);
    g.red_elems = (esz == sizeof(cplx) ? 2 : 1) * std::max(c->maxD * c->maxD, BP_BLOCK);
    g.buf_elems = std::max(c->max_site, 1LL);
    size_t base = sizeof(double) * g.red_elems + esz * (g.mats_elems + g.out_elems) + 16;
    size_t with = base + 2 * esz * g.buf_elems;
    g.use_lds = with <= 64 * 1024 ? 1 : 0;
    g.lds_bytes = g.use_lds ? with : base;
    g.grid = std::max(1, std::min(nitems, g.use_lds ? 65535 : 2048));
    g.ws_stride = 2 * g.buf_elems;
    return g;
}

static int ensure_ws(bp_context *c, const GenLaunch &g, size_t esz) {
    if (g.use_lds) return BP_OK;
    size_t need = esz * (size_t)g.ws_stride * (size_t)g.grid;
    if (need > c->ws_bytes) {
        RETCHK(dev_alloc(&c->d_ws, need));
        c->ws_bytes = need;
    }
    return BP_OK;
}

// Work items [q0, q1): message slots, or entries of `list` (a bucket's slots) when given.
template <typename T>
static int launch_update_generic(bp_context *c, const T *sites, const T *mold, T *mnew,
                                 const NetCtl &ctl, int q0, int q1, const int *list = nullptr,
                                 int flags = 0) {
    if (q1 <= q0) return BP_OK;
    GenLaunch g = gen_launch(c, q1 - q0, sizeof(T));
    RETCHK(ensure_ws(c, g, sizeof(T)));
    if constexpr (std::is_same<T, cplx>::value)
        hipLaunchKernelGGL(k_update_cplx<BP_BLOCK>, dim3(g.grid), dim3(BP_BLOCK), g.lds_bytes,
                           c->stream, c->topo(), sites, mold, mnew, ctl, c->d_res_fid,
                           c->d_res_abs, q0, q1, list, g.mats_elems, g.out_elems, g.red_elems,
                           g.buf_elems, g.use_lds, (T *)c->d_ws, g.ws_stride, flags);
    else
        hipLaunchKernelGGL(k_update_generic<T>, dim3(g.grid), dim3(BP_BLOCK), g.lds_bytes,
                           c->stream, c->topo(), sites, mold, mnew, ctl, c->d_res_fid,
                           c->d_res_abs, q0, q1, list, g.mats_elems, g.out_elems, g.red_elems,
                           g.buf_elems, g.use_lds, (T *)c->d_ws, g.ws_stride, flags);
    HIPCHK(hipGetLastError());
    return BP_OK;
}

// ---- generic buckets

static const char *dtype_name(int dt) {
    return dt == BP_F64 ? "f64" : dt == BP_F32 ? "f32" : dt == BP_C128 ? "c128" : "bf16";
}

static constexpr size_t GEN_LDS_MAX = 160 * 1024;  // LDS per workgroup on gfx950
static constexpr size_t GEN_LDS_DEFAULT = 64 * 1024;

static GenLaunch gen_launch_bucket(const GenBucket &b, size_t esz) {
    GenLaunch g;
    g.mats_elems = std::max(b.max_mats, 1);
    g.out_elems = std::max(b.maxD * b.maxD, 1);
    // complex kernels (16-byte elements) keep interleaved (re, im) partial sums
    g.red_elems = (esz == sizeof(cplx) ? 2 : 1) * std::max(b.maxD * b.maxD, b.bs);
    g.buf_elems = std::max(b.max_site, 1LL);
    const size_t base = sizeof(double) * g.red_elems + esz * (g.mats_elems + g.out_elems) + 16;
    const size_t with = base + 2 * esz * g.buf_elems;
    g.use_lds = (b.lds == 2 && with <= GEN_LDS_MAX) || (b.lds == 1 && with <= GEN_LDS_DEFAULT);
    g.lds_bytes = g.use_lds ? with : base;
    g.grid = std::max(1, std::min((int)b.slots.size(), g.use_lds ? 65535 : 2048));
    g.ws_stride = 2 * g.buf_elems;
    return g;
}

static bool bucket_lds_fits(const GenBucket &b, size_t esz, int mode) {
    GenBucket t = b;
    t.lds = mode;
    return gen_launch_bucket(t, esz).use_lds == 1;
}

// Raise the dynamic-LDS limit of every generic-kernel instance once per process.
static int gen_set_lds_limits() {
    static int done = 0;
    if (done) return BP_OK;
    const void *fns[9] = {(const void *)k_update_generic<float, 128>,
                          (const void *)k_update_generic<float, 256>,
                          (const void *)k_update_generic<float, 512>,
                          (const void *)k_update_generic<double, 128>,
                          (const void *)k_update_generic<double, 256>,
                          (const void *)k_update_generic<double, 512>,
                          (const void *)k_update_cplx<128>,
                          (const void *)k_update_cplx<256>,
                          (const void *)k_update_cplx<512>};
    for (const void *f : fns)
        HIPCHK(hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)GEN_LDS_MAX));
    done = 1;
    return BP_OK;
}

static std::string bucket_key(int dtype, const GenBucket &b) {
    std::ostringstream o;
    o << "gen|" << dtype_name(dtype) << "|d" << b.d << "|D" << b.maxD;
    return o.str();
}

static void free_buckets(bp_context *c) {
    bump_epoch();
    for (auto &b : c->gb)
        if (b.d_slots) (void)hipFree(b.d_slots);
    c->gb.clear();
    c->gb_path = -1;
}

// Buckets of the generic slots of `path` (all slots, or the slots outside the vertex bucket).
static int build_buckets(bp_context *c, int path) {
    if (c->gb_path == path) return BP_OK;
    free_buckets(c);
    if (path == BP_PATH_VERTEX) { c->gb_path = path; return BP_OK; }
    std::vector<int> all;
    const std::vector<int> *slots = &c->vplan.gen_slots;
    if (path != BP_PATH_MIXED) {
        all.resize(c->nports);
        for (int q = 0; q < c->nports; ++q) all[q] = q;
        slots = &all;
    }
    std::map<std::pair<int, int>, int> index;
    for (int q : *slots) {
        const int u = c->owner[q];
        const int k = c->port_off[u + 1] - c->port_off[u];
        int maxD = 1, mats = 0;
        for (int p = c->port_off[u]; p < c->port_off[u + 1]; ++p) {
            const int D = c->edge_dim[c->port_edge[p]];
            maxD = std::max(maxD, D);
            mats += c->plain ? D : D * D;
        }
        const auto key = std::make_pair(c->plain ? 1 : c->phys_dim[u], maxD);
        auto it = index.find(key);
        if (it == index.end()) {
            GenBucket b;
            b.d = key.first;
            b.maxD = maxD;
            it = index.emplace(key, (int)c->gb.size()).first;
            c->gb.push_back(b);
        }
        GenBucket &b = c->gb[it->second];
        b.k = std::max(b.k, k);
        b.slots.push_back(q);
        b.max_site = std::max(b.max_site, c->site_off[u + 1] - c->site_off[u]);
        b.max_mats = std::max(b.max_mats, mats);
    }
    for (auto &b : c->gb) {
        b.key = bucket_key(c->dtype, b);
        auto t = c->tune.find(b.key);
        if (t != c->tune.end()) {
            b.bs = t->second.block;
            b.lds = t->second.lds;
        }
        RETCHK(upload(&b.d_slots, b.slots));
    }
    c->gb_path = path;
    return BP_OK;
}

template <typename T>
static int launch_bucket(bp_context *c, const GenBucket &b, const T *sites, const T *mold,
                         T *mnew, const NetCtl &ctl, int flags) {
    if (b.slots.empty()) return BP_OK;
    const GenLaunch g = gen_launch_bucket(b, sizeof(T));
    const int n = (int)b.slots.size();
#define GEN_B(BS)                                                                               \
    if constexpr (std::is_same<T, cplx>::value)                                                 \
        hipLaunchKernelGGL((k_update_cplx<BS>), dim3(g.grid), dim3(BS), g.lds_bytes, c->stream,  \
                           c->topo(), sites, mold, mnew, ctl, c->d_res_fid, c->d_res_abs, 0, n, \
                           b.d_slots, g.mats_elems, g.out_elems, g.red_elems, g.buf_elems,      \
                           g.use_lds, (T *)c->d_ws, g.ws_stride, flags);                        \
    else                                                                                        \
        hipLaunchKernelGGL((k_update_generic<T, BS>), dim3(g.grid), dim3(BS), g.lds_bytes,     \
                           c->stream, c->topo(), sites, mold, mnew, ctl, c->d_res_fid,          \
                           c->d_res_abs, 0, n, b.d_slots, g.mats_elems, g.out_elems,            \
                           g.red_elems, g.buf_elems, g.use_lds, (T *)c->d_ws, g.ws_stride, flags)
    switch (b.bs) {
        case 128: GEN_B(128); break;
        case 512: GEN_B(512); break;
        default: GEN_B(256); break;
    }
#undef GEN_B
    HIPCHK(hipGetLastError());
    return BP_OK;
}

// Workspace for every bucket (one buffer, buckets run in stream order).
static int ensure_bucket_ws(bp_context *c, size_t esz) {
    size_t need = 0;
    for (const auto &b : c->gb) {
        const GenLaunch g = gen_launch_bucket(b, esz);
        if (!g.use_lds) need = std::max(need, esz * (size_t)g.ws_stride * (size_t)g.grid);
    }
    if (need > c->ws_bytes) {
        RETCHK(dev_alloc(&c->d_ws, need));
        c->ws_bytes = need;
    }
    return BP_OK;
}

template <typename T>
static int launch_buckets(bp_context *c, const T *sites, const T *mold, T *mnew,
                          const NetCtl &ctl, int flags) {
    for (const auto &b : c->gb) RETCHK(launch_bucket<T>(c, b, sites, mold, mnew, ctl, flags));
    return BP_OK;
}

template <typename T>
static int launch_vertex_scalars(bp_context *c, const T *sites, const T *msgs) {
    GenLaunch g = gen_launch(c, c->nv, sizeof(T));
    RETCHK(ensure_ws(c, g, sizeof(T)));
    if constexpr (std::is_same<T, cplx>::value)
        hipLaunchKernelGGL(k_vertex_scalar_cplx, dim3(g.grid), dim3(BP_BLOCK), g.lds_bytes,
                           c->stream, c->topo(), sites, msgs, c->d_zv, g.mats_elems, g.out_elems,
                           g.red_elems, g.buf_elems, g.use_lds, (T *)c->d_ws, g.ws_stride);
    else
        hipLaunchKernelGGL(k_vertex_scalar_generic<T>, dim3(g.grid), dim3(BP_BLOCK), g.lds_bytes,
                           c->stream, c->topo(), sites, msgs, c->d_zv, g.mats_elems, g.out_elems,
                           g.red_elems, g.buf_elems, g.use_lds, (T *)c->d_ws, g.ws_stride);
    HIPCHK(hipGetLastError());
    return BP_OK;
}

template <typename T>
static int launch_edge_scalars(bp_context *c, const T *msgs) {
    const int blocks = (c->nports + 255) / 256;
    if (blocks > 0) {
        if constexpr (std::is_same<T, cplx>::value)
            hipLaunchKernelGGL(k_edge_scalars_cplx, dim3(blocks), dim3(256), 0, c->stream,
                               c->topo(), msgs, c->d_ze);
        else
            hipLaunchKernelGGL(k_edge_scalars<T>, dim3(blocks), dim3(256), 0, c->stream,
                               c->topo(), msgs, c->d_ze);
    }
    HIPCHK(hipGetLastError());
    return BP_OK;
}

// ------------------------------------------------------------------ path selection

// Resolved dispatch: every network on the generic kernel, every network on the vertex
// kernel, or the vertex bucket plus the generic kernel for the remaining networks.


// The FP64 vertex plan for this context's networks (bond dimensions enabled: the context's
// vertex mask in FP64 contexts; 6..8 and 9..16 for the polish of FP32/BF16 contexts): the
// D <= 8 kernel (vertex64.h) when every network's largest bond dimension is 6..8, else the
// large-bond kernel (vertex64l.h) when it is 9..16 for every network, with the padded FP64
// slices when sites are present.  Allocates: never called inside a graph capture
// (prepare_sweep runs first).
static int ensure_v64(bp_context *c) {
    if (!c->have_topo) return BP_OK;
    if (!c->v64_tried) {
        const unsigned mask = c->dtype == BP_F64 ? c->vx_mask : 0x1C0u;
        c->v64_kind = vertex64_plan_build(&c->v64, c->plain, c->nv, c->port_off.data(),
                                          c->port_edge.data(), c->edge_dim.data(),
                                          c->phys_dim.data(), c->port_rev.data(),
                                          c->vnet.data(), mask) == 0 &&
                              vertex64_set_lds_limit(&c->v64) == 0
                          ? 1 : 0;
        if (!c->v64_kind) {
            const unsigned lmask = c->dtype == BP_F64 ? c->vx_mask : 0x1FE00u;
            if (vertex64l_plan_build(&c->v64l, c->plain, c->nv, c->port_off.data(),
                                     c->port_edge.data(), c->edge_dim.data(),
                                     c->phys_dim.data(), c->port_rev.data(), c->vnet.data(),
                                     lmask) == 0)
                c->v64_kind = 2;
        }
        c->v64_plan_ok = c->v64_kind != 0;
        c->v64_tried = true;
        c->v64_sites = false;
        bump_epoch();
    }
    if (c->v64_plan_ok && c->have_sites && !c->v64_sites) {
        const int rc = c->v64_kind == 1
                           ? vertex64_upload_sites(&c->v64, c->stream, c->d_sites64, c->d_site_off,
                                                   c->d_port_off, c->d_port_edge, c->d_edge_dim)
                           : vertex64l_upload_sites(&c->v64l, c->stream, c->d_sites64,
                                                    c->d_site_off, c->d_port_off,
                                                    c->d_port_edge, c->d_edge_dim);
        if (rc != 0) return set_err(BP_ERR_HIP, "FP64 vertex site layout failed");
        HIPCHK(hipStreamSynchronize(c->stream));
        c->v64_sites = true;
        bump_epoch();
    }
    c->v64_ok = c->v64_plan_ok && c->v64_sites;
    return BP_OK;
}

static void v64_invalidate(bp_context *c) {
    vertex64_release(&c->v64);
    vertex64l_release(&c->v64l);
    c->v64_kind = 0;
    c->v64_tried = c->v64_plan_ok = c->v64_sites = c->v64_ok = false;
}

// One FP64 sweep on the context's FP64 vertex kernel (ensure_v64 has made it ready).
static int v64_sweep(bp_context *c, const double *src, double *dst, const NetCtl &ctl) {
    const int rc = c->v64_kind == 1
                       ? vertex64_sweep(&c->v64, c->stream, c->d_msg_off, src, dst, ctl,
                                        c->d_res_fid, c->d_res_abs)
                       : vertex64l_sweep(&c->v64l, c->stream, c->d_msg_off, src, dst, ctl,
                                         c->d_res_fid, c->d_res_abs);
    return rc != 0 ? set_err(BP_ERR_HIP, "FP64 vertex sweep launch failed") : BP_OK;
}

// Flags of the generic kernel for this context (BF16 operands, no normalisation).
static int gen_flags(const bp_context *c, bool bf16) {
    return (bf16 ? GEN_FLAG_BF16 : 0) | (c->normalize ? 0 : GEN_FLAG_NO_NORMALIZE);
}

static int resolve_path(bp_context *c) {
    // the vertex kernels always normalise: unnormalised updates run on the generic kernel
    if (!c->normalize || c->is_cplx()) return c->path == BP_PATH_VERTEX ? -1 : BP_PATH_GENERIC;
    if (c->dtype == BP_F64) {  // the FP64 vertex kernel covers every network or none
        if (c->path != BP_PATH_GENERIC && c->v64_ok) return BP_PATH_VERTEX;
        return c->path == BP_PATH_VERTEX ? -1 : BP_PATH_GENERIC;
    }
    if (c->path == BP_PATH_GENERIC || !c->vplan_ok) return c->path == BP_PATH_VERTEX ? -1 : BP_PATH_GENERIC;
    if (c->vplan.all) return BP_PATH_VERTEX;
    // AUTO: vertex kernel for its bucket, generic kernel for the rest
    return c->path == BP_PATH_VERTEX ? -1 : BP_PATH_MIXED;
}

// Allocate whatever workspace the next sweeps need, before any graph capture (no
// allocation may happen inside a captured region).
static int prepare_sweep(bp_context *c) {
    if (c->dtype == BP_F64) RETCHK(ensure_v64(c));
    const int path = resolve_path(c);
    if (path < 0) return BP_OK;  // reported by sweep_once
    RETCHK(gen_set_lds_limits());
    RETCHK(build_buckets(c, path));
    if (c->vplan_ok && c->fold_n != c->nnet) {
        // folded residual accumulators, FOLD_BUFS rotating buffers (outside any graph
        // capture); every run starts by clearing them (k_reset_nets)
        const size_t nb = sizeof(unsigned long long) * FOLD_BUFS * FOLD_SLOTS * (size_t)c->nnet;
        RETCHK(dev_alloc((void **)&c->d_fold_acc, nb));
        HIPCHK(hipMemset(c->d_fold_acc, 0, nb));
        c->fold_n = c->nnet;
    }
    if (path != BP_PATH_VERTEX)
        RETCHK(ensure_bucket_ws(c, c->elt()));
    return BP_OK;
}

// Convergence control for the sweep kernels (disabled: plain sweeps).
static NetCtl make_ctl(bp_context *c, bool enabled) {
    NetCtl k;
    k.done = c->d_done;
    k.enabled = enabled ? 1 : 0;
    return k;
}

// The same bookkeeping after a folded sweep: the per-network maxima are already in acc
// (FOLD_SLOTS bit patterns of values >= 0 or NaN per network, which order like the values);
// one thread per network, the slots reset for the next sweep.
__global__ void k_reduce_folded(unsigned long long *acc, const RunParams *rp, int nnet,
                                int *done, int *iter, double *delta, int *at) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nnet) return;
    unsigned long long b = 0ull;
    for (int k = 0; k < FOLD_SLOTS; ++k) {
        const unsigned long long v = acc[(long long)n * FOLD_SLOTS + k];
        b = v > b ? v : b;
        acc[(long long)n * FOLD_SLOTS + k] = 0ull;
    }
    if (done[n]) return;
    const double m = __longlong_as_double((long long)b);
    const int it = iter[n] + 1;
    iter[n] = it;
    delta[n] = m;
    if (rp->tol >= 0.0 && m < rp->tol) {
        done[n] = 1;
        at[n] = it;
    } else if (rp->maxiter > 0 && it >= rp->maxiter) {
        done[n] = 1;
    }
}

// After the last folded sweep (run index k) of a sequence: its stopping rule, and the rest
// of the sequence's StopAfterIteration.
static int launch_reduce_folded(bp_context *c, int k) {
    unsigned long long *acc =
        c->d_fold_acc + (size_t)(k % FOLD_BUFS) * FOLD_SLOTS * (size_t)c->fold_n;
    hipLaunchKernelGGL(k_reduce_folded, dim3((c->nnet + 63) / 64), dim3(64), 0, c->stream,
                       acc, (const RunParams *)c->d_runp, c->nnet, c->d_done,
                       c->d_iter, c->d_delta, c->d_at);
    HIPCHK(hipGetLastError());
    return BP_OK;
}

// Per-network bookkeeping after a sweep: one workgroup of 1024 threads per network.
static int launch_reduce(bp_context *c) {
    hipLaunchKernelGGL(k_reduce_nets, dim3(c->nnet), dim3(1024), 0, c->stream, c->d_pbeg,
                       c->d_res_fid, c->d_res_abs, c->d_runp, c->d_done, c->d_iter, c->d_delta,
                       c->d_at);
    HIPCHK(hipGetLastError());
    return BP_OK;
}

// Site operands of the generic message kernel: FP32, or BF16-rounded FP32 in BF16 contexts
// (scalars and gradients keep the unrounded FP32 copy).
static const float *generic_sites32(const bp_context *c) {
    return (const float *)(c->dtype == BP_BF16 ? c->d_sites_bf : c->d_sites);
}

// One Jacobi sweep on the working dtype: reads d_msg[cur], writes d_msg[cur^1].
// Whether the residual reduction can ride on the sweep kernel: the D <= 8 four-port kernel
// alone updates every message of every network.
static bool fold_possible(const bp_context *c) {
    return c->vplan_ok && c->vplan.all && c->vplan.nv_k3 == 0 && c->vplan.nv16 == 0 &&
           c->vplan.ns_total() == 0 &&
           c->path != BP_PATH_GENERIC && c->d_fold_acc && c->fold_n == c->nnet &&
           c->vplan.nv > 0 && c->vplan.d_lead && c->fold_on;
}

// One Jacobi sweep; with `folded` non-null and the bookkeeping foldable, the sweep kernel
// also performs it (*folded = true): sweep `k` of the run folds its residuals into buffer
// k % FOLD_BUFS and, with `apply_prev`, first applies the stopping rule of sweep k - 1; the
// sequence's last sweep is followed by launch_reduce_folded(c, k).
static int sweep_once(bp_context *c, const NetCtl &done, bool *folded = nullptr, int k = 0,
                      bool apply_prev = false) {
    const int path = resolve_path(c);
    if (path < 0) return set_err(BP_ERR_UNSUPPORTED, "vertex path not available: %s",
                                 c->vplan.why.c_str());
    if (c->gb_path != path)  // normally done by prepare_sweep, outside any graph capture
        RETCHK(prepare_sweep(c));
    void *src = c->d_msg[c->cur], *dst = c->d_msg[c->cur ^ 1];
    SweepFold fold{};
    if (folded) *folded = false;
    if (folded && done.enabled && path == BP_PATH_VERTEX && fold_possible(c)) {
        const size_t stride = FOLD_SLOTS * (size_t)c->fold_n;
        fold.acc = c->d_fold_acc + (size_t)(k % FOLD_BUFS) * stride;
        fold.rp = c->d_runp;
        fold.prev = apply_prev ? c->d_fold_acc + (size_t)((k + FOLD_BUFS - 1) % FOLD_BUFS) * stride
                               : nullptr;
        fold.clear = c->d_fold_acc + (size_t)((k + 1) % FOLD_BUFS) * stride;
        fold.lead = c->vplan.d_lead;
        fold.iter = c->d_iter;
        fold.delta = c->d_delta;
        fold.at = c->d_at;
        *folded = true;
    }
    if (path == BP_PATH_VERTEX && c->dtype == BP_F64) {
        RETCHK(v64_sweep(c, (const double *)src, (double *)dst, done));
    } else if (path == BP_PATH_VERTEX || path == BP_PATH_MIXED) {
        RETCHK(vertex_sweep(&c->vplan, c->stream, c->topo(), c->d_vsites, src, dst, done,
                            c->d_res_fid, c->d_res_abs, fold));
        RETCHK(vertex_k3_sweep(&c->vplan, c->stream, c->topo(), src, dst, done,
                               c->d_res_fid, c->d_res_abs));
        RETCHK(vertex16_sweep(&c->vplan, c->stream, c->topo(), src, dst, done, c->d_res_fid,
                              c->d_res_abs));
        if (vertex_s_sweep(&c->vplan, c->stream, c->topo(), src, dst, done, c->d_res_fid,
                           c->d_res_abs) != 0)
            return set_err(BP_ERR_HIP, "small-bond sweep launch failed");
        HIPCHK(hipGetLastError());
    }
    if (path != BP_PATH_VERTEX) {
        if (c->is_cplx())
            RETCHK(launch_buckets<cplx>(c, (const cplx *)c->d_sites, (const cplx *)src,
                                        (cplx *)dst, done, gen_flags(c, false)));
        else if (c->dtype == BP_F64)
            RETCHK(launch_buckets<double>(c, (const double *)c->d_sites, (const double *)src,
                                          (double *)dst, done, gen_flags(c, false)));
        else
            RETCHK(launch_buckets<float>(c, generic_sites32(c), (const float *)src, (float *)dst,
                                         done, gen_flags(c, c->dtype == BP_BF16)));
    }
    c->cur ^= 1;
    return BP_OK;
}

// Padded site layouts of the vertex buckets (bond dimension <= 8 with its degree <= 3 copies,
// and the large-bond bucket: 9..12 in FP32, 9..16 in BF16).
static int upload_vertex_sites(bp_context *c) {
    bump_epoch();
    if (vertex_upload_sites(&c->vplan, c->stream, c->d_sites64, &c->d_vsites, c->d_site_off,
                            c->d_port_off, c->d_port_edge, c->d_edge_dim) != BP_OK ||
        vertex_k3_upload_sites(&c->vplan, c->stream, c->d_sites64, c->d_site_off,
                               c->d_port_off, c->d_port_edge, c->d_edge_dim) != BP_OK ||
        vertex16_upload_sites(&c->vplan, c->stream, c->d_sites64, c->d_site_off, c->d_port_off,
                              c->d_port_edge, c->d_edge_dim) != BP_OK ||
        vertex_s_upload_sites(&c->vplan, c->stream, c->d_sites64, c->d_site_off,
                              c->d_port_off, c->d_port_edge, c->d_edge_dim) != 0)
        return set_err(BP_ERR_HIP, "vertex-path site layout failed");
    HIPCHK(hipStreamSynchronize(c->stream));
    return BP_OK;
}

// Vertex-path plan for the current mask (may decline; the reason is kept for bp_set_path
// errors), with the padded site layout when sites are present; generic buckets follow.
static int rebuild_vertex_plan(bp_context *c) {
    bump_epoch();
    v64_invalidate(c);
    vertex_plan_release(&c->vplan);
    free_buckets(c);
    if (c->is_cplx()) {  // complex arithmetic runs on the generic kernel only
        c->vplan_ok = false;
        c->vplan.why = "complex contexts run on the generic kernel";
        return BP_OK;
    }
    c->vplan_ok = vertex_plan_build(&c->vplan, c->dtype, c->plain, c->nv, c->port_off.data(),
                                    c->port_edge.data(), c->edge_dim.data(), c->phys_dim.data(),
                                    c->site_off.data(), c->msg_off.data(), c->port_rev.data(),
                                    c->vnet.data(), c->vx_mask, c->vx_small) == BP_OK;
    if (c->vplan_ok && (vertex_set_lds_limits(&c->vplan) != BP_OK ||
                        (c->vplan.nv16 > 0 && vertex16_set_lds_limit() != BP_OK))) {
        c->vplan_ok = false;
        c->vplan.why = "could not raise the LDS limit of the vertex kernel";
    }
    c->vplan.variant = vx_variant_of(c);
    c->vplan.v16_waves = c->v16_waves;
    if (c->vplan_ok && c->have_sites) RETCHK(upload_vertex_sites(c));
    return BP_OK;
}

// ------------------------------------------------------------------ C ABI

static int refresh_sites(bp_context *c);

// The device address of `scratch` on `device` (current), looked up once per device.
static int scratch_setup(int device) {
    if (device < 0 || device >= BP_SCRATCH_DEVICES)
        return set_err(BP_ERR_UNSUPPORTED, "device %d beyond the %d the library tracks", device,
                       BP_SCRATCH_DEVICES);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return BP_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_err(BP_ERR_HIP, "hipGetSymbolAddress(scratch) failed on device %d: %s",
                       device, hipGetErrorString(e));
    g_scratch_ptr[device] = (double *)p;
    return BP_OK;
}

extern "C" {

const char *bp_last_error(void) { return g_err.c_str(); }

int bp_version(void) { return 10000; }

int bp_create(bp_context **out, int device, int dtype) {
    if (!out) return set_err(BP_ERR_ARG, "out is NULL");
    if (dtype != BP_F64 && dtype != BP_F32 && dtype != BP_BF16 && dtype != BP_C128)
        return set_err(BP_ERR_ARG, "unknown dtype %d", dtype);
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev)
        return set_err(BP_ERR_ARG, "device %d out of range (%d devices)", device, ndev);
    HIPCHK(hipSetDevice(device));
    RETCHK(scratch_setup(device));
    bp_context *c = new bp_context();
    c->device = device;
    c->dtype = dtype;
    // FP64 contexts: the FP64 vertex kernels for largest bond dimensions 6..16 by default
    // (below 6 the zero padding to 8 costs more than the generic kernel saves)
    if (dtype == BP_F64) c->vx_mask = 0x1FFC0u;
    if (dtype == BP_C128) c->vx_mask = 0u;
    if (hipStreamCreateWithFlags(&c->own_stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess) {
        delete c;
        return set_err(BP_ERR_HIP, "stream/event creation failed");
    }
    c->stream = c->own_stream;
    if (dev_alloc((void **)&c->d_runp, sizeof(RunParams)) != BP_OK) {
        const std::string e = g_err;
        (void)bp_destroy(c);
        return set_err(BP_ERR_HIP, "%s", e.c_str());
    }
    *out = c;
    return BP_OK;
}

int bp_destroy(bp_context *c) {
    if (!c) return BP_OK;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    free_all(c);
    if (c->h_state) (void)hipHostFree(c->h_state);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->own_stream) (void)hipStreamDestroy(c->own_stream);
    delete c;
    return BP_OK;
}

int bp_set_stream(bp_context *c, void *s) {
    if (!c) return set_err(BP_ERR_ARG, "ctx is NULL");
    c->stream = s ? (hipStream_t)s : c->own_stream;
    return BP_OK;
}

int bp_set_topology(bp_context *c, int nv, int ne, const int *port_off, const int *port_nbr,
                    const int *port_edge, const int *port_rev, const int *edge_dim,
                    const int *phys_dim, const int *vertex_net, int nnet, int plain) {
    if (!c) return set_err(BP_ERR_ARG, "ctx is NULL");
    if (nv <= 0 || ne < 0 || !port_off) return set_err(BP_ERR_ARG, "empty topology");
    if (nnet <= 0) nnet = 1;
    HIPCHK(hipSetDevice(c->device));
    c->nv = nv;
    c->ne = ne;
    c->plain = plain ? 1 : 0;
    c->nnet = nnet;
    c->port_off.assign(port_off, port_off + nv + 1);
    if (c->port_off[0] != 0) return set_err(BP_ERR_ARG, "port_off[0] must be 0");
    for (int u = 0; u < nv; ++u) {
        const int k = c->port_off[u + 1] - c->port_off[u];
        if (k < 0) return set_err(BP_ERR_ARG, "port_off not monotone at %d", u);
        if (k + 1 > BP_MAXR)
            return set_err(BP_ERR_UNSUPPORTED, "vertex %d has degree %d > %d", u, k,
                           BP_MAXR - 1);
    }
    c->nports = c->port_off[nv];
    const int np = c->nports;
    c->port_nbr.assign(port_nbr, port_nbr + np);
    c->port_edge.assign(port_edge, port_edge + np);
    c->port_rev.assign(port_rev, port_rev + np);
    c->edge_dim.assign(edge_dim, edge_dim + ne);
    c->phys_dim.assign(nv, 1);
    if (!plain) {
        if (!phys_dim) return set_err(BP_ERR_ARG, "phys_dim is NULL for a norm network");
        c->phys_dim.assign(phys_dim, phys_dim + nv);
    }
    c->vnet.assign(nv, 0);
    if (vertex_net) c->vnet.assign(vertex_net, vertex_net + nv);
    // validation
    c->owner.assign(np, 0);
    for (int u = 0; u < nv; ++u)
        for (int q = c->port_off[u]; q < c->port_off[u + 1]; ++q) c->owner[q] = u;
    for (int q = 0; q < np; ++q) {
        const int w = c->port_nbr[q], r = c->port_rev[q], e = c->port_edge[q];
        if (w < 0 || w >= nv || r < 0 || r >= np || e < 0 || e >= ne)
            return set_err(BP_ERR_ARG, "port %d out of range", q);
        if (c->owner[r] != w || c->port_rev[r] != q || c->port_nbr[r] != c->owner[q] ||
            c->port_edge[r] != e)
            return set_err(BP_ERR_ARG, "port %d: inconsistent reverse port", q);
    }
    c->maxD = 0;
    for (int e = 0; e < ne; ++e) {
        if (c->edge_dim[e] <= 0) return set_err(BP_ERR_ARG, "edge %d has dim %d", e, c->edge_dim[e]);
        c->maxD = std::max(c->maxD, c->edge_dim[e]);
    }
    if (c->maxD > 64) return set_err(BP_ERR_UNSUPPORTED, "bond dimension %d > 64", c->maxD);
    // offsets
    c->site_off.assign(nv + 1, 0);
    c->max_site = 0;
    c->max_mats = 0;
    for (int u = 0; u < nv; ++u) {
        long long n = c->phys_dim[u];
        int mats = 0;
        for (int q = c->port_off[u]; q < c->port_off[u + 1]; ++q) {
            const int D = c->edge_dim[c->port_edge[q]];
            n *= D;
            mats += plain ? D : D * D;
        }
        c->site_off[u + 1] = c->site_off[u] + n;
        c->max_site = std::max(c->max_site, n);
        c->max_mats = std::max(c->max_mats, mats);
    }
    c->msg_off.assign(np + 1, 0);
    for (int q = 0; q < np; ++q) {
        const int D = c->edge_dim[c->port_edge[q]];
        c->msg_off[q + 1] = c->msg_off[q] + (plain ? D : (long long)D * D);
    }
    c->nsite = c->site_off[nv];
    c->nmsg = c->msg_off[np];
    // networks must be contiguous in vertex order
    c->vbeg.assign(nnet + 1, 0);
    c->pbeg.assign(nnet + 1, 0);
    for (int u = 0; u < nv; ++u) {
        const int n = c->vnet[u];
        if (n < 0 || n >= nnet) return set_err(BP_ERR_ARG, "vertex_net[%d]=%d out of range", u, n);
        if (u > 0 && n < c->vnet[u - 1])
            return set_err(BP_ERR_ARG, "networks must be contiguous and ordered in vertex order");
    }
    {
        int u = 0;
        for (int n = 0; n < nnet; ++n) {
            c->vbeg[n] = u;
            while (u < nv && c->vnet[u] == n) ++u;
        }
        c->vbeg[nnet] = nv;
        for (int n = 0; n <= nnet; ++n) c->pbeg[n] = c->port_off[c->vbeg[n]];
        for (int q = 0; q < np; ++q)
            if (c->vnet[c->port_nbr[q]] != c->vnet[c->owner[q]])
                return set_err(BP_ERR_ARG, "edge crosses networks at port %d", q);
    }
    RETCHK(upload(&c->d_port_off, c->port_off));
    RETCHK(upload(&c->d_port_nbr, c->port_nbr));
    RETCHK(upload(&c->d_port_edge, c->port_edge));
    RETCHK(upload(&c->d_port_rev, c->port_rev));
    RETCHK(upload(&c->d_edge_dim, c->edge_dim));
    RETCHK(upload(&c->d_phys_dim, c->phys_dim));
    RETCHK(upload(&c->d_owner, c->owner));
    RETCHK(upload(&c->d_vnet, c->vnet));
    RETCHK(upload(&c->d_vbeg, c->vbeg));
    RETCHK(upload(&c->d_pbeg, c->pbeg));
    RETCHK(upload(&c->d_site_off, c->site_off));
    RETCHK(upload(&c->d_msg_off, c->msg_off));
    const size_t es = c->elt();
    RETCHK(dev_alloc(&c->d_msg[0], es * c->nmsg));
    RETCHK(dev_alloc(&c->d_msg[1], es * c->nmsg));
    RETCHK(dev_alloc((void **)&c->d_res_fid, sizeof(double) * np));
    RETCHK(dev_alloc((void **)&c->d_res_abs, sizeof(double) * np));
    RETCHK(dev_alloc((void **)&c->d_zv, sizeof(double) * c->hw() * nv));
    RETCHK(dev_alloc((void **)&c->d_ze, sizeof(double) * c->hw() * std::max(ne, 1)));
    RETCHK(dev_alloc((void **)&c->d_lre, sizeof(double) * nnet));
    RETCHK(dev_alloc((void **)&c->d_lim, sizeof(double) * nnet));
    // per-network run state in one block (one copy back per run): delta, done, iter, at
    RETCHK(dev_alloc((void **)&c->d_delta, run_state_bytes(nnet)));
    c->d_done = (int *)(c->d_delta + nnet);
    c->d_iter = c->d_done + nnet;
    c->d_at = c->d_iter + nnet;
    RETCHK(dev_alloc((void **)&c->d_sites64, sizeof(double) * c->hw() * c->nsite));
    RETCHK(dev_alloc(&c->d_sites, es * c->nsite));
    if (c->dtype == BP_BF16) RETCHK(dev_alloc((void **)&c->d_sites_bf, sizeof(float) * c->nsite));
    if (c->d_msg64[0]) { (void)hipFree(c->d_msg64[0]); c->d_msg64[0] = nullptr; }
    if (c->d_msg64[1]) { (void)hipFree(c->d_msg64[1]); c->d_msg64[1] = nullptr; }
    c->fp64_active = false;
    c->cur = 0;
    RETCHK(rebuild_vertex_plan(c));
    c->have_topo = true;
    c->have_sites = false;
    c->have_msgs = false;
    return BP_OK;
}

int bp_upload_sites(bp_context *c, const double *sites) {
    if (!c || !sites) return set_err(BP_ERR_ARG, "NULL argument");
    if (!c->have_topo) return set_err(BP_ERR_STATE, "bp_set_topology first");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(c->d_sites64, sites, sizeof(double) * c->hw() * c->nsite,
                          hipMemcpyHostToDevice, c->stream));
    RETCHK(refresh_sites(c));
    c->have_sites = true;
    return BP_OK;
}

int bp_download_sites(bp_context *c, double *sites) {
    if (!c || !sites) return set_err(BP_ERR_ARG, "NULL argument");
    if (!c->have_sites) return set_err(BP_ERR_STATE, "no site tensors");
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipMemcpyAsync(sites, c->d_sites64, sizeof(double) * c->hw() * c->nsite,
                          hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return BP_OK;
}
}  // extern "C"

// the working copies of the FP64 master sites: the context's arithmetic type and the
// vertex kernels' padded layouts
static int refresh_sites(bp_context *c) {
    const int blocks = (int)std::min<long long>((c->nsite + 255) / 256, 65535);
    if (c->dtype == BP_F64 || c->is_cplx()) {
        HIPCHK(hipMemcpyAsync(c->d_sites, c->d_sites64, sizeof(double) * c->hw() * c->nsite,
                              hipMemcpyDeviceToDevice, c->stream));
    } else {
        hipLaunchKernelGGL((k_convert<double, float>), dim3(std::max(blocks, 1)), dim3(256), 0,
                           c->stream, c->d_sites64, (float *)c->d_sites, c->nsite);
        HIPCHK(hipGetLastError());
        if (c->dtype == BP_BF16) {
            hipLaunchKernelGGL(k_round_bf16, dim3(std::max(blocks, 1)), dim3(256), 0, c->stream,
                               c->d_sites64, c->d_sites_bf, c->nsite);
            HIPCHK(hipGetLastError());
        }
    }
    if (c->vplan_ok) RETCHK(upload_vertex_sites(c));
    c->v64_sites = false;
    c->v64_ok = false;
    if (c->v64_tried) {
        c->have_sites = true;  // the padded FP64 slices follow the new sites now
        RETCHK(ensure_v64(c));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return BP_OK;
}
extern "C" {

int bp_init_messages(bp_context *c, int kind, const double *host) {
    if (!c) return set_err(BP_ERR_ARG, "ctx is NULL");
    if (!c->have_topo) return set_err(BP_ERR_STATE, "bp_set_topology first");
    HIPCHK(hipSetDevice(c->device));
    c->cur = 0;
    c->fp64_active = false;
    if (kind == BP_INIT_UPLOAD) {
        if (!host) return set_err(BP_ERR_ARG, "host buffer required for BP_INIT_UPLOAD");
        if (c->dtype == BP_F64 || c->is_cplx()) {
            HIPCHK(hipMemcpyAsync(c->d_msg[0], host, sizeof(double) * c->hw() * c->nmsg,
                                  hipMemcpyHostToDevice, c->stream));
        } else {
            double *tmp = nullptr;
            HIPCHK(hipMalloc((void **)&tmp, sizeof(double) * c->nmsg));
            HIPCHK(hipMemcpyAsync(tmp, host, sizeof(double) * c->nmsg, hipMemcpyHostToDevice,
                                  c->stream));
            const int blocks = (int)std::min<long long>((c->nmsg + 255) / 256, 65535);
            hipLaunchKernelGGL((k_convert<double, float>), dim3(std::max(blocks, 1)), dim3(256),
                               0, c->stream, tmp, (float *)c->d_msg[0], c->nmsg);
            HIPCHK(hipStreamSynchronize(c->stream));
            HIPCHK(hipFree(tmp));
        }
    } else if (kind == BP_INIT_IDENTITY || kind == BP_INIT_ONES) {
        if (c->nports == 0) {
            // no edges: nothing to initialise
        } else if (c->is_cplx())
            hipLaunchKernelGGL(k_init_messages<cplx>, dim3(c->nports), dim3(64), 0, c->stream,
                               c->topo(), (cplx *)c->d_msg[0], kind);
        else if (c->dtype == BP_F64)
            hipLaunchKernelGGL(k_init_messages<double>, dim3(c->nports), dim3(64), 0, c->stream,
                               c->topo(), (double *)c->d_msg[0], kind);
        else
            hipLaunchKernelGGL(k_init_messages<float>, dim3(c->nports), dim3(64), 0, c->stream,
                               c->topo(), (float *)c->d_msg[0], kind);
        HIPCHK(hipGetLastError());
    } else {
        return set_err(BP_ERR_ARG, "unknown init kind %d", kind);
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    c->have_msgs = true;
    return BP_OK;
}

int bp_download_messages(bp_context *c, double *out) {
    if (!c || !out) return set_err(BP_ERR_ARG, "NULL argument");
    if (!c->have_msgs) return set_err(BP_ERR_STATE, "no messages");
    HIPCHK(hipSetDevice(c->device));
    if (c->fp64_active) {
        HIPCHK(hipMemcpyAsync(out, c->d_msg64[c->cur64], sizeof(double) * c->nmsg,
                              hipMemcpyDeviceToHost, c->stream));
    } else if (c->dtype == BP_F64 || c->is_cplx()) {
        HIPCHK(hipMemcpyAsync(out, c->d_msg[c->cur], sizeof(double) * c->hw() * c->nmsg,
                              hipMemcpyDeviceToHost, c->stream));
    } else {
        std::vector<float> tmp(c->nmsg);
        HIPCHK(hipMemcpyAsync(tmp.data(), c->d_msg[c->cur], sizeof(float) * c->nmsg,
                              hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (long long i = 0; i < c->nmsg; ++i) out[i] = tmp[i];
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return BP_OK;
}

// Slot ranges of the message buffer: checks the slots and sizes a host buffer.
static int slot_span(const bp_context *c, int n, const int *slots, long long *total) {
    if (n < 0 || (n > 0 && !slots)) return set_err(BP_ERR_ARG, "bad slot list");
    long long t = 0;
    for (int i = 0; i < n; ++i) {
        if (slots[i] < 0 || slots[i] >= c->nports)
            return set_err(BP_ERR_ARG, "slot %d out of range", slots[i]);
        t += c->msg_off[slots[i] + 1] - c->msg_off[slots[i]];
    }
    *total = t;
    return BP_OK;
}

// Copies between the slots' messages and a packed FP64 host buffer (dir 0: host -> device).
static int slot_copy(bp_context *c, int n, const int *slots, double *host, int dir) {
    long long total = 0;
    RETCHK(slot_span(c, n, slots, &total));
    HIPCHK(hipSetDevice(c->device));
    // FP64 and complex messages copy as they are (hw doubles per element); FP32 converts
    const bool f64 = c->fp64_active || c->dtype == BP_F64 || c->is_cplx();
    const int hw = c->fp64_active ? 1 : c->hw();
    char *base = c->fp64_active ? (char *)c->d_msg64[c->cur64] : (char *)c->d_msg[c->cur];
    const size_t esz = f64 ? sizeof(double) * hw : sizeof(float);
    std::vector<float> tmp(f64 ? 0 : (size_t)total);
    if (!f64 && dir == 0)
        for (long long i = 0; i < total; ++i) tmp[i] = (float)host[i];
    long long off = 0;
    for (int i = 0; i < n; ++i) {
        const long long mo = c->msg_off[slots[i]], len = c->msg_off[slots[i] + 1] - mo;
        char *dev = base + esz * mo;
        char *h = f64 ? (char *)(host + hw * off) : (char *)(tmp.data() + off);
        if (dir == 0)
            HIPCHK(hipMemcpyAsync(dev, h, esz * len, hipMemcpyHostToDevice, c->stream));
        else
            HIPCHK(hipMemcpyAsync(h, dev, esz * len, hipMemcpyDeviceToHost, c->stream));
        off += len;
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    if (!f64 && dir == 1)
        for (long long i = 0; i < total; ++i) host[i] = tmp[i];
    return BP_OK;
}

int bp_set_slot_messages(bp_context *c, int n, const int *slots, const double *data) {
    if (!c || (n > 0 && !data)) return set_err(BP_ERR_ARG, "NULL argument");
    if (!c->have_msgs) return set_err(BP_ERR_STATE, "no messages (bp_init_messages first)");
    return slot_copy(c, n, slots, const_cast<double *>(data), 0);
}

int bp_get_slot_messages(bp_context *c, int n, const int *slots, double *out) {
    if (!c || (n > 0 && !out)) return set_err(BP_ERR_ARG, "NULL argument");
    if (!c->have_msgs) return set_err(BP_ERR_STATE, "no messages");
    return slot_copy(c, n, slots, out, 1);
}

int bp_set_path(bp_context *c, int path) {
    if (!c) return set_err(BP_ERR_ARG, "ctx is NULL");
    if (path != BP_PATH_AUTO && path != BP_PATH_GENERIC && path != BP_PATH_VERTEX)
        return set_err(BP_ERR_ARG, "unknown path %d", path);
    if (path == BP_PATH_VERTEX && !c->normalize)
        return set_err(BP_ERR_UNSUPPORTED, "the vertex kernels always normalise (option normalize is 0)");
    if (path == BP_PATH_VERTEX && c->is_cplx())
        return set_err(BP_ERR_UNSUPPORTED, "complex contexts run on the generic kernel");
    if (path == BP_PATH_VERTEX && c->dtype == BP_F64) {
        RETCHK(ensure_v64(c));
        if (!c->v64_plan_ok)
            return set_err(BP_ERR_UNSUPPORTED, "FP64 vertex path not available: %s; %s",
                           c->v64.why.c_str(), c->v64l.why.c_str());
    } else if (path == BP_PATH_VERTEX && !(c->vplan_ok && c->vplan.all)) {
        return set_err(BP_ERR_UNSUPPORTED, "vertex path not available for every network: %s",
                       c->vplan.why.c_str());
    }
    bump_epoch();
    c->path = path;
    return BP_OK;
}

int bp_set_option(bp_context *c, const char *name, int value) {
    if (!c || !name) return set_err(BP_ERR_ARG, "NULL argument");
    bump_epoch();
    if (strcmp(name, "vertex_variant") == 0) {
        if (value < 0 || value > 5) return set_err(BP_ERR_ARG, "vertex_variant is 0..5");
        if (value >= 2 && value <= 4 && c->dtype != BP_F32)
            return set_err(BP_ERR_UNSUPPORTED, "vertex_variant %d (MFMA open contractions) is FP32 only",
                           value);
        if (value == 5 && c->dtype != BP_F32 && c->dtype != BP_BF16)
            return set_err(BP_ERR_UNSUPPORTED, "vertex_variant 5 is FP32 or BF16");
        c->vx_variant = value;
        c->vplan.variant = value;
        return BP_OK;
    }
    if (strcmp(name, "vertex16_waves") == 0) {
        if (value != 4 && value != 8 && value != 16)
            return set_err(BP_ERR_ARG, "vertex16_waves is 4, 8 or 16");
        c->v16_waves = value;
        c->vplan.v16_waves = value;
        return BP_OK;
    }
    if (strcmp(name, "fold_bookkeeping") == 0) {
        if (value != 0 && value != 1) return set_err(BP_ERR_ARG, "fold_bookkeeping is 0 or 1");
        c->fold_on = value;
        return BP_OK;
    }
    if (strcmp(name, "normalize") == 0) {
        if (value != 0 && value != 1) return set_err(BP_ERR_ARG, "normalize is 0 or 1");
        if (!value && c->path == BP_PATH_VERTEX)
            return set_err(BP_ERR_UNSUPPORTED, "the vertex kernels always normalise");
        c->normalize = value;
        return BP_OK;
    }
    if (strcmp(name, "vertex_bond_mask") == 0) {
        c->vx_mask = (unsigned)value & 0x1FFFEu;
        if (c->have_topo) RETCHK(rebuild_vertex_plan(c));
        return BP_OK;
    }
    if (strcmp(name, "small_bond") == 0) {
        if (value != 0 && value != 1) return set_err(BP_ERR_ARG, "small_bond is 0 or 1");
        c->vx_small = value;
        if (c->have_topo) RETCHK(rebuild_vertex_plan(c));
        return BP_OK;
    }
    return set_err(BP_ERR_ARG, "unknown option %s", name);
}

int bp_get_option(bp_context *c, const char *name, int *value) {
    if (!c || !name || !value) return set_err(BP_ERR_ARG, "NULL argument");
    if (strcmp(name, "vertex_variant") == 0) { *value = vx_variant_of(c); return BP_OK; }
    if (strcmp(name, "vertex16_waves") == 0) { *value = c->v16_waves; return BP_OK; }
    if (strcmp(name, "vertex_bond_mask") == 0) { *value = (int)c->vx_mask; return BP_OK; }
    if (strcmp(name, "fold_bookkeeping") == 0) { *value = c->fold_on; return BP_OK; }
    if (strcmp(name, "normalize") == 0) { *value = c->normalize; return BP_OK; }
    if (strcmp(name, "small_bond") == 0) { *value = c->vx_small; return BP_OK; }
    return set_err(BP_ERR_ARG, "unknown option %s", name);
}

int bp_get_path(bp_context *c, int *path) {
    if (!c || !path) return set_err(BP_ERR_ARG, "NULL argument");
    if (c->dtype == BP_F64 && c->have_sites) RETCHK(ensure_v64(c));
    *path = resolve_path(c);
    if (*path < 0)
        return set_err(BP_ERR_UNSUPPORTED, "no kernel path for this configuration: %s",
                       c->vplan.why.c_str());
    return BP_OK;
}

static int ready(bp_context *c) {
    if (!c) return set_err(BP_ERR_ARG, "ctx is NULL");
    if (!c->have_topo || !c->have_sites || !c->have_msgs)
        return set_err(BP_ERR_STATE, "topology, sites and messages must be set first");
    return BP_OK;
}

// Entry points whose kernels compute in real arithmetic only.
static int real_only(const bp_context *c, const char *what) {
    if (c->is_cplx())
        return set_err(BP_ERR_UNSUPPORTED, "%s: complex contexts support BP, scalars, log-norm "
                       "and iterate_diff only", what);
    return BP_OK;
}

int bp_sweep(bp_context *c) {
    RETCHK(ready(c));
    HIPCHK(hipSetDevice(c->device));
    c->fp64_active = false;
    RETCHK(sweep_once(c, make_ctl(c, false)));
    HIPCHK(hipStreamSynchronize(c->stream));
    return BP_OK;
}

// Sweeps per captured chunk of bp_run (between host checks of the done flags with a
// tolerance); a multiple of FOLD_BUFS, so every chunk starts on accumulator residue 0.
static constexpr int RUN_CHUNK = 16;
static_assert(RUN_CHUNK % FOLD_BUFS == 0, "a captured sweep chunk must span whole accumulator cycles");

// Capture and instantiate n sweeps (each with its bookkeeping) starting at the current
// parity into *out.  A chunk starts on accumulator residue 0 (runs queue whole chunks of
// 16, a multiple of FOLD_BUFS), so sweep i of the chunk is captured as sweep i.  Capture
// executes nothing: the parity is restored, and each replay flips it n times.
static int capture_sweeps(bp_context *c, const NetCtl &ctl, int n, hipGraphExec_t *out) {
    hipGraph_t graph = nullptr;
    const int cur0 = c->cur;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = BP_OK;
    bool folded = false;
    for (int i = 0; i < n && rc == BP_OK; ++i) {
        rc = sweep_once(c, ctl, &folded, i, i > 0);
        if (rc == BP_OK && !folded) rc = launch_reduce(c);
    }
    if (rc == BP_OK && folded) rc = launch_reduce_folded(c, n - 1);
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    c->cur = cur0;
    if (rc != BP_OK) {
        if (graph) (void)hipGraphDestroy(graph);
        return rc;
    }
    if (e != hipSuccess) return set_err(BP_ERR_HIP, "graph capture failed: %s", hipGetErrorString(e));
    const hipError_t ie = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (ie != hipSuccess) {
        *out = nullptr;
        return set_err(BP_ERR_HIP, "graph instantiate failed: %s", hipGetErrorString(ie));
    }
    // upload now, so that the first replay does not pay for it
    HIPCHK(hipGraphUpload(*out, c->stream));
    c->graph_epoch = g_epoch.load();
    return BP_OK;
}

int bp_prepare_run(bp_context *c, int maxiter) {
    RETCHK(ready(c));
    if (maxiter < 0) return set_err(BP_ERR_ARG, "maxiter must be >= 0");
    HIPCHK(hipSetDevice(c->device));
    RETCHK(prepare_sweep(c));
    if (c->graph_epoch != g_epoch.load()) graphs_release(c);
    const NetCtl ctl = make_ctl(c, true);
    const int tail = maxiter % RUN_CHUNK;
    const int cur0 = c->cur;
    for (int par = 0; par < 2; ++par) {  // a run may start on either message buffer
        c->cur = par;
        int rc = BP_OK;
        if (maxiter >= RUN_CHUNK && !c->gexec[par])
            rc = capture_sweeps(c, ctl, RUN_CHUNK, &c->gexec[par]);
        if (rc == BP_OK && tail > 0 && c->gtail_n[par] != tail) {
            if (c->gtail[par]) (void)hipGraphExecDestroy(c->gtail[par]);
            c->gtail[par] = nullptr;
            c->gtail_n[par] = 0;
            rc = capture_sweeps(c, ctl, tail, &c->gtail[par]);
            if (rc == BP_OK) c->gtail_n[par] = tail;
        }
        if (rc != BP_OK) {
            c->cur = cur0;
            return rc;
        }
    }
    c->cur = cur0;
    return BP_OK;
}

int bp_run(bp_context *c, int maxiter, double tol, int criterion, int use_graph, int *out_sweeps,
           double *out_delta, int *out_at) {
    RETCHK(ready(c));
    if (maxiter <= 0 && tol < 0.0)
        return set_err(BP_ERR_ARG, "at least one of maxiter or tol must be given");
    if (criterion != BP_CRIT_FIDELITY && criterion != BP_CRIT_MAXABS)
        return set_err(BP_ERR_ARG, "unknown criterion %d", criterion);
    HIPCHK(hipSetDevice(c->device));
    c->fp64_active = false;
    RETCHK(prepare_sweep(c));
    const int cap = maxiter > 0 ? maxiter : 100000;  // tol-only runs are capped
    // The run state (delta, done, iter, at: one block) comes back through a pinned buffer.
    const int nn = c->nnet;
    const size_t sbytes = run_state_bytes(nn);
    if (c->h_state_bytes < sbytes) {
        if (c->h_state) (void)hipHostFree(c->h_state);
        c->h_state = nullptr;
        c->h_state_bytes = 0;
        HIPCHK(hipHostMalloc((void **)&c->h_state, sbytes, hipHostMallocDefault));
        c->h_state_bytes = sbytes;
    }
    RunParams rp;
    rp.maxiter = cap;
    rp.crit = criterion;
    rp.tol = tol;
    hipLaunchKernelGGL(k_reset_nets, dim3((c->nnet + 255) / 256), dim3(256), 0, c->stream,
                       c->nnet, c->d_done, c->d_iter, c->d_delta, c->d_at, c->d_fold_acc,
                       c->fold_n, c->d_runp, rp);
    const NetCtl ctl = make_ctl(c, true);
    HIPCHK(hipGetLastError());
    // a captured chunk bakes in the folded accumulator indices k % FOLD_BUFS, so every
    // replay must start on residue 0
    constexpr int chunk = RUN_CHUNK;
    // With a tolerance the host reads the run state after every chunk and stops once every
    // network has stopped; without one no network stops before maxiter, so every chunk is
    // queued at once and the state is read once at the end (no idle gaps between chunks).
    const bool check_chunks = tol >= 0.0;
    if (c->graph_epoch != g_epoch.load()) graphs_release(c);
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    int launched = 0;
    while (launched < cap) {
        const int n = std::min(chunk, cap - launched);
        const bool tail = use_graph && n < chunk && c->gtail[c->cur] && c->gtail_n[c->cur] == n;
        if (use_graph && (n == chunk || tail)) {
            hipGraphExec_t &gexec = tail ? c->gtail[c->cur] : c->gexec[c->cur];
            if (!gexec) RETCHK(capture_sweeps(c, ctl, n, &gexec));
            HIPCHK(hipGraphLaunch(gexec, c->stream));
            if (n % 2) c->cur ^= 1;
        } else {
            bool folded = false;
            for (int i = 0; i < n; ++i) {
                RETCHK(sweep_once(c, ctl, &folded, launched + i, i > 0));
                if (!folded) RETCHK(launch_reduce(c));
            }
            if (folded) RETCHK(launch_reduce_folded(c, launched + n - 1));
        }
        launched += n;
        if (!check_chunks && launched < cap) continue;
        HIPCHK(hipEventRecord(c->ev1, c->stream));
        HIPCHK(hipMemcpyAsync(c->h_state, c->d_delta, sbytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        const int *done = (const int *)(c->h_state + sizeof(double) * nn);
        bool all = true;
        for (int i = 0; i < nn; ++i) all = all && done[i];
        if (all) break;
    }
    if (launched == 0) {
        HIPCHK(hipEventRecord(c->ev1, c->stream));
        HIPCHK(hipMemcpyAsync(c->h_state, c->d_delta, sbytes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    std::vector<char> st(c->h_state, c->h_state + sbytes);
    const char *p = st.data();
    if (out_delta) std::memcpy(out_delta, p, sizeof(double) * nn);
    p += sizeof(double) * nn + sizeof(int) * nn;  // past delta and done
    if (out_sweeps) std::memcpy(out_sweeps, p, sizeof(int) * nn);
    p += sizeof(int) * nn;
    if (out_at) std::memcpy(out_at, p, sizeof(int) * nn);
    return BP_OK;
}

// ------------------------------------------------------------------ autotuner
//
// Choices, each timed with HIP events on the context's own data (min of 3 launches after
// a warm-up) and kept in a JSON cache, one entry per line:
//   "vx|<dtype>|d<d>|D<D>": {"vertex": 0|1, "us": t_chosen, "us_alt": t_other}
//       networks whose largest bond dimension is D run on the vertex kernels (padded to 8,
//       12 or 16; FP64: the FP64 vertex kernel, D <= 8) or on the generic kernel,
//       whichever sweeps faster;
//   "vxk|<f32|bf16>|d<d>": {"variant": 0..5, "us": t_chosen, "us_alt": t_next}
//       the D <= 8 vertex sweep kernel: 0 two LDS buffers per wave, 1 one buffer per wave
//       with two waves per SIMD, 2 and 3 (FP32 only) open contractions on MFMA with the
//       bra from LDS or from the padded site copies in HBM, 4 (FP32 only) and 5 half or
//       all of them on MFMA in place (the partner message absorbed into the bra);
//   "v16w|bf16|d<d>": {"waves": 4|8|16, "us": t_best, "us_alt": t_worst}
//       waves per slice workgroup of the bond-dimension-16 kernel;
//   "gen|<dtype>|d<d>|D<D>": {"block": 128|256|512, "lds": 0|1|2, "us": t, "us_alt": t0}
//       workgroup size and intermediate storage (global, LDS up to 64 KiB, LDS up to
//       160 KiB) of the generic kernel for the bucket of slots whose owner has physical
//       dimension d and largest bond dimension D (t0: the default's time).

static void tune_load(bp_context *c, const char *path) {
    std::ifstream f(path);
    if (!f) return;
    std::string line;
    while (std::getline(f, line)) {
        const size_t a = line.find('"');
        if (a == std::string::npos) continue;
        const size_t b = line.find('"', a + 1);
        if (b == std::string::npos) continue;
        const std::string key = line.substr(a + 1, b - a - 1);
        if (key.rfind("gen|", 0) != 0 && key.rfind("vx|", 0) != 0 &&
            key.rfind("vxk|", 0) != 0 && key.rfind("v16w|", 0) != 0)
            continue;
        TuneEntry e;
        auto num = [&](const char *name, double *out) {
            const std::string pat = std::string("\"") + name + "\":";
            const size_t p = line.find(pat, b + 1);
            if (p == std::string::npos) return false;
            *out = atof(line.c_str() + p + pat.size());
            return true;
        };
        double v = 0.0;
        if (num("block", &v)) e.block = (int)v;
        if (num("waves", &v)) e.block = (int)v;
        if (num("lds", &v)) e.lds = (int)v;
        if (num("vertex", &v)) e.vertex = (int)v;
        if (num("variant", &v)) e.vertex = (int)v;
        if (num("us", &v)) e.us = v;
        if (num("us_alt", &v)) e.us_alt = v;
        if (key.rfind("v16w|", 0) == 0) {
            if (e.block != 4 && e.block != 8 && e.block != 16) e.block = 8;
        } else if (e.block != 128 && e.block != 256 && e.block != 512) {
            e.block = BP_BLOCK;
        }
        c->tune[key] = e;
    }
}

static int tune_save(const bp_context *c, const char *path) {
    std::ofstream f(path);
    if (!f) return set_err(BP_ERR_ARG, "cannot write the autotuner cache %s", path);
    f << "{\n";
    size_t i = 0;
    for (const auto &kv : c->tune) {
        const TuneEntry &e = kv.second;
        f << "  \"" << kv.first << "\": {";
        if (kv.first.rfind("gen|", 0) == 0) f << "\"block\": " << e.block << ", \"lds\": " << e.lds;
        else if (kv.first.rfind("vxk|", 0) == 0) f << "\"variant\": " << e.vertex;
        else if (kv.first.rfind("v16w|", 0) == 0) f << "\"waves\": " << e.block;
        else f << "\"vertex\": " << e.vertex;
        f << ", \"us\": " << e.us << ", \"us_alt\": " << e.us_alt << "}"
          << (++i < c->tune.size() ? "," : "") << "\n";
    }
    f << "}\n";
    f.close();
    return f ? BP_OK : set_err(BP_ERR_ARG, "cannot write the autotuner cache %s", path);
}

// One full plain sweep (every network), in ms: the minimum over `reps` of the mean over `per`
// back-to-back sweeps (one event pair around several launches keeps the event overhead and
// a single launch's jitter out of the comparison); -1 on error
static double tune_time_sweep(bp_context *c, int reps, int per = 4) {
    if (prepare_sweep(c) != BP_OK) return -1.0;
    const NetCtl ctl = make_ctl(c, false);
    if (sweep_once(c, ctl) != BP_OK) return -1.0;
    double best = 1e30;
    for (int r = 0; r < reps; ++r) {
        if (hipEventRecord(c->ev0, c->stream) != hipSuccess) return -1.0;
        for (int k = 0; k < per; ++k)
            if (sweep_once(c, ctl) != BP_OK) return -1.0;
        if (hipEventRecord(c->ev1, c->stream) != hipSuccess ||
            hipEventSynchronize(c->ev1) != hipSuccess)
            return -1.0;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
        best = std::min(best, (double)ms / per);
    }
    return best;
}

// min over `reps` of one launch of bucket b (current messages in, the other buffer out)
static double tune_time_bucket(bp_context *c, const GenBucket &b, int reps) {
    const NetCtl ctl = make_ctl(c, false);
    void *src = c->d_msg[c->cur], *dst = c->d_msg[c->cur ^ 1];
    auto launch = [&]() {
        if (c->is_cplx())
            return launch_bucket<cplx>(c, b, (const cplx *)c->d_sites, (const cplx *)src,
                                       (cplx *)dst, ctl, 0);
        if (c->dtype == BP_F64)
            return launch_bucket<double>(c, b, (const double *)c->d_sites, (const double *)src,
                                         (double *)dst, ctl, 0);
        return launch_bucket<float>(c, b, generic_sites32(c), (const float *)src, (float *)dst,
                                    ctl, c->dtype == BP_BF16);
    };
    if (launch() != BP_OK) return -1.0;
    double best = 1e30;
    for (int r = 0; r < reps; ++r) {
        if (hipEventRecord(c->ev0, c->stream) != hipSuccess) return -1.0;
        if (launch() != BP_OK) return -1.0;
        if (hipEventRecord(c->ev1, c->stream) != hipSuccess ||
            hipEventSynchronize(c->ev1) != hipSuccess)
            return -1.0;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
        best = std::min(best, (double)ms);
    }
    return best;
}

int bp_autotune(bp_context *c, const char *cache_path, int mode, int *out_tuned) {
    RETCHK(ready(c));
    if (mode < 0 || mode > 2) return set_err(BP_ERR_ARG, "unknown autotune mode %d", mode);
    HIPCHK(hipSetDevice(c->device));
    if (cache_path && mode < 2) tune_load(c, cache_path);
    if (mode == 2) c->tune.clear();
    int tuned = 0;
    const size_t es = c->elt();
    // the timed sweeps advance the messages: keep a copy and restore it afterwards
    void *snap = nullptr;
    HIPCHK(hipMalloc(&snap, es * std::max<long long>(c->nmsg, 1)));
    const int cur0 = c->cur;
    const bool fp64_0 = c->fp64_active;
    HIPCHK(hipMemcpyAsync(snap, c->d_msg[cur0], es * c->nmsg, hipMemcpyDeviceToDevice, c->stream));
    c->fp64_active = false;
    int rc = BP_OK;
    // 1. vertex kernels or generic kernel, per largest bond dimension (<= 12 FP32, <= 16 BF16)
    if (!c->plain && !c->is_cplx() && c->path == BP_PATH_AUTO) {
        std::map<int, int> dmax_d;  // largest bond dimension -> physical dimension
        for (int n = 0; n < c->nnet && rc == BP_OK; ++n) {
            int dm = 0, deg = 0, d = c->phys_dim[c->vbeg[n]];
            for (int u = c->vbeg[n]; u < c->vbeg[n + 1]; ++u) {
                deg = std::max(deg, c->port_off[u + 1] - c->port_off[u]);
                for (int q = c->port_off[u]; q < c->port_off[u + 1]; ++q)
                    dm = std::max(dm, c->edge_dim[c->port_edge[q]]);
            }
            const int dlim = c->dtype == BP_BF16 ? 16 : c->dtype == BP_F32 ? 12 : 16;
            if (dm >= 1 && dm <= dlim && deg <= 4 && d <= 4) dmax_d[dm] = d;
        }
        for (const auto &kv : dmax_d) {
            const int D = kv.first;
            std::ostringstream o;
            o << "vx|" << dtype_name(c->dtype) << "|d" << kv.second << "|D" << D;
            const std::string key = o.str();
            auto it = c->tune.find(key);
            if (it == c->tune.end() && mode >= 1) {
                c->vx_mask |= 1u << D;
                rc = rebuild_vertex_plan(c);
                const double t_on = rc == BP_OK ? tune_time_sweep(c, 3) : -1.0;
                c->vx_mask &= ~(1u << D);
                if (rc == BP_OK) rc = rebuild_vertex_plan(c);
                const double t_off = rc == BP_OK ? tune_time_sweep(c, 3) : -1.0;
                if (rc != BP_OK || t_on < 0.0 || t_off < 0.0) {
                    if (rc == BP_OK) rc = set_err(BP_ERR_HIP, "autotuner sweep failed");
                    break;
                }
                TuneEntry e;
                e.vertex = t_on <= t_off;
                e.us = 1e3 * std::min(t_on, t_off);
                e.us_alt = 1e3 * std::max(t_on, t_off);
                it = c->tune.emplace(key, e).first;
                ++tuned;
            }
            if (it != c->tune.end()) {
                if (it->second.vertex) c->vx_mask |= 1u << D;
                else c->vx_mask &= ~(1u << D);
            }
        }
        if (rc == BP_OK) rc = rebuild_vertex_plan(c);
    }
    // 1b. D <= 8 vertex sweep kernel variant (FP32 or BF16), timed on full sweeps
    if (rc == BP_OK && (c->dtype == BP_F32 || c->dtype == BP_BF16) && c->vplan_ok &&
        c->vplan.nv > c->vplan.nv_k3 && c->path != BP_PATH_GENERIC) {
        std::ostringstream o;
        o << "vxk|" << (c->dtype == BP_F32 ? "f32" : "bf16") << "|d" << c->vplan.d;
        const std::string key = o.str();
        auto it = c->tune.find(key);
        if (it == c->tune.end() && mode >= 1) {
            // FP32: six kernels (two with VALU open contractions, two with MFMA ones staged
            // through LDS, two with half or all of them on MFMA in place); BF16: variants 0,
            // 1 and 5
            const int nvar = 6;
            const bool cand[6] = {true, true, c->dtype == BP_F32, c->dtype == BP_F32,
                                  c->dtype == BP_F32, true};
            double t[6] = {-1.0, -1.0, -1.0, -1.0, -1.0, -1.0};
            int best = 0;
            // the candidates in turn, three rounds, each keeping its best (drift in clocks or
            // in the rest of the device then weighs on every candidate alike)
            for (int round = 0; round < 3 && rc == BP_OK; ++round)
                for (int v = 0; v < nvar && rc == BP_OK; ++v) {
                    if (!cand[v]) continue;
                    c->vplan.variant = v;
                    const double tv = tune_time_sweep(c, 1);
                    if (tv < 0.0) rc = set_err(BP_ERR_HIP, "autotuner sweep failed");
                    else if (t[v] < 0.0 || tv < t[v]) t[v] = tv;
                }
            for (int v = 1; v < nvar && rc == BP_OK; ++v)
                if (cand[v] && t[v] < t[best]) best = v;
            if (rc == BP_OK) {
                double alt = 1e30;
                for (int v = 0; v < nvar; ++v)
                    if (cand[v] && v != best) alt = std::min(alt, t[v]);
                TuneEntry e;
                e.vertex = best;  // the chosen variant
                e.us = 1e3 * t[best];
                e.us_alt = 1e3 * alt;  // the runner-up
                it = c->tune.emplace(key, e).first;
                ++tuned;
            }
        }
        if (it != c->tune.end()) c->vx_variant = it->second.vertex;
        c->vplan.variant = vx_variant_of(c);
    }
    // 1c. waves per slice workgroup of the bond-dimension-16 kernel (BF16)
    if (rc == BP_OK && c->vplan_ok && c->vplan.nv16 > 0 && c->vplan.bf16 &&
        c->path != BP_PATH_GENERIC) {
        std::ostringstream o;
        o << "v16w|bf16|d" << c->vplan.d;
        const std::string key = o.str();
        auto it = c->tune.find(key);
        if (it == c->tune.end() && mode >= 1) {
            const int cand[3] = {4, 8, 16};
            double best = 1e30, worst = 0.0;
            int bw = c->v16_waves;
            for (int k = 0; k < 3 && rc == BP_OK; ++k) {
                c->vplan.v16_waves = cand[k];
                const double t = tune_time_sweep(c, 3);
                if (t < 0.0) { rc = set_err(BP_ERR_HIP, "autotuner sweep failed"); break; }
                if (t < best) { best = t; bw = cand[k]; }
                worst = std::max(worst, t);
            }
            if (rc == BP_OK) {
                TuneEntry e;
                e.block = bw;
                e.us = 1e3 * best;
                e.us_alt = 1e3 * worst;
                it = c->tune.emplace(key, e).first;
                ++tuned;
            }
        }
        if (it != c->tune.end()) c->v16_waves = it->second.block;
        c->vplan.v16_waves = c->v16_waves;
    }
    // 2. generic buckets: workgroup size and storage
    const int path = resolve_path(c);
    if (rc == BP_OK && path >= 0 && path != BP_PATH_VERTEX && mode >= 1) {
        rc = prepare_sweep(c);
        for (auto &b : c->gb) {
            if (rc != BP_OK) break;
            if (c->tune.count(b.key)) continue;  // applied by build_buckets
            const int bs0 = b.bs, lds0 = b.lds;
            double t0 = -1.0, best = 1e30;
            int best_bs = bs0, best_lds = lds0;
            for (int bs : {128, 256, 512}) {
                for (int lds : {2, 1, 0}) {
                    if (lds && !bucket_lds_fits(b, es, lds)) continue;
                    if (lds == 2 && bucket_lds_fits(b, es, 1)) continue;  // same launch as 1
                    b.bs = bs;
                    b.lds = lds;
                    rc = ensure_bucket_ws(c, es);
                    if (rc != BP_OK) break;
                    const double t = tune_time_bucket(c, b, 3);
                    if (t < 0.0) { rc = set_err(BP_ERR_HIP, "autotuner launch failed"); break; }
                    if (bs == BP_BLOCK && lds == (bucket_lds_fits(b, es, 1) ? 1 : 0)) t0 = t;
                    if (t < best) { best = t; best_bs = bs; best_lds = lds; }
                }
                if (rc != BP_OK) break;
            }
            b.bs = best_bs;
            b.lds = best_lds;
            TuneEntry e;
            e.block = best_bs;
            e.lds = best_lds;
            e.us = 1e3 * best;
            e.us_alt = 1e3 * (t0 >= 0.0 ? t0 : best);
            c->tune[b.key] = e;
            ++tuned;
        }
        if (rc == BP_OK) rc = ensure_bucket_ws(c, es);
    }
    // restore the messages
    (void)hipMemcpyAsync(c->d_msg[cur0], snap, es * c->nmsg, hipMemcpyDeviceToDevice, c->stream);
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(snap);
    c->cur = cur0;
    c->fp64_active = fp64_0;
    bump_epoch();  // kernel choices may have changed
    if (rc != BP_OK) return rc;
    if (cache_path && mode >= 1) RETCHK(tune_save(c, cache_path));
    if (out_tuned) *out_tuned = tuned;
    return BP_OK;
}

int bp_time_sweeps(bp_context *c, int n, double *avg_ms) {
    RETCHK(ready(c));
    if (n <= 0 || !avg_ms) return set_err(BP_ERR_ARG, "n must be positive");
    HIPCHK(hipSetDevice(c->device));
    c->fp64_active = false;
    RETCHK(prepare_sweep(c));
    std::vector<hipEvent_t> ev(2 * n);
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    // one untimed sweep first: a kernel's first launch may also load its code object
    int rc = sweep_once(c, make_ctl(c, false));
    for (int i = 0; i < n && rc == BP_OK; ++i) {
        (void)hipEventRecord(ev[2 * i], c->stream);
        rc = sweep_once(c, make_ctl(c, false));
        (void)hipEventRecord(ev[2 * i + 1], c->stream);
    }
    (void)hipStreamSynchronize(c->stream);
    double tot = 0.0;
    for (int i = 0; i < n && rc == BP_OK; ++i) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]) != hipSuccess) rc = BP_ERR_HIP;
        tot += ms;
    }
    for (auto &e : ev) (void)hipEventDestroy(e);
    if (rc != BP_OK) return rc == BP_ERR_HIP ? set_err(rc, "event timing failed") : rc;
    *avg_ms = tot / n;
    return BP_OK;
}

int bp_last_run_ms(bp_context *c, double *ms) {
    if (!c || !ms) return set_err(BP_ERR_ARG, "NULL argument");
    *ms = c->last_ms;
    return BP_OK;
}

int bp_update_slot(bp_context *c, int slot) {
    RETCHK(ready(c));
    if (slot < 0 || slot >= c->nports) return set_err(BP_ERR_ARG, "slot %d out of range", slot);
    HIPCHK(hipSetDevice(c->device));
    c->fp64_active = false;
    void *m = c->d_msg[c->cur];
    if (c->is_cplx())
        RETCHK(launch_update_generic<cplx>(c, (const cplx *)c->d_sites, (const cplx *)m,
                                           (cplx *)m, make_ctl(c, false), slot, slot + 1,
                                           nullptr, gen_flags(c, false)));
    else if (c->dtype == BP_F64)
        RETCHK(launch_update_generic<double>(c, (const double *)c->d_sites, (const double *)m,
                                             (double *)m, make_ctl(c, false), slot, slot + 1,
                                             nullptr, gen_flags(c, false)));
    else
        RETCHK(launch_update_generic<float>(c, generic_sites32(c), (const float *)m,
                                            (float *)m, make_ctl(c, false), slot, slot + 1,
                                            nullptr, gen_flags(c, c->dtype == BP_BF16)));
    HIPCHK(hipStreamSynchronize(c->stream));
    return BP_OK;
}

int bp_residuals(bp_context *c, double *fid, double *maxabs) {
    if (!c) return set_err(BP_ERR_ARG, "ctx is NULL");
    if (!c->have_topo) return set_err(BP_ERR_STATE, "no topology");
    HIPCHK(hipSetDevice(c->device));
    if (fid) HIPCHK(hipMemcpy(fid, c->d_res_fid, sizeof(double) * c->nports, hipMemcpyDeviceToHost));
    if (maxabs)
        HIPCHK(hipMemcpy(maxabs, c->d_res_abs, sizeof(double) * c->nports, hipMemcpyDeviceToHost));
    return BP_OK;
}

// current messages as a device pointer + whether it is FP64
static const void *current_msgs(bp_context *c, bool *is64) {
    if (c->fp64_active) { *is64 = true; return c->d_msg64[c->cur64]; }
    *is64 = c->dtype == BP_F64;
    return c->d_msg[c->cur];
}

// The adjoint runs on the vertex kernels when the FP32 bucket holds every network (the
// generic kernels otherwise, and always for FP64 messages or a forced generic path).
static bool vertex_adjoint_ok(const bp_context *c, bool is64) {
    return !is64 && c->dtype == BP_F32 && c->vplan_ok && c->vplan.all && !c->vplan.bf16 &&
           c->vplan.nv16 == 0 && c->vplan.ns_total() == 0 && c->path != BP_PATH_GENERIC &&
           c->d_vsites != nullptr;
}

// per-slot distances of the current messages against `other` (device buffer) into
// fid / mab (the context's residual buffers when null)
static int diff_launch(bp_context *c, const void *other, bool other64, double *fid = nullptr,
                       double *mab = nullptr) {
    bool is64 = false;
    const void *m = current_msgs(c, &is64);
    if (!fid) fid = c->d_res_fid;
    if (!mab) mab = c->d_res_abs;
    if (c->nports == 0) return BP_OK;
    const dim3 g(c->nports), b(64);
    if (c->is_cplx() && !c->fp64_active)  // complex messages, `other` complex too
        hipLaunchKernelGGL(k_diff_slots_cplx, g, b, 0, c->stream, c->topo(), (const cplx *)m,
                           (const cplx *)other, fid, mab);
    else if (is64 && other64)
        hipLaunchKernelGGL((k_diff_slots<double, double>), g, b, 0, c->stream, c->topo(),
                           (const double *)m, (const double *)other, fid, mab);
    else if (is64)
        hipLaunchKernelGGL((k_diff_slots<double, float>), g, b, 0, c->stream, c->topo(),
                           (const double *)m, (const float *)other, fid, mab);
    else if (other64)
        hipLaunchKernelGGL((k_diff_slots<float, double>), g, b, 0, c->stream, c->topo(),
                           (const float *)m, (const double *)other, fid, mab);
    else
        hipLaunchKernelGGL((k_diff_slots<float, float>), g, b, 0, c->stream, c->topo(),
                           (const float *)m, (const float *)other, fid, mab);
    HIPCHK(hipGetLastError());
    return BP_OK;
}

int bp_iterate_diff(bp_context *c, const double *other, double *out) {
    RETCHK(ready(c));
    if (!other || !out) return set_err(BP_ERR_ARG, "NULL argument");
    HIPCHK(hipSetDevice(c->device));
    double *d_other = nullptr;
    const long long nd = c->hw() * c->nmsg;  // doubles in `other`
    HIPCHK(hipMalloc((void **)&d_other, sizeof(double) * std::max<long long>(nd, 1)));
    int rc = BP_OK;
    if (hipMemcpyAsync(d_other, other, sizeof(double) * nd, hipMemcpyHostToDevice,
                       c->stream) != hipSuccess)
        rc = set_err(BP_ERR_HIP, "copy failed");
    if (rc == BP_OK) rc = diff_launch(c, d_other, true);
    std::vector<double> fid(c->nports);
    if (rc == BP_OK && c->nports > 0 &&
        hipMemcpyAsync(fid.data(), c->d_res_fid, sizeof(double) * c->nports,
                       hipMemcpyDeviceToHost, c->stream) != hipSuccess)
        rc = set_err(BP_ERR_HIP, "copy failed");
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(d_other);
    if (rc != BP_OK) return rc;
    double mx = -INFINITY;  // maximum over edges (an empty cache gives -Inf, as maximum() would fail)
    for (double v : fid) mx = (v > mx || v != v) ? v : mx;
    *out = mx;
    return BP_OK;
}

int bp_run_sequential(bp_context *c, const int *order, int norder, int maxiter, double tol,
                      int criterion, int *out_sweeps, double *out_delta, int *out_at) {
    RETCHK(ready(c));
    if (!order || norder <= 0) return set_err(BP_ERR_ARG, "empty schedule");
    if (maxiter <= 0 && tol < 0.0)
        return set_err(BP_ERR_ARG, "at least one of maxiter or tol must be given");
    if (c->nnet != 1) return set_err(BP_ERR_UNSUPPORTED, "sequential runs take one network");
    for (int i = 0; i < norder; ++i)
        if (order[i] < 0 || order[i] >= c->nports)
            return set_err(BP_ERR_ARG, "schedule entry %d out of range", i);
    HIPCHK(hipSetDevice(c->device));
    c->fp64_active = false;
    const int cap = maxiter > 0 ? maxiter : 100000;
    const size_t es = c->elt();
    void *prev = nullptr;
    HIPCHK(hipMalloc(&prev, es * std::max<long long>(c->nmsg, 1)));
    std::vector<double> res(c->nports);
    const NetCtl off = make_ctl(c, false);
    int it = 0, at = -1, rc = BP_OK;
    double delta = INFINITY;
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    while (rc == BP_OK && it < cap) {
        if (hipMemcpyAsync(prev, c->d_msg[c->cur], es * c->nmsg, hipMemcpyDeviceToDevice,
                           c->stream) != hipSuccess) {
            rc = set_err(BP_ERR_HIP, "snapshot copy failed");
            break;
        }
        void *m = c->d_msg[c->cur];
        for (int i = 0; i < norder && rc == BP_OK; ++i) {
            if (c->is_cplx())
                rc = launch_update_generic<cplx>(c, (const cplx *)c->d_sites, (const cplx *)m,
                                                 (cplx *)m, off, order[i], order[i] + 1, nullptr,
                                                 gen_flags(c, false));
            else if (c->dtype == BP_F64)
                rc = launch_update_generic<double>(c, (const double *)c->d_sites,
                                                   (const double *)m, (double *)m, off, order[i],
                                                   order[i] + 1, nullptr, gen_flags(c, false));
            else
                rc = launch_update_generic<float>(c, generic_sites32(c), (const float *)m,
                                                  (float *)m, off, order[i], order[i] + 1,
                                                  nullptr, gen_flags(c, c->dtype == BP_BF16));
        }
        if (rc != BP_OK) break;
        ++it;
        if (tol >= 0.0) {
            rc = diff_launch(c, prev, c->dtype == BP_F64);
            if (rc != BP_OK) break;
            const double *src = criterion == BP_CRIT_FIDELITY ? c->d_res_fid : c->d_res_abs;
            if (hipMemcpyAsync(res.data(), src, sizeof(double) * c->nports, hipMemcpyDeviceToHost,
                               c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess) {
                rc = set_err(BP_ERR_HIP, "residual copy failed");
                break;
            }
            double mx = 0.0;
            for (double v : res) mx = (v > mx || v != v) ? v : mx;
            delta = mx;
            if (delta < tol) {
                at = it;
                break;
            }
        }
    }
    (void)hipEventRecord(c->ev1, c->stream);
    (void)hipEventSynchronize(c->ev1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, c->ev0, c->ev1);
    c->last_ms = ms;
    (void)hipFree(prev);
    if (rc != BP_OK) return rc;
    if (out_sweeps) *out_sweeps = it;
    if (out_delta) *out_delta = delta;
    if (out_at) *out_at = at;
    return BP_OK;
}

// Vertex scalars on the FP64 large-bond kernel (passes 1 and 2) when every network's largest
// bond dimension is 9..16 and the plan builds: FP64 messages directly, FP32 messages through
// an FP64 copy in a polish buffer (free while the FP32 messages are current).  *done tells
// whether it ran.
static int v64l_scalars(bp_context *c, const void *m, bool is64, bool *done) {
    *done = false;
    if (c->plain || c->path == BP_PATH_GENERIC || c->maxD < 9 || c->maxD > 16) return BP_OK;
    RETCHK(ensure_v64(c));
    if (!c->v64_ok || c->v64_kind != 2) return BP_OK;
    const double *m64 = (const double *)m;
    if (!is64) {
        if (!c->d_msg64[0]) RETCHK(dev_alloc((void **)&c->d_msg64[0], sizeof(double) * c->nmsg));
        const int blocks = (int)std::max<long long>(1, std::min<long long>((c->nmsg + 255) / 256, 65535));
        hipLaunchKernelGGL((k_convert<float, double>), dim3(blocks), dim3(256), 0, c->stream,
                           (const float *)m, c->d_msg64[0], c->nmsg);
        m64 = c->d_msg64[0];
    }
    if (vertex64l_scalars(&c->v64l, c->stream, c->d_msg_off, m64, c->d_zv) != 0)
        return set_err(BP_ERR_HIP, "FP64 vertex scalar launch failed");
    *done = true;
    return BP_OK;
}

static int scalars(bp_context *c) {
    bool is64 = false;
    const void *m = current_msgs(c, &is64);
    if (c->is_cplx()) {
        RETCHK(launch_vertex_scalars<cplx>(c, (const cplx *)c->d_sites64, (const cplx *)m));
        RETCHK(launch_edge_scalars<cplx>(c, (const cplx *)m));
        return BP_OK;
    }
    // largest bond dimensions 9..16: FP64 scalars on the large-bond kernel, in every dtype
    bool done = false;
    RETCHK(v64l_scalars(c, m, is64, &done));
    if (is64) {
        if (!done) RETCHK(launch_vertex_scalars<double>(c, c->d_sites64, (const double *)m));
        RETCHK(launch_edge_scalars<double>(c, (const double *)m));
        return BP_OK;
    }
    if (!done && vertex_adjoint_ok(c, false) && vd_set_limits(c->vplan.d) == 0) {
        // FP32 networks on the D <= 8 vertex kernels: the message-form dual kernel without
        // overrides gives Z_u = <G_{u, port 0}, m_in,0> from the unnormalised port-0 result
        const VdCtx vc{&c->vplan, c->stream, c->topo(), (const float *)c->d_vsites};
        if (vd_msg(vc, (const float *)m, nullptr, OP_I, 1.0, nullptr, 0, nullptr, nullptr,
                   nullptr, nullptr, c->d_zv) != 0)
            return set_err(BP_ERR_HIP, "vertex scalar launch failed");
        done = true;
    }
    if (!done) RETCHK(launch_vertex_scalars<float>(c, (const float *)c->d_sites, (const float *)m));
    RETCHK(launch_edge_scalars<float>(c, (const float *)m));
    return BP_OK;
}

int bp_vertex_scalars(bp_context *c, double *zv) {
    RETCHK(ready(c));
    if (!zv) return set_err(BP_ERR_ARG, "NULL argument");
    HIPCHK(hipSetDevice(c->device));
    RETCHK(scalars(c));
    HIPCHK(hipMemcpyAsync(zv, c->d_zv, sizeof(double) * c->hw() * c->nv, hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return BP_OK;
}

int bp_edge_scalars(bp_context *c, double *ze) {
    RETCHK(ready(c));
    if (!ze) return set_err(BP_ERR_ARG, "NULL argument");
    HIPCHK(hipSetDevice(c->device));
    RETCHK(scalars(c));
    if (c->ne > 0)
        HIPCHK(hipMemcpyAsync(ze, c->d_ze, sizeof(double) * c->hw() * c->ne,
                              hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return BP_OK;
}

int bp_lognorm(bp_context *c, double *out_re, double *out_im) {
    RETCHK(ready(c));
    if (!out_re) return set_err(BP_ERR_ARG, "NULL argument");
    HIPCHK(hipSetDevice(c->device));
    RETCHK(scalars(c));
    if (c->is_cplx())
        hipLaunchKernelGGL(k_lognorm_nets_cplx, dim3(c->nnet), dim3(BP_BLOCK), 0, c->stream,
                           c->topo(), c->d_vbeg, c->d_pbeg, c->d_zv, c->d_ze, c->d_lre, c->d_lim);
    else
        hipLaunchKernelGGL(k_lognorm_nets, dim3(c->nnet), dim3(BP_BLOCK), 0, c->stream,
                           c->topo(), c->d_vbeg, c->d_pbeg, c->d_zv, c->d_ze, c->d_lre,
                           c->d_lim);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out_re, c->d_lre, sizeof(double) * c->nnet, hipMemcpyDeviceToHost,
                          c->stream));
    if (out_im)
        HIPCHK(hipMemcpyAsync(out_im, c->d_lim, sizeof(double) * c->nnet, hipMemcpyDeviceToHost,
                              c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return BP_OK;
}

// out[0] = max over v[0, n) with a NaN kept (-Inf when n = 0): one workgroup, no atomics.
__global__ __launch_bounds__(256) void k_max_slots(const double *v, int n, double *out) {
    __shared__ double red[256];
    double mx = -INFINITY;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double x = v[i];
        mx = (x > mx || x != x) ? x : mx;
    }
    red[threadIdx.x] = mx;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) {
            const double o = red[threadIdx.x + h];
            if (o > red[threadIdx.x] || o != o) red[threadIdx.x] = o;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

int bp_iterate_diff_noalloc(bp_context *c, const double *other, double *out) {
    RETCHK(ready(c));
    if (!other || !out) return set_err(BP_ERR_ARG, "NULL argument");
    // scratch layout: fid [nports] | max-abs [nports] | staged `other` [hw * nmsg doubles;
    // its offset, 16 * nports bytes, keeps complex entries 16-byte aligned] | result [1]
    const long long nd = c->hw() * c->nmsg;
    const long long need = 2LL * c->nports + nd + 1;
    if (need > BP_SCRATCH_DOUBLES)
        return set_err(BP_ERR_ARG,
                       "bp_iterate_diff_noalloc: 2 * %d slots + %lld message doubles + 1 exceed "
                       "the %d-double scratch; use bp_iterate_diff",
                       c->nports, nd, BP_SCRATCH_DOUBLES);
    HIPCHK(hipSetDevice(c->device));
    double *sc = g_scratch_ptr[c->device];
    if (!sc) return set_err(BP_ERR_STATE, "no scratch address for device %d", c->device);
    double *fid = sc, *mab = sc + c->nports;
    double *stage = mab + c->nports, *res = stage + nd;
    if (nd > 0)
        HIPCHK(hipMemcpyAsync(stage, other, sizeof(double) * nd, hipMemcpyHostToDevice,
                              c->stream));
    RETCHK(diff_launch(c, stage, true, fid, mab));
    hipLaunchKernelGGL(k_max_slots, dim3(1), dim3(256), 0, c->stream, fid, c->nports, res);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, res, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return BP_OK;
}

int bp_polish_fp64(bp_context *c, int nsweeps, double *out_delta) {
    RETCHK(ready(c));
    RETCHK(real_only(c, "bp_polish_fp64"));
    HIPCHK(hipSetDevice(c->device));
    if (!c->d_msg64[0]) RETCHK(dev_alloc((void **)&c->d_msg64[0], sizeof(double) * c->nmsg));
    if (!c->d_msg64[1]) RETCHK(dev_alloc((void **)&c->d_msg64[1], sizeof(double) * c->nmsg));
    if (!c->fp64_active) {
        const int blocks = (int)std::max<long long>(1, std::min<long long>((c->nmsg + 255) / 256, 65535));
        if (c->dtype == BP_F64)
            HIPCHK(hipMemcpyAsync(c->d_msg64[0], c->d_msg[c->cur], sizeof(double) * c->nmsg,
                                  hipMemcpyDeviceToDevice, c->stream));
        else
            hipLaunchKernelGGL((k_convert<float, double>), dim3(blocks), dim3(256), 0, c->stream,
                               (const float *)c->d_msg[c->cur], c->d_msg64[0], c->nmsg);
        c->cur64 = 0;
        c->fp64_active = true;
    }
    // FP64 sweeps on an FP64 vertex kernel when one covers the networks (largest bond
    // dimensions 6..8 or 9..16 in FP32/BF16 contexts), else on the generic kernel
    if (c->normalize) RETCHK(ensure_v64(c));
    const bool v64 = c->normalize && c->v64_ok && c->path != BP_PATH_GENERIC;
    for (int i = 0; i < nsweeps; ++i) {
        if (v64) {
            RETCHK(v64_sweep(c, c->d_msg64[c->cur64], c->d_msg64[c->cur64 ^ 1],
                             make_ctl(c, false)));
        } else {
            RETCHK(launch_update_generic<double>(c, c->d_sites64, c->d_msg64[c->cur64],
                                                 c->d_msg64[c->cur64 ^ 1], make_ctl(c, false), 0,
                                                 c->nports, nullptr, gen_flags(c, false)));
        }
        c->cur64 ^= 1;
    }
    if (out_delta) {
        std::vector<double> r(c->nports);
        HIPCHK(hipMemcpyAsync(r.data(), c->d_res_abs, sizeof(double) * c->nports,
                              hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (int n = 0; n < c->nnet; ++n) {
            double m = 0;
            for (int q = c->pbeg[n]; q < c->pbeg[n + 1]; ++q) m = std::max(m, r[q]);
            out_delta[n] = nsweeps > 0 ? m : 0.0;
        }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return BP_OK;
}

static int grad_lognorm_impl(bp_context *c, const GradOut &go) {
    RETCHK(ready(c));
    RETCHK(real_only(c, "log-norm gradient"));
    if (!c->normalize) return set_err(BP_ERR_UNSUPPORTED, "gradients need normalised BP (option normalize)");
    if (c->plain) return set_err(BP_ERR_UNSUPPORTED, "gradient needs a norm network");
    HIPCHK(hipSetDevice(c->device));
    bool is64 = false;
    const void *m = current_msgs(c, &is64);
    if (vertex_adjoint_ok(c, is64)) {
        const VdCtx vc{&c->vplan, c->stream, c->topo(), (const float *)c->d_vsites};
        const int rv = vadj_grad_lognorm(&c->adj, vc, (const float *)m, c->nsite, go);
        return rv == BP_OK ? BP_OK : set_err(rv, "log-norm gradient failed (status %d)", rv);
    }
    const int rc = adjoint_grad_lognorm(&c->adj, c->stream, c->topo(), c->max_site, c->max_mats,
                                        c->maxD, is64 ? BP_F64 : c->dtype,
                                        is64 ? (const void *)c->d_sites64
                                             : (const void *)c->d_sites,
                                        m, c->nsite, go);
    return rc == BP_OK ? BP_OK : set_err(rc, "log-norm gradient failed (status %d)", rc);
}

static int tfim_impl(bp_context *c, double J, double h, double *energy, const GradOut &go,
                     double adj_tol, int adj_maxiter, int *out_iters, double *out_res) {
    RETCHK(ready(c));
    RETCHK(real_only(c, "TFIM energy"));
    if (!energy) return set_err(BP_ERR_ARG, "NULL argument");
    if (c->plain) return set_err(BP_ERR_UNSUPPORTED, "energy needs a norm network");
    if (!c->normalize && go.any()) return set_err(BP_ERR_UNSUPPORTED, "gradients need normalised BP (option normalize)");
    for (int u = 0; u < c->nv; ++u)
        if (c->phys_dim[u] != 2) return set_err(BP_ERR_UNSUPPORTED, "TFIM needs phys dim 2");
    HIPCHK(hipSetDevice(c->device));
    bool is64 = false;
    const void *m = current_msgs(c, &is64);
    if (vertex_adjoint_ok(c, is64)) {
        const VdCtx vc{&c->vplan, c->stream, c->topo(), (const float *)c->d_vsites};
        const int rv = vadj_tfim(&c->adj, vc, c->ne, (const float *)m, c->nsite, c->nmsg, J, h,
                                 energy, go, adj_tol, adj_maxiter, out_iters, out_res);
        return rv == BP_OK ? BP_OK : set_err(rv, "TFIM energy/gradient failed (status %d)", rv);
    }
    const int rc = adjoint_tfim(&c->adj, c->stream, c->topo(), c->ne, c->max_site, c->max_mats,
                                c->maxD, is64 ? BP_F64 : c->dtype,
                                is64 ? (const void *)c->d_sites64 : (const void *)c->d_sites, m,
                                c->nsite, c->nmsg, J, h, energy, go, adj_tol, adj_maxiter,
                                out_iters, out_res);
    return rc == BP_OK ? BP_OK : set_err(rc, "TFIM energy/gradient failed (status %d)", rc);
}

// operators: ld x ld blocks, ld >= every physical dimension
static int expect_impl(bp_context *c, const double *ops_v, const double *ops_a,
                       const double *ops_b, int ld, double *out) {
    RETCHK(ready(c));
    RETCHK(real_only(c, "expectation values"));
    if (!out || (!ops_v && (!ops_a || !ops_b))) return set_err(BP_ERR_ARG, "NULL argument");
    if (c->plain) return set_err(BP_ERR_UNSUPPORTED, "expectation values need a norm network");
    for (int u = 0; u < c->nv; ++u)
        if (c->phys_dim[u] > ld)
            return set_err(BP_ERR_ARG, "operator block %d smaller than physical dimension %d", ld,
                           c->phys_dim[u]);
    HIPCHK(hipSetDevice(c->device));
    bool is64 = false;
    const void *m = current_msgs(c, &is64);
    const int rc = bp_expect(&c->adj, c->stream, c->topo(), c->ne, c->nmsg, c->max_site,
                             c->max_mats, c->maxD, is64 ? BP_F64 : c->dtype,
                             is64 ? (const void *)c->d_sites64 : (const void *)c->d_sites, m,
                             ops_v, ops_a, ops_b, ld, out);
    return rc == BP_OK ? BP_OK : set_err(rc, "expectation values failed (status %d)", rc);
}

int bp_expect_vertex(bp_context *c, const double *ops, int ld, double *out) {
    if (!ops) return set_err(BP_ERR_ARG, "NULL argument");
    return expect_impl(c, ops, nullptr, nullptr, ld, out);
}

int bp_expect_edge(bp_context *c, const double *ops_first, const double *ops_second, int ld,
                   double *out) {
    return expect_impl(c, nullptr, ops_first, ops_second, ld, out);
}

static int local_impl(bp_context *c, const double *ops_v, const double *ops_first,
                      const double *ops_second, int ld, double *energy, const GradOut &go,
                      double adj_tol, int adj_maxiter, int *out_iters, double *out_res) {
    RETCHK(ready(c));
    RETCHK(real_only(c, "local energy"));
    if (!energy || (!ops_v && (!ops_first || !ops_second)))
        return set_err(BP_ERR_ARG, "NULL argument");
    if ((ops_first == nullptr) != (ops_second == nullptr))
        return set_err(BP_ERR_ARG, "edge operators come in (first, second) pairs");
    if (c->plain) return set_err(BP_ERR_UNSUPPORTED, "energies need a norm network");
    if (!c->normalize && go.any()) return set_err(BP_ERR_UNSUPPORTED, "gradients need normalised BP (option normalize)");
    for (int u = 0; u < c->nv; ++u)
        if (c->phys_dim[u] > ld)
            return set_err(BP_ERR_ARG, "operator block %d smaller than physical dimension %d", ld,
                           c->phys_dim[u]);
    HIPCHK(hipSetDevice(c->device));
    bool is64 = false;
    const void *m = current_msgs(c, &is64);
    const int rc = adjoint_local(&c->adj, c->stream, c->topo(), c->ne, c->max_site, c->max_mats,
                                 c->maxD, is64 ? BP_F64 : c->dtype,
                                 is64 ? (const void *)c->d_sites64 : (const void *)c->d_sites, m,
                                 c->nsite, c->nmsg, ops_v, ops_first, ops_second, ld, energy, go,
                                 adj_tol, adj_maxiter, out_iters, out_res);
    return rc == BP_OK ? BP_OK : set_err(rc, "local energy/gradient failed (status %d)", rc);
}

static GradOut device_grad(const bp_context *c, void *grad) {
    GradOut go;
    go.dev = grad;
    go.dev_f64 = c->dtype == BP_F64;
    return go;
}

int bp_local_energy(bp_context *c, const double *ops_v, const double *ops_first,
                    const double *ops_second, int ld, double *energy, double *grad,
                    double adj_tol, int adj_maxiter, int *out_iters, double *out_res) {
    GradOut go;
    go.host = grad;
    return local_impl(c, ops_v, ops_first, ops_second, ld, energy, go, adj_tol, adj_maxiter,
                      out_iters, out_res);
}

int bp_local_energy_device(bp_context *c, const double *ops_v, const double *ops_first,
                           const double *ops_second, int ld, double *energy, void *grad,
                           double adj_tol, int adj_maxiter, int *out_iters, double *out_res) {
    if (!c || !grad) return set_err(BP_ERR_ARG, "NULL argument");
    return local_impl(c, ops_v, ops_first, ops_second, ld, energy, device_grad(c, grad), adj_tol,
                      adj_maxiter, out_iters, out_res);
}

int bp_grad_lognorm(bp_context *c, double *grad) {
    if (!grad) return set_err(BP_ERR_ARG, "NULL argument");
    GradOut go;
    go.host = grad;
    return grad_lognorm_impl(c, go);
}

int bp_grad_lognorm_device(bp_context *c, void *grad) {
    if (!c || !grad) return set_err(BP_ERR_ARG, "NULL argument");
    return grad_lognorm_impl(c, device_grad(c, grad));
}

int bp_tfim_energy(bp_context *c, double J, double h, double *energy, double *grad,
                   double adj_tol, int adj_maxiter, int *out_iters, double *out_res) {
    GradOut go;
    go.host = grad;
    return tfim_impl(c, J, h, energy, go, adj_tol, adj_maxiter, out_iters, out_res);
}

int bp_tfim_energy_device(bp_context *c, double J, double h, double *energy, void *grad,
                          double adj_tol, int adj_maxiter, int *out_iters, double *out_res) {
    if (!c || !grad) return set_err(BP_ERR_ARG, "NULL argument");
    return tfim_impl(c, J, h, energy, device_grad(c, grad), adj_tol, adj_maxiter, out_iters,
                     out_res);
}

// ------------------------------------------------------------- exact contraction

int bp_cn_plan_create(bp_context *c, int nt, const int *ranks, const int *labels, int n_labels,
                      const int *label_dim, const int *order, int order_alg,
                      long long max_elems, bp_cn_plan **out) {
    if (!c || !out) return set_err(BP_ERR_ARG, "NULL argument");
    *out = nullptr;
    HIPCHK(hipSetDevice(c->device));
    bp_cn_plan *p = new bp_cn_plan();
    p->device = c->device;
    std::string msg;
    int rc = cn_build_plan(nt, ranks, labels, n_labels, label_dim, order, order_alg,
                           max_elems, *p, msg);
    if (rc == 0) rc = cn_alloc_plan(*p, msg);
    if (rc != 0) {
        cn_free_plan(*p);
        delete p;
        return set_err(rc == 1 ? BP_ERR_ARG : rc == 2 ? BP_ERR_HIP : rc == 4 ? BP_ERR_UNSUPPORTED
                                                                            : BP_ERR_NOMEM,
                       "contraction plan: %s", msg.c_str());
    }
    *out = p;
    return BP_OK;
}

int bp_cn_plan_info(const bp_cn_plan *p, int *out_rank, int *out_labels, long long *out_size,
                    long long *in_size, double *flops, double *bytes,
                    long long *workspace_bytes) {
    if (!p) return set_err(BP_ERR_ARG, "plan is NULL");
    if (out_rank) *out_rank = (int)p->out_labels.size();
    if (out_labels)
        for (size_t i = 0; i < p->out_labels.size(); ++i) out_labels[i] = p->out_labels[i];
    if (out_size) *out_size = p->out_size;
    if (in_size) *in_size = p->in_total;
    if (flops) *flops = p->flops;
    if (bytes) *bytes = p->bytes;
    if (workspace_bytes) *workspace_bytes = (long long)sizeof(double) * p->arena_elems;
    return BP_OK;
}

int bp_cn_plan_destroy(bp_cn_plan *p) {
    if (!p) return BP_OK;
    (void)hipSetDevice(p->device);
    cn_free_plan(*p);
    delete p;
    return BP_OK;
}

// enqueue the plan over device inputs `in`, copy the result to `out` (kind: a hipMemcpyKind),
// synchronise; the launches' device time goes to last_ms
static int cn_run(bp_context *c, bp_cn_plan *p, const double *in, void *out,
                  hipMemcpyKind kind) {
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    const double *res = nullptr;
    HIPCHK(cn_launch(*p, in, c->stream, &res));
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    HIPCHK(hipMemcpyAsync(out, res, sizeof(double) * p->out_size, kind, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->last_ms = ms;
    return BP_OK;
}

int bp_cn_execute(bp_context *c, bp_cn_plan *p, const double *data, double *out) {
    if (!c || !p || !out || (p->in_total > 0 && !data)) return set_err(BP_ERR_ARG, "NULL argument");
    if (c->device != p->device)
        return set_err(BP_ERR_ARG, "plan was made on device %d, context is on %d", p->device,
                       c->device);
    HIPCHK(hipSetDevice(c->device));
    if (!p->staging && p->in_total > 0)
        HIPCHK(hipMalloc((void **)&p->staging, sizeof(double) * p->in_total));
    if (p->in_total > 0)
        HIPCHK(hipMemcpyAsync(p->staging, data, sizeof(double) * p->in_total,
                              hipMemcpyHostToDevice, c->stream));
    return cn_run(c, p, p->staging, out, hipMemcpyDeviceToHost);
}

int bp_cn_execute_device(bp_context *c, bp_cn_plan *p, const void *data, void *out) {
    if (!c || !p || !out || (p->in_total > 0 && !data)) return set_err(BP_ERR_ARG, "NULL argument");
    if (c->device != p->device)
        return set_err(BP_ERR_ARG, "plan was made on device %d, context is on %d", p->device,
                       c->device);
    HIPCHK(hipSetDevice(c->device));
    return cn_run(c, p, (const double *)data, out, hipMemcpyDeviceToDevice);
}

int bp_contract_network(bp_context *c, int nt, const int *ranks, const int *labels,
                        int n_labels, const int *label_dim, const double *data,
                        const int *order, int order_alg, long long max_elems, int *out_rank,
                        int *out_labels, double *out, long long out_cap) {
    if (!out_rank || !out_labels || !out) return set_err(BP_ERR_ARG, "NULL argument");
    bp_cn_plan *p = nullptr;
    RETCHK(bp_cn_plan_create(c, nt, ranks, labels, n_labels, label_dim, order, order_alg,
                             max_elems, &p));
    int rc = BP_OK;
    if (p->out_size > out_cap)
        rc = set_err(BP_ERR_ARG, "result has %lld entries, out_cap is %lld", p->out_size,
                     out_cap);
    if (rc == BP_OK) rc = bp_cn_execute(c, p, data, out);
    if (rc == BP_OK)
        bp_cn_plan_info(p, out_rank, out_labels, nullptr, nullptr, nullptr, nullptr, nullptr);
    bp_cn_plan_destroy(p);
    return rc;
}

// ------------------------------------------------------------- gate application

int bp_gate_sizes(int ngates, const int *desc, long long *in_size, long long *out_size) {
    if (ngates < 0 || (ngates > 0 && !desc)) return set_err(BP_ERR_ARG, "invalid gate list");
    for (int g = 0; g < ngates; ++g) {
        GtLayout L;
        if (gt_layout(desc + (long long)g * GD_N, L))
            return set_err(BP_ERR_ARG, "gate %d: invalid descriptor", g);
        if (in_size) in_size[g] = L.in_total;
        if (out_size) out_size[g] = L.out_total;
    }
    return BP_OK;
}

int bp_apply_gates(bp_context *c, int ngates, const int *desc, const double *in, double *out,
                   int *chi, double rtol) {
    if (!c) return set_err(BP_ERR_ARG, "ctx is NULL");
    if (ngates < 0 || (ngates > 0 && (!desc || !in || !out || !chi)))
        return set_err(BP_ERR_ARG, "NULL argument");
    if (ngates == 0) return BP_OK;
    std::vector<long long> offs((size_t)ngates * GO_N);
    long long tin = 0, tout = 0, tws = 0, lneed = 0;
    for (int g = 0; g < ngates; ++g) {
        GtLayout L;
        if (gt_layout(desc + (long long)g * GD_N, L))
            return set_err(BP_ERR_ARG, "gate %d: invalid descriptor (ports <= %d, matching "
                                       "shared bond, min(m1, m2) <= %d)", g, GT_KMAX, GT_QMAX);
        offs[(size_t)g * GO_N + GO_IN] = tin;
        offs[(size_t)g * GO_N + GO_OUT] = tout;
        offs[(size_t)g * GO_N + GO_WS] = tws;
        tin += L.in_total;
        tout += L.out_total;
        tws += (L.ws_total + 31) / 32 * 32;
        lneed = std::max(lneed, gt_lds_need(L));
    }
    // Jacobi sweeps staged in LDS up to 120 KiB per workgroup (matrices beyond stay in HBM)
    const long long lds_cap = std::min<long long>(lneed, 120 * 1024 / 8);
    const size_t lds_bytes = sizeof(double) * (size_t)lds_cap;
    HIPCHK(hipSetDevice(c->device));
    int *d_desc = nullptr, *d_chi = nullptr;
    long long *d_offs = nullptr;
    double *d_in = nullptr, *d_out = nullptr, *d_ws = nullptr;
    auto release = [&]() {
        if (d_desc) (void)hipFree(d_desc);
        if (d_chi) (void)hipFree(d_chi);
        if (d_offs) (void)hipFree(d_offs);
        if (d_in) (void)hipFree(d_in);
        if (d_out) (void)hipFree(d_out);
        if (d_ws) (void)hipFree(d_ws);
    };
    if (hipMalloc((void **)&d_desc, sizeof(int) * GD_N * (size_t)ngates) != hipSuccess ||
        hipMalloc((void **)&d_chi, sizeof(int) * (size_t)ngates) != hipSuccess ||
        hipMalloc((void **)&d_offs, sizeof(long long) * offs.size()) != hipSuccess ||
        hipMalloc((void **)&d_in, sizeof(double) * std::max(tin, 1LL)) != hipSuccess ||
        hipMalloc((void **)&d_out, sizeof(double) * std::max(tout, 1LL)) != hipSuccess ||
        hipMalloc((void **)&d_ws, sizeof(double) * std::max(tws, 1LL)) != hipSuccess) {
        release();
        return set_err(BP_ERR_NOMEM, "gate workspace (%lld doubles) does not fit", tws);
    }
    hipError_t e = hipMemcpyAsync(d_desc, desc, sizeof(int) * GD_N * (size_t)ngates,
                                  hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_offs, offs.data(), sizeof(long long) * offs.size(),
                           hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(d_in, in, sizeof(double) * tin, hipMemcpyHostToDevice, c->stream);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_gate,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
    if (e == hipSuccess) {
        hipLaunchKernelGGL(k_gate, dim3((unsigned)ngates), dim3(GT_NT), lds_bytes, c->stream,
                           (const int *)d_desc, (const long long *)d_offs, (const double *)d_in,
                           d_out, d_ws, d_chi, rtol, lds_cap);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out, d_out, sizeof(double) * tout, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(chi, d_chi, sizeof(int) * (size_t)ngates, hipMemcpyDeviceToHost,
                           c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
    release();
    if (e != hipSuccess) return set_err(BP_ERR_HIP, "gate application: %s", hipGetErrorString(e));
    c->last_ms = ms;
    return BP_OK;
}

int bp_apply_gates_resident(bp_context *c, int ngates, const int *kind, const int *where,
                            const double *gates, int maxdim, int normalize, double rtol,
                            int *out_chi, double *out_s, int s_stride) {
    RETCHK(ready(c));
    RETCHK(real_only(c, "bp_apply_gates_resident"));
    if (ngates < 0 || (ngates > 0 && (!kind || !where || !gates)))
        return set_err(BP_ERR_ARG, "NULL argument");
    if (c->plain) return set_err(BP_ERR_UNSUPPORTED, "gates act on norm networks");
    if (out_s && s_stride < 1) return set_err(BP_ERR_ARG, "s_stride must be positive");
    if (ngates == 0) return BP_OK;
    std::vector<int> desc((size_t)ngates * GD_N, 0);
    std::vector<long long> offs((size_t)ngates * GO_N), rec((size_t)ngates * GS_N, 0), seg;
    std::vector<char> used(c->nv, 0);
    long long tin = 0, tout = 0, tws = 0, lneed = 0, gsum = 0;
    for (int g = 0; g < ngates; ++g) {
        int *d = desc.data() + (size_t)g * GD_N;
        long long *rr = rec.data() + (size_t)g * GS_N;
        int us[2], ps[2], ks[2];
        if (kind[g] == 2) {
            const int q1 = where[g];
            if (q1 < 0 || q1 >= c->nports) return set_err(BP_ERR_ARG, "gate %d: slot %d", g, q1);
            us[0] = c->owner[q1];
            us[1] = c->port_nbr[q1];
            const int q2 = c->port_rev[q1];
            ps[0] = q1 - c->port_off[us[0]];
            ps[1] = q2 - c->port_off[us[1]];
        } else if (kind[g] == 1) {
            if (where[g] < 0 || where[g] >= c->nv)
                return set_err(BP_ERR_ARG, "gate %d: vertex %d", g, where[g]);
            us[0] = where[g];
            ps[0] = -1;
        } else {
            return set_err(BP_ERR_ARG, "gate %d: kind %d", g, kind[g]);
        }
        const int ns = kind[g];
        for (int s2 = 0; s2 < ns; ++s2) {
            if (used[us[s2]])
                return set_err(BP_ERR_ARG, "gate %d touches vertex %d of an earlier gate of the "
                                           "same call", g, us[s2]);
            used[us[s2]] = 1;
            ks[s2] = c->port_off[us[s2] + 1] - c->port_off[us[s2]];
            if (ks[s2] > GT_KMAX)
                return set_err(BP_ERR_UNSUPPORTED, "vertex %d has %d ports (at most %d)",
                               us[s2], ks[s2], GT_KMAX);
        }
        d[GD_KIND] = ns;
        d[GD_D1] = c->phys_dim[us[0]];
        d[GD_D2] = ns == 2 ? c->phys_dim[us[1]] : 0;
        d[GD_K1] = ks[0];
        d[GD_K2] = ns == 2 ? ks[1] : 0;
        d[GD_P1] = ps[0];
        d[GD_P2] = ns == 2 ? ps[1] : -1;
        d[GD_NORM] = normalize ? 1 : 0;
        for (int s2 = 0; s2 < ns; ++s2)
            for (int j = 0; j < ks[s2]; ++j)
                d[(s2 ? GD_DIMS2 : GD_DIMS1) + j] =
                    c->edge_dim[c->port_edge[c->port_off[us[s2]] + j]];
        const int De = ns == 2 ? d[GD_DIMS1 + ps[0]] : 0;
        d[GD_MAXDIM] = maxdim > 0 && maxdim < De ? maxdim : De;  // the bond keeps De
        GtLayout L;
        if (gt_layout(d, L)) return set_err(BP_ERR_ARG, "gate %d: unsupported shape", g);
        offs[(size_t)g * GO_N + GO_IN] = tin;
        offs[(size_t)g * GO_N + GO_OUT] = tout;
        offs[(size_t)g * GO_N + GO_WS] = tws;
        // input segments: sites, the messages into every port (but the bond's), the gate
        long long at = tin;
        for (int s2 = 0; s2 < ns; ++s2) {
            const long long sz = c->site_off[us[s2] + 1] - c->site_off[us[s2]];
            seg.insert(seg.end(), {0, c->site_off[us[s2]], at, sz});
            at += sz;
        }
        for (int s2 = 0; s2 < ns; ++s2) {
            if (ns == 1 && !normalize) break;
            for (int j = 0; j < ks[s2]; ++j) {
                if (ns == 2 && j == ps[s2]) continue;
                const int q = c->port_off[us[s2]] + j, sl = c->port_rev[q];
                const long long len = c->msg_off[sl + 1] - c->msg_off[sl];
                seg.insert(seg.end(), {1, c->msg_off[sl], at, len});
                at += len;
            }
        }
        const long long glen = ns == 2 ? (long long)d[GD_D1] * d[GD_D2] * d[GD_D1] * d[GD_D2]
                                       : (long long)d[GD_D1] * d[GD_D1];
        seg.insert(seg.end(), {2, gsum, at, glen});
        at += glen;
        gsum += glen;
        if (at - tin != L.in_total) return set_err(BP_ERR_ARG, "gate %d: packing mismatch", g);
        rr[GS_KIND] = ns;
        rr[GS_OUT] = tout;
        rr[GS_CAP1] = L.out1;
        rr[GS_CAP2] = L.out2;
        rr[GS_SITE1] = c->site_off[us[0]];
        rr[GS_SITE2] = ns == 2 ? c->site_off[us[1]] : 0;
        rr[GS_D1] = d[GD_D1];
        rr[GS_D2] = d[GD_D2];
        rr[GS_K1] = ks[0];
        rr[GS_K2] = ns == 2 ? ks[1] : 0;
        rr[GS_P1] = ps[0];
        rr[GS_P2] = ns == 2 ? ps[1] : -1;
        rr[GS_DE] = De;
        if (ns == 2) {
            const int q1 = where[g];
            rr[GS_MSG1] = c->msg_off[q1];
            rr[GS_MSG2] = c->msg_off[c->port_rev[q1]];
        }
        for (int j = 0; j < GT_KMAX; ++j) {
            rr[GS_DIMS1 + j] = d[GD_DIMS1 + j];
            rr[GS_DIMS2 + j] = d[GD_DIMS2 + j];
        }
        if (out_s && ns == 2 && De > s_stride)
            return set_err(BP_ERR_ARG, "s_stride %d below the bond dimension %d", s_stride, De);
        tin = at;
        tout += L.out_total;
        tws += (L.ws_total + 31) / 32 * 32;
        lneed = std::max(lneed, gt_lds_need(L));
    }
    const long long lds_cap = std::min<long long>(lneed, 120 * 1024 / 8);
    const size_t lds_bytes = sizeof(double) * (size_t)lds_cap;
    const int nseg = (int)(seg.size() / 4);
    HIPCHK(hipSetDevice(c->device));
    bool is64 = false;
    const void *msgs = current_msgs(c, &is64);
    int *d_desc = nullptr, *d_chi = nullptr;
    long long *d_offs = nullptr, *d_seg = nullptr, *d_rec = nullptr;
    double *d_gates = nullptr, *d_in = nullptr, *d_out = nullptr, *d_ws = nullptr, *d_s = nullptr;
    auto release = [&]() {
        for (void *ptr : {(void *)d_desc, (void *)d_chi, (void *)d_offs, (void *)d_seg,
                          (void *)d_rec, (void *)d_gates, (void *)d_in, (void *)d_out,
                          (void *)d_ws, (void *)d_s})
            if (ptr) (void)hipFree(ptr);
    };
    const size_t sbytes = out_s ? sizeof(double) * (size_t)ngates * s_stride : 0;
    if (hipMalloc((void **)&d_desc, sizeof(int) * desc.size()) != hipSuccess ||
        hipMalloc((void **)&d_chi, sizeof(int) * (size_t)ngates) != hipSuccess ||
        hipMalloc((void **)&d_offs, sizeof(long long) * offs.size()) != hipSuccess ||
        hipMalloc((void **)&d_seg, sizeof(long long) * std::max<size_t>(seg.size(), 1)) != hipSuccess ||
        hipMalloc((void **)&d_rec, sizeof(long long) * rec.size()) != hipSuccess ||
        hipMalloc((void **)&d_gates, sizeof(double) * std::max(gsum, 1LL)) != hipSuccess ||
        hipMalloc((void **)&d_in, sizeof(double) * std::max(tin, 1LL)) != hipSuccess ||
        hipMalloc((void **)&d_out, sizeof(double) * std::max(tout, 1LL)) != hipSuccess ||
        hipMalloc((void **)&d_ws, sizeof(double) * std::max(tws, 1LL)) != hipSuccess ||
        (out_s && hipMalloc((void **)&d_s, sbytes) != hipSuccess)) {
        release();
        return set_err(BP_ERR_NOMEM, "gate workspace (%lld doubles) does not fit", tws);
    }
    hipError_t e = hipSuccess;
    auto up = [&](void *dst, const void *src, size_t n) {
        if (e == hipSuccess) e = hipMemcpyAsync(dst, src, n, hipMemcpyHostToDevice, c->stream);
    };
    up(d_desc, desc.data(), sizeof(int) * desc.size());
    up(d_offs, offs.data(), sizeof(long long) * offs.size());
    up(d_seg, seg.data(), sizeof(long long) * seg.size());
    up(d_rec, rec.data(), sizeof(long long) * rec.size());
    up(d_gates, gates, sizeof(double) * gsum);
    if (e == hipSuccess && d_s) e = hipMemsetAsync(d_s, 0, sbytes, c->stream);
    if (e == hipSuccess)
        e = hipFuncSetAttribute((const void *)k_gate,
                                hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds_bytes);
    if (e == hipSuccess) e = hipEventRecord(c->ev0, c->stream);
    const unsigned gblocks = (unsigned)std::min(nseg, 65535);
    double *m64 = nullptr;
    if (e == hipSuccess) {
        if (is64)
            hipLaunchKernelGGL(k_gate_gather<double>, dim3(gblocks), dim3(256), 0, c->stream,
                               (const long long *)d_seg, nseg, (const double *)c->d_sites64,
                               (const double *)msgs, (const double *)d_gates, d_in);
        else
            hipLaunchKernelGGL(k_gate_gather<float>, dim3(gblocks), dim3(256), 0, c->stream,
                               (const long long *)d_seg, nseg, (const double *)c->d_sites64,
                               (const float *)msgs, (const double *)d_gates, d_in);
        hipLaunchKernelGGL(k_gate, dim3((unsigned)ngates), dim3(GT_NT), lds_bytes, c->stream,
                           (const int *)d_desc, (const long long *)d_offs, (const double *)d_in,
                           d_out, d_ws, d_chi, rtol, lds_cap);
        // the working messages, and the FP64 polish messages too when they are current
        if (c->fp64_active) {
            m64 = c->d_msg64[c->cur64];
            if (c->dtype == BP_F64)
                hipLaunchKernelGGL(k_gate_scatter<double>, dim3((unsigned)ngates), dim3(256), 0,
                                   c->stream, (const long long *)d_rec, (const int *)d_chi,
                                   (const double *)d_out, c->d_sites64, (double *)c->d_msg[c->cur],
                                   m64, d_s, s_stride);
            else
                hipLaunchKernelGGL(k_gate_scatter<float>, dim3((unsigned)ngates), dim3(256), 0,
                                   c->stream, (const long long *)d_rec, (const int *)d_chi,
                                   (const double *)d_out, c->d_sites64, (float *)c->d_msg[c->cur],
                                   m64, d_s, s_stride);
        } else if (is64) {
            hipLaunchKernelGGL(k_gate_scatter<double>, dim3((unsigned)ngates), dim3(256), 0,
                               c->stream, (const long long *)d_rec, (const int *)d_chi,
                               (const double *)d_out, c->d_sites64, (double *)c->d_msg[c->cur],
                               (double *)nullptr, d_s, s_stride);
        } else {
            hipLaunchKernelGGL(k_gate_scatter<float>, dim3((unsigned)ngates), dim3(256), 0,
                               c->stream, (const long long *)d_rec, (const int *)d_chi,
                               (const double *)d_out, c->d_sites64, (float *)c->d_msg[c->cur],
                               (double *)nullptr, d_s, s_stride);
        }
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(c->ev1, c->stream);
    if (e == hipSuccess && out_chi)
        e = hipMemcpyAsync(out_chi, d_chi, sizeof(int) * (size_t)ngates, hipMemcpyDeviceToHost,
                           c->stream);
    if (e == hipSuccess && out_s)
        e = hipMemcpyAsync(out_s, d_s, sbytes, hipMemcpyDeviceToHost, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    float ms = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, c->ev0, c->ev1);
    release();
    if (e != hipSuccess) return set_err(BP_ERR_HIP, "resident gates: %s", hipGetErrorString(e));
    c->last_ms = ms;
    RETCHK(refresh_sites(c));
    return BP_OK;
}

}  // extern "C"
