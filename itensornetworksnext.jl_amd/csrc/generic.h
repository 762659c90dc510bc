// generic.h -- runtime-shaped BP kernels: any bond dimension per edge, any degree up to
// BP_MAXR-1, FP64 or FP32, norm (double-layer, matrix messages) or plain (vector
// messages) networks.  One workgroup per directed message (or per vertex for the
// scalars), grid-stride over work items.  Intermediates live in LDS when two copies of
// the largest site tensor fit, otherwise in a per-workgroup global workspace.
//
// Message update restated from SimpleMessageUpdate
// (src/beliefpropagation/beliefpropagation.jl:242-257): for slot q = (u -> v),
//   out[a,a'] = sum A_u[s,..a..,X] conj(A_u)[s,..a'..,X'] prod_{w != v} m_{w->u}[x_w,x'_w]
// contracted as k-1 ket-side mode products followed by one open contraction, then
// divided by its sum when the sum is nonzero.
#pragma once

#include "common.h"

// out[pre, b, post] = sum_a in[pre, a, post] * m[a, b]; m is na x nb (row-major).
// trans != 0 uses m[b, a] instead (m square).
#define GEN_FLAG_BF16 1          // BF16 product operands (FP32 kernels)
#define GEN_FLAG_NO_NORMALIZE 2  // skip the sum normalisation

template <typename T>
__device__ void block_mode_product(const T *in, T *out, const int *dims, int r, int ax,
                                   const T *m, int nb, int trans = 0) {
    long long pre = 1, post = 1;
    for (int i = 0; i < ax; ++i) pre *= dims[i];
    for (int i = ax + 1; i < r; ++i) post *= dims[i];
    const int na = dims[ax];
    const long long ncol = pre * post;
    constexpr int NB = sizeof(T) > 8 ? 8 : 16;  // outputs per thread (complex: 8, no spills)
    for (long long c = threadIdx.x; c < ncol; c += blockDim.x) {
        const long long p = c / post, t = c - p * post;
        const T *src = in + p * na * post + t;
        T *dst = out + p * (long long)nb * post + t;
        for (int b0 = 0; b0 < nb; b0 += NB) {
            T acc[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) acc[b] = T(0);
            if (trans) {
                for (int a = 0; a < na; ++a) {
                    const T x = src[(long long)a * post];
#pragma unroll
                    for (int b = 0; b < NB; ++b)
                        if (b0 + b < nb) acc[b] += x * m[(long long)(b0 + b) * na + a];
                }
            } else {
                for (int a = 0; a < na; ++a) {
                    const T x = src[(long long)a * post];
                    const T *mr = m + (long long)a * nb + b0;
#pragma unroll
                    for (int b = 0; b < NB; ++b)
                        if (b0 + b < nb) acc[b] += x * mr[b];
                }
            }
#pragma unroll
            for (int b = 0; b < NB; ++b)
                if (b0 + b < nb) dst[(long long)(b0 + b) * post] = acc[b];
        }
    }
}

// out[a, a'] = sum_{pre,post} cur[pre, a, post] * B[pre, a', post] on axis ax.
// red: LDS scratch of max(n*n, blockDim) doubles; out: n*n (LDS).
template <typename T>
__device__ void block_open_contract(const T *cur, const T *B, const int *dims, int r, int ax,
                                    double *red, T *out) {
    long long pre = 1, post = 1;
    for (int i = 0; i < ax; ++i) pre *= dims[i];
    for (int i = ax + 1; i < r; ++i) post *= dims[i];
    const int n = dims[ax];
    const long long ncol = pre * post;
    const int pairs = n * n;
    const int G = pairs >= (int)blockDim.x ? 1 : (int)blockDim.x / pairs;
    for (int item = threadIdx.x; item < pairs * G; item += blockDim.x) {
        const int pr = item % pairs, g = item / pairs;
        const int a = pr / n, b = pr - a * n;
        T acc = T(0);
        for (long long c = g; c < ncol; c += G) {
            const long long p = c / post, t = c - p * post;
            acc += cur[(p * n + a) * post + t] * B[(p * n + b) * post + t];
        }
        red[item] = (double)acc;
    }
    __syncthreads();
    for (int pr = threadIdx.x; pr < pairs; pr += blockDim.x) {
        T s = T(0);
        for (int g = 0; g < G; ++g) s += (T)red[g * pairs + pr];
        out[pr] = s;
    }
    __syncthreads();
}

struct VertexShape {
    int r;
    int k;
    int dims[BP_MAXR];
    long long n;
};

__device__ __forceinline__ VertexShape vertex_shape(const DevTopo &tp, int u) {
    VertexShape s;
    const int base = tp.port_off[u];
    s.k = tp.port_off[u + 1] - base;
    s.r = s.k + 1;
    s.dims[0] = tp.plain ? 1 : tp.phys_dim[u];
    s.n = s.dims[0];
    for (int j = 0; j < s.k; ++j) {
        s.dims[1 + j] = tp.edge_dim[tp.port_edge[base + j]];
        s.n *= s.dims[1 + j];
    }
    return s;
}

// Dynamic LDS carve-up shared by the generic kernels.
template <typename T>
struct GenSmem {
    T *mats;      // incoming matrices (sum over ports of D^2, or D for plain)
    T *out;       // one message (max D^2)
    double *red;  // reduction scratch (max(maxD^2, blockDim) doubles)
    T *buf0, *buf1;
};

template <typename T>
__device__ GenSmem<T> carve(char *smem, int mats_elems, int out_elems, int red_elems,
                            long long buf_elems, int use_lds, T *ws, long long ws_stride) {
    GenSmem<T> s;
    size_t off = 0;
    s.red = reinterpret_cast<double *>(smem + off);
    off += sizeof(double) * red_elems;
    s.mats = reinterpret_cast<T *>(smem + off);
    off += sizeof(T) * mats_elems;
    s.out = reinterpret_cast<T *>(smem + off);
    off += sizeof(T) * out_elems;
    off = (off + 15) & ~size_t(15);
    if (use_lds) {
        s.buf0 = reinterpret_cast<T *>(smem + off);
        s.buf1 = s.buf0 + buf_elems;
    } else {
        s.buf0 = ws + (long long)blockIdx.x * ws_stride;
        s.buf1 = s.buf0 + buf_elems;
    }
    return s;
}

// x rounded to BF16 (round to nearest even; infinities and NaN pass through).
__device__ __forceinline__ float bf16r(float x) {
    unsigned u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return x;
    u += 0x7fffu + ((u >> 16) & 1u);
    return __uint_as_float(u & 0xffff0000u);
}

template <typename T>
__device__ __forceinline__ void block_round_bf16(T *x, long long n) {
    for (long long i = threadIdx.x; i < n; i += blockDim.x) x[i] = (T)bf16r((float)x[i]);
}

// Load the incoming message of every port of u into LDS; returns per-port offsets.
template <typename T>
__device__ void load_incoming(const DevTopo &tp, int u, const VertexShape &vs, const T *msgs,
                             T *mats, int *moff, int skip) {
    const int base = tp.port_off[u];
    int off = 0;
    for (int j = 0; j < vs.k; ++j) {
        moff[j] = off;
        const int D = vs.dims[1 + j];
        const int len = tp.plain ? D : D * D;
        if (j != skip) {
            const T *src = msgs + tp.msg_off[tp.port_rev[base + j]];
            for (int i = threadIdx.x; i < len; i += blockDim.x) mats[off + i] = src[i];
        }
        off += len;
    }
}

// One directed message per workgroup iteration: update, normalise, residual.  Work items
// i in [item_begin, item_end) are message slots, or slot_list[i] when a list is given (the
// slots of one dispatch bucket).
// flags: GEN_FLAG_NO_NORMALIZE keeps the unnormalised message (SimpleMessageUpdate with
// normalize = false, beliefpropagation.jl:237-253); GEN_FLAG_BF16 (FP32 only): every
// product operand is rounded to BF16 and sums stay in FP32, as
// on the vertex kernels: the sites arrive rounded, each incoming message is scaled by
// 1 / max|m| (the scale cancels in the sum normalisation) and rounded, and every partial
// product is rounded before it enters the next product.
template <typename T, int BS = BP_BLOCK>
__global__ __launch_bounds__(BS) void k_update_generic(
    DevTopo tp, const T *__restrict__ sites, const T *mold, T *mnew,
    NetCtl ctl, double *__restrict__ res_fid, double *__restrict__ res_abs,
    int item_begin, int item_end, const int *__restrict__ slot_list, int mats_elems,
    int out_elems, int red_elems, long long buf_elems, int use_lds, T *ws, long long ws_stride,
    int flags) {
    extern __shared__ __align__(16) char smem[];
    GenSmem<T> sm = carve<T>(smem, mats_elems, out_elems, red_elems, buf_elems, use_lds, ws,
                             ws_stride);
    const int bf16 = flags & GEN_FLAG_BF16;
    const bool normalize = !(flags & GEN_FLAG_NO_NORMALIZE);
    for (int it = item_begin + blockIdx.x; it < item_end; it += gridDim.x) {
        const int q = slot_list ? slot_list[it] : it;
        const int u = tp.owner[q];
        const int base = tp.port_off[u];
        const int p = q - base;
        const long long mo = tp.msg_off[q];
        const int len = (int)(tp.msg_off[q + 1] - mo);
        if (ctl.enabled && ctl.done[tp.vnet[u]]) {
            for (int i = threadIdx.x; i < len; i += blockDim.x) mnew[mo + i] = mold[mo + i];
            if (threadIdx.x == 0) {
                res_fid[q] = 0.0;
                res_abs[q] = 0.0;
            }
            continue;
        }
        VertexShape vs = vertex_shape(tp, u);
        int moff[BP_MAXR];
        load_incoming(tp, u, vs, mold, sm.mats, moff, p);
        __syncthreads();
        if (bf16) {
            for (int j = 0; j < vs.k; ++j) {
                if (j == p) continue;
                const int D = vs.dims[1 + j];
                const int mlen = tp.plain ? D : D * D;
                double mx = 0.0;
                for (int i = threadIdx.x; i < mlen; i += blockDim.x)
                    mx = fmax(mx, fabs((double)sm.mats[moff[j] + i]));
                mx = block_max(mx, sm.red);
                const T sc = mx > 0.0 ? (T)(1.0 / mx) : T(1);
                for (int i = threadIdx.x; i < mlen; i += blockDim.x)
                    sm.mats[moff[j] + i] = (T)bf16r((float)(sm.mats[moff[j] + i] * sc));
                __syncthreads();
            }
        }
        const T *A = sites + tp.site_off[u];
        const T *cur = A;
        T *bufs[2] = {sm.buf0, sm.buf1};
        int which = 0;
        for (int j = 0; j < vs.k; ++j) {
            if (j == p) continue;
            const int nb = tp.plain ? 1 : vs.dims[1 + j];
            block_mode_product(cur, bufs[which], vs.dims, vs.r, 1 + j, sm.mats + moff[j], nb);
            if (tp.plain) vs.dims[1 + j] = 1;
            __syncthreads();
            if (bf16) {
                long long n = 1;
                for (int i = 0; i < vs.r; ++i) n *= vs.dims[i];
                block_round_bf16(bufs[which], n);
                __syncthreads();
            }
            cur = bufs[which];
            which ^= 1;
        }
        if (tp.plain) {
            for (int i = threadIdx.x; i < len; i += blockDim.x) sm.out[i] = cur[i];
            __syncthreads();
        } else {
            block_open_contract(cur, A, vs.dims, vs.r, 1 + p, sm.red, sm.out);
        }
        // sum normalisation (beliefpropagation.jl:248-253) and residuals
        double ls = 0.0;
        for (int i = threadIdx.x; i < len; i += blockDim.x) ls += (double)sm.out[i];
        const double S = block_sum(ls, sm.red);
        const T* old = mold + mo;
        double dno = 0.0, dnn = 0.0, doo = 0.0, dmx = 0.0;
        for (int i = threadIdx.x; i < len; i += blockDim.x) {
            T x = sm.out[i];
            if (normalize && S != 0.0) x = (T)((double)x / S);
            const double xo = (double)old[i], xn = (double)x;  // read before an in-place write
            mnew[mo + i] = x;
            dno += xn * xo;
            dnn += xn * xn;
            doo += xo * xo;
            const double d = fabs(xn - xo);
            dmx = (d > dmx || d != d) ? d : dmx;
        }
        dno = block_sum(dno, sm.red);
        dnn = block_sum(dnn, sm.red);
        doo = block_sum(doo, sm.red);
        dmx = block_max(dmx, sm.red);
        if (threadIdx.x == 0) {
            const double c = dno / sqrt(dnn * doo);
            const double fid = 1.0 - c * c;
            res_fid[q] = fid;
            res_abs[q] = dmx;
        }
        __syncthreads();
    }
}

// Z_v = full contraction of vertex v with all incoming messages (FP64 final sum).
template <typename T>
__global__ __launch_bounds__(BP_BLOCK) void k_vertex_scalar_generic(
    DevTopo tp, const T *__restrict__ sites, const T *__restrict__ msgs, double *__restrict__ zv,
    int mats_elems, int out_elems, int red_elems, long long buf_elems, int use_lds, T *ws,
    long long ws_stride) {
    extern __shared__ __align__(16) char smem[];
    GenSmem<T> sm = carve<T>(smem, mats_elems, out_elems, red_elems, buf_elems, use_lds, ws,
                             ws_stride);
    for (int u = blockIdx.x; u < tp.nv; u += gridDim.x) {
        VertexShape vs = vertex_shape(tp, u);
        int moff[BP_MAXR];
        load_incoming(tp, u, vs, msgs, sm.mats, moff, -1);
        __syncthreads();
        const T *A = sites + tp.site_off[u];
        const T *cur = A;
        T *bufs[2] = {sm.buf0, sm.buf1};
        int which = 0;
        for (int j = 0; j < vs.k; ++j) {
            const int nb = tp.plain ? 1 : vs.dims[1 + j];
            block_mode_product(cur, bufs[which], vs.dims, vs.r, 1 + j, sm.mats + moff[j], nb);
            if (tp.plain) vs.dims[1 + j] = 1;
            __syncthreads();
            cur = bufs[which];
            which ^= 1;
        }
        double acc = 0.0;
        if (tp.plain) {
            if (threadIdx.x == 0) acc = (double)cur[0];
        } else {
            for (long long i = threadIdx.x; i < vs.n; i += blockDim.x)
                acc += (double)cur[i] * (double)A[i];
        }
        acc = block_sum(acc, sm.red);
        if (threadIdx.x == 0) zv[u] = acc;
        __syncthreads();
    }
}

// Z_e = <m_uv, m_vu> for every undirected edge (one thread per port pair).
template <typename T>
__global__ void k_edge_scalars(DevTopo tp, const T *__restrict__ msgs, double *__restrict__ ze) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= tp.nports) return;
    const int r = tp.port_rev[q];
    if (q > r) return;
    const long long a = tp.msg_off[q], b = tp.msg_off[r];
    const int len = (int)(tp.msg_off[q + 1] - a);
    double s = 0.0;
    for (int i = 0; i < len; ++i) s += (double)msgs[a + i] * (double)msgs[b + i];
    ze[tp.port_edge[q]] = s;
}
