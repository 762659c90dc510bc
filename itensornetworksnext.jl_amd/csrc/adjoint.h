// adjoint.h -- fixed-point adjoint of the BP map and the BP energy estimator on the GPU.
//
// The forward map is one normalised Jacobi sweep F(m; A) (beliefpropagation.jl:242-257);
// at its fixed point m* = F(m*; A) an objective L(m*, A) has the gradient
//     dL/dA = dL/dA|_m + J_A^T lambda,     lambda = J_m^T lambda + dL/dm,
// solved by adjoint sweeps.  The reference has no AD; this follows the implicit-function
// theorem for the same fixed point.  All contractions reuse the runtime-shaped
// building blocks of generic.h; the per-vertex kernels below evaluate
//   open(op.A, A, M, j)[x,x'] = sum (op.A)[s,..x..,X] A[s,..x'..,X'] prod_{i != j} M_i[x_i,x'_i]
// with M_i the incoming message of port i, except one port p that may carry an override
// matrix W (a cotangent or an environment of the other side of an edge), and op a
// 2x2 operator on the physical index (identity, Pauli X, Pauli Z).
//
// Energy: E = sum_<uv> -J <Z_u Z_v>_BP + sum_v -h <X_v>_BP with <O>_BP the ratio of the
// operator-inserted region contraction to the plain one (edge region {u,v} with the
// boundary messages, vertex region {v} with its incoming messages).
#pragma once

#include <cmath>
#include <type_traits>
#include <vector>

#include "common.h"
#include "generic.h"

enum { OP_I = 0, OP_X = 1, OP_Z = 2 };

// Ends an FP32 adjoint iteration at its rounding floor.  The Jacobian products round at
// about 1e-7 of their size, so the relative change stalls near 1e-7 (8e-8 at 32x32 D=8):
// once it is below 1e-5 and has not fallen by 30% over 32 sweeps, more sweeps do not
// converge further; the caller gets the residual reached.
struct AdjFloor {
    double best = INFINITY;
    int best_it = 0;
    bool stalled(double r, int it) {
        if (r < 0.7 * best) {
            best = r;
            best_it = it;
            return false;
        }
        return r < 1e-5 && it - best_it >= 32;
    }
};

__device__ __forceinline__ double phys_op(int op, int sp, int s) {
    if (op == OP_X) return sp != s ? 1.0 : 0.0;
    if (op == OP_Z) return sp == s ? (s == 0 ? 1.0 : -1.0) : 0.0;
    return sp == s ? 1.0 : 0.0;
}

// out[s', X] = sum_s op[s', s] A[s, X]   (tr: op[s, s'])
template <typename T>
__device__ void block_apply_op(const T *A, T *out, int d, long long rest, int op, int tr) {
    for (long long i = threadIdx.x; i < d * rest; i += blockDim.x) {
        const int sp = (int)(i / rest);
        const long long x = i - sp * rest;
        double acc = 0.0;
        for (int s = 0; s < d; ++s)
            acc += (tr ? phys_op(op, s, sp) : phys_op(op, sp, s)) * (double)A[s * rest + x];
        out[i] = (T)acc;
    }
}

template <typename T>
struct AdjSm {
    double *red;
    T *mats, *ov, *out;
    T *b0, *b1, *b2;
};

template <typename T>
__device__ AdjSm<T> adj_carve(char *smem, int mats_elems, int ov_elems, int red_elems,
                              long long buf_elems, int use_lds, T *ws, long long ws_stride) {
    AdjSm<T> s;
    size_t off = 0;
    s.red = reinterpret_cast<double *>(smem);
    off += sizeof(double) * red_elems;
    s.mats = reinterpret_cast<T *>(smem + off);
    off += sizeof(T) * mats_elems;
    s.ov = reinterpret_cast<T *>(smem + off);
    off += sizeof(T) * ov_elems;
    s.out = reinterpret_cast<T *>(smem + off);
    off += sizeof(T) * ov_elems;
    off = (off + 15) & ~size_t(15);
    T *base = use_lds ? reinterpret_cast<T *>(smem + off) : ws + (long long)blockIdx.x * ws_stride;
    s.b0 = base;
    s.b1 = base + buf_elems;
    s.b2 = base + 2 * buf_elems;
    return s;
}

// Mode products over every port except `open` (override W at port ovp), result pointer.
template <typename T>
__device__ const T *absorb_ports(const T *ket, const VertexShape &vs, const T *mats,
                                 const int *moff, int ovp, const T *ov, int open, int trans,
                                 T *b0, T *b1) {
    const T *cur = ket;
    T *bufs[2] = {b0, b1};
    int w = 0;
    for (int j = 0; j < vs.k; ++j) {
        if (j == open) continue;
        const T *m = j == ovp ? ov : mats + moff[j];
        block_mode_product(cur, bufs[w], vs.dims, vs.r, 1 + j, m, vs.dims[1 + j], trans);
        __syncthreads();
        cur = bufs[w];
        w ^= 1;
    }
    return cur;
}

// ket = op.A in b2 (or A itself for the identity)
template <typename T>
__device__ const T *ket_tensor(const T *A, const VertexShape &vs, int op, T *b2) {
    if (op == OP_I) return A;
    block_apply_op(A, b2, vs.dims[0], vs.n / vs.dims[0], op, 0);
    __syncthreads();
    return b2;
}

// ---- unnormalised open contraction for every slot: G[q] = open(op.A, A, m_in, p), S[q] = sum
template <typename T>
__global__ __launch_bounds__(BP_BLOCK) void k_adj_open(DevTopo tp, const T *sites, const T *msgs,
                                                        int op, T *G, double *S, int mats_elems,
                                                        int ov_elems, int red_elems,
                                                        long long buf_elems, int use_lds, T *ws,
                                                        long long ws_stride) {
    extern __shared__ __align__(16) char smem[];
    AdjSm<T> sm = adj_carve<T>(smem, mats_elems, ov_elems, red_elems, buf_elems, use_lds, ws,
                               ws_stride);
    for (int q = blockIdx.x; q < tp.nports; q += gridDim.x) {
        const int u = tp.owner[q];
        const int p = q - tp.port_off[u];
        VertexShape vs = vertex_shape(tp, u);
        int moff[BP_MAXR];
        load_incoming(tp, u, vs, msgs, sm.mats, moff, p);
        __syncthreads();
        const T *A = sites + tp.site_off[u];
        const T *ket = ket_tensor(A, vs, op, sm.b2);
        const T *cur = absorb_ports(ket, vs, sm.mats, moff, -1, sm.ov, p, 0, sm.b0, sm.b1);
        block_open_contract(cur, A, vs.dims, vs.r, 1 + p, sm.red, sm.out);
        const long long mo = tp.msg_off[q];
        const int len = (int)(tp.msg_off[q + 1] - mo);
        double ls = 0.0;
        for (int i = threadIdx.x; i < len; i += blockDim.x) {
            G[mo + i] = sm.out[i];
            ls += (double)sm.out[i];
        }
        ls = block_sum(ls, sm.red);
        if (threadIdx.x == 0) S[q] = ls;
        __syncthreads();
    }
}

// ---- vertex scalars with an operator: zv[u] = open(op.A, A, m_in, -1)
template <typename T>
__global__ __launch_bounds__(BP_BLOCK) void k_adj_vscalar(DevTopo tp, const T *sites,
                                                           const T *msgs, int op, double *zv,
                                                           int mats_elems, int ov_elems,
                                                           int red_elems, long long buf_elems,
                                                           int use_lds, T *ws,
                                                           long long ws_stride) {
    extern __shared__ __align__(16) char smem[];
    AdjSm<T> sm = adj_carve<T>(smem, mats_elems, ov_elems, red_elems, buf_elems, use_lds, ws,
                               ws_stride);
    for (int u = blockIdx.x; u < tp.nv; u += gridDim.x) {
        VertexShape vs = vertex_shape(tp, u);
        int moff[BP_MAXR];
        load_incoming(tp, u, vs, msgs, sm.mats, moff, -1);
        __syncthreads();
        const T *A = sites + tp.site_off[u];
        const T *ket = ket_tensor(A, vs, op, sm.b2);
        const T *cur = absorb_ports(ket, vs, sm.mats, moff, -1, sm.ov, -1, 0, sm.b0, sm.b1);
        double acc = 0.0;
        for (long long i = threadIdx.x; i < vs.n; i += blockDim.x)
            acc += (double)cur[i] * (double)A[i];
        acc = block_sum(acc, sm.red);
        if (threadIdx.x == 0) zv[u] = acc;
        __syncthreads();
    }
}

// ---- expectation values with arbitrary operators (ld x ld row-major blocks, the top-left
// d x d corner used): ket = O.A, out[s', X] = sum_s O[s', s] A[s, X], in b2
template <typename T>
__device__ const T *ket_tensor_m(const T *A, const VertexShape &vs, const double *opm, int ld,
                                 T *b2) {
    const int d = vs.dims[0];
    const long long rest = vs.n / d;
    for (long long i = threadIdx.x; i < vs.n; i += blockDim.x) {
        const int sp = (int)(i / rest);
        const long long x = i - sp * rest;
        double acc = 0.0;
        for (int t = 0; t < d; ++t) acc += opm[sp * ld + t] * (double)A[t * rest + x];
        b2[i] = (T)acc;
    }
    __syncthreads();
    return b2;
}

// zo[u] = open(O_u.A, A, m_in, -1): the operator-inserted vertex scalar
template <typename T>
__global__ __launch_bounds__(BP_BLOCK) void k_expect_vscalar(DevTopo tp, const T *sites,
                                                              const T *msgs, const double *ops,
                                                              int ld, double *zo, int mats_elems,
                                                              int ov_elems, int red_elems,
                                                              long long buf_elems, int use_lds,
                                                              T *ws, long long ws_stride) {
    extern __shared__ __align__(16) char smem[];
    AdjSm<T> sm = adj_carve<T>(smem, mats_elems, ov_elems, red_elems, buf_elems, use_lds, ws,
                               ws_stride);
    for (int u = blockIdx.x; u < tp.nv; u += gridDim.x) {
        VertexShape vs = vertex_shape(tp, u);
        int moff[BP_MAXR];
        load_incoming(tp, u, vs, msgs, sm.mats, moff, -1);
        __syncthreads();
        const T *A = sites + tp.site_off[u];
        const T *ket = ket_tensor_m(A, vs, ops + (long long)u * ld * ld, ld, sm.b2);
        const T *cur = absorb_ports(ket, vs, sm.mats, moff, -1, sm.ov, -1, 0, sm.b0, sm.b1);
        double acc = 0.0;
        for (long long i = threadIdx.x; i < vs.n; i += blockDim.x)
            acc += (double)cur[i] * (double)A[i];
        acc = block_sum(acc, sm.red);
        if (threadIdx.x == 0) zo[u] = acc;
        __syncthreads();
    }
}

// G[q] = open(O.A, A, m_in, p) for every slot q (u -> v through port p of u), unnormalised,
// with O = ops_first[e] when q is the lower slot of its edge e, else ops_second[e]
template <typename T>
__global__ __launch_bounds__(BP_BLOCK) void k_expect_open(DevTopo tp, const T *sites,
                                                           const T *msgs, const double *ops_first,
                                                           const double *ops_second, int ld,
                                                           T *G, int mats_elems, int ov_elems,
                                                           int red_elems, long long buf_elems,
                                                           int use_lds, T *ws,
                                                           long long ws_stride) {
    extern __shared__ __align__(16) char smem[];
    AdjSm<T> sm = adj_carve<T>(smem, mats_elems, ov_elems, red_elems, buf_elems, use_lds, ws,
                               ws_stride);
    for (int q = blockIdx.x; q < tp.nports; q += gridDim.x) {
        const int u = tp.owner[q];
        const int p = q - tp.port_off[u];
        const int e = tp.port_edge[q];
        const double *opm = (q < tp.port_rev[q] ? ops_first : ops_second) + (long long)e * ld * ld;
        VertexShape vs = vertex_shape(tp, u);
        int moff[BP_MAXR];
        load_incoming(tp, u, vs, msgs, sm.mats, moff, p);
        __syncthreads();
        const T *A = sites + tp.site_off[u];
        const T *ket = ket_tensor_m(A, vs, opm, ld, sm.b2);
        const T *cur = absorb_ports(ket, vs, sm.mats, moff, -1, sm.ov, p, 0, sm.b0, sm.b1);
        block_open_contract(cur, A, vs.dims, vs.r, 1 + p, sm.red, sm.out);
        const long long mo = tp.msg_off[q];
        const int len = (int)(tp.msg_off[q + 1] - mo);
        for (int i = threadIdx.x; i < len; i += blockDim.x) G[mo + i] = sm.out[i];
        __syncthreads();
    }
}

__global__ void k_expect_ratio(const double *num, const double *den, int n, double *out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = num[i] / den[i];
}

// ---- cotangent through the sum normalisation f = g / S:  gg = (lam - <lam, f>) / S
template <typename T>
__global__ void k_adj_gg(DevTopo tp, const T *lam, const T *G, const double *S, T *gg) {
    const int q = blockIdx.x;
    if (q >= tp.nports) return;
    const long long mo = tp.msg_off[q];
    const int len = (int)(tp.msg_off[q + 1] - mo);
    const double s = S[q];
    double dot = 0.0;
    for (int i = threadIdx.x; i < len; i += blockDim.x)
        dot += (double)lam[mo + i] * (double)G[mo + i];
    dot = wave_sum(dot);  // one wave per slot
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
        const double l = (double)lam[mo + i];
        gg[mo + i] = (T)(s != 0.0 ? (l - dot / s) / s : l);
    }
}

// ---- message cotangents: for every port p carrying W (slot base+p, or its reverse slot
// when wrev) with coefficient coefq[base+p], and every other port j:
//     mbar[in-slot of j] += coef * open(op.A, A, {p <- W, others <- m_in}, j)
// With W == nullptr there is one term per vertex with coefficient coefv[u] and no
// override (every port j gets open(op.A, A, m_in, j)).  The ket operator is `op` (I, X, Z),
// or an ld x ld matrix: opv[u] for the vertex terms, opq[base+p] for the port terms.
template <typename T>
__global__ __launch_bounds__(BP_BLOCK) void k_adj_msgvjp(
    DevTopo tp, const T *sites, const T *msgs, const T *W, int wrev, const double *coefq,
    const double *coefv, int op, T *mbar, int mats_elems, int ov_elems, int red_elems,
    long long buf_elems, int use_lds, T *ws, long long ws_stride, const double *opv,
    const double *opq, int ld) {
    extern __shared__ __align__(16) char smem[];
    AdjSm<T> sm = adj_carve<T>(smem, mats_elems, ov_elems, red_elems, buf_elems, use_lds, ws,
                               ws_stride);
    for (int u = blockIdx.x; u < tp.nv; u += gridDim.x) {
        VertexShape vs = vertex_shape(tp, u);
        const int base = tp.port_off[u];
        int moff[BP_MAXR];
        load_incoming(tp, u, vs, msgs, sm.mats, moff, -1);
        __syncthreads();
        const T *A = sites + tp.site_off[u];
        const bool per_term = W && opq;
        const T *ket = per_term ? A
                       : (!W && opv) ? ket_tensor_m(A, vs, opv + (long long)u * ld * ld, ld, sm.b2)
                                     : ket_tensor(A, vs, op, sm.b2);
        const int nterm = W ? vs.k : 1;
        for (int t = 0; t < nterm; ++t) {
            const int ovp = W ? t : -1;
            const double c = W ? coefq[base + t] : coefv[u];
            if (c == 0.0) continue;  // uniform over the workgroup
            if (W) {
                const int src = wrev ? tp.port_rev[base + t] : base + t;
                const long long wo = tp.msg_off[src];
                const int wl = (int)(tp.msg_off[src + 1] - wo);
                for (int i = threadIdx.x; i < wl; i += blockDim.x) sm.ov[i] = W[wo + i];
                __syncthreads();
            }
            if (per_term) ket = ket_tensor_m(A, vs, opq + (long long)(base + t) * ld * ld, ld, sm.b2);
            for (int j = 0; j < vs.k; ++j) {
                if (j == ovp) continue;
                const T *cur = absorb_ports(ket, vs, sm.mats, moff, ovp, sm.ov, j, 0, sm.b0,
                                            sm.b1);
                block_open_contract(cur, A, vs.dims, vs.r, 1 + j, sm.red, sm.out);
                const int qin = tp.port_rev[base + j];
                const long long mo = tp.msg_off[qin];
                const int len = (int)(tp.msg_off[qin + 1] - mo);
                for (int i = threadIdx.x; i < len; i += blockDim.x)
                    mbar[mo + i] += (T)(c * (double)sm.out[i]);
                __syncthreads();
            }
        }
    }
}

// ---- site cotangents: abar_u += coef * [ op^T . (A x_all M^T) + (op.A) x_all M ]
// (ket and bra derivatives of <op.A, A>_M for a real site tensor), same term structure
// and operator choice (op, or the matrices opv / opq) as k_adj_msgvjp.
template <typename T>
__global__ __launch_bounds__(BP_BLOCK) void k_adj_sitevjp(
    DevTopo tp, const T *sites, const T *msgs, const T *W, int wrev, const double *coefq,
    const double *coefv, int op, T *abar, int mats_elems, int ov_elems, int red_elems,
    long long buf_elems, int use_lds, T *ws, long long ws_stride, const double *opv,
    const double *opq, int ld) {
    extern __shared__ __align__(16) char smem[];
    AdjSm<T> sm = adj_carve<T>(smem, mats_elems, ov_elems, red_elems, buf_elems, use_lds, ws,
                               ws_stride);
    for (int u = blockIdx.x; u < tp.nv; u += gridDim.x) {
        VertexShape vs = vertex_shape(tp, u);
        const int base = tp.port_off[u];
        int moff[BP_MAXR];
        load_incoming(tp, u, vs, msgs, sm.mats, moff, -1);
        __syncthreads();
        const T *A = sites + tp.site_off[u];
        T *out = abar + tp.site_off[u];
        const int d = vs.dims[0];
        const long long rest = vs.n / d;
        const int nterm = W ? vs.k : 1;
        for (int t = 0; t < nterm; ++t) {
            const int ovp = W ? t : -1;
            const double c = W ? coefq[base + t] : coefv[u];
            if (c == 0.0) continue;
            if (W) {
                const int src = wrev ? tp.port_rev[base + t] : base + t;
                const long long wo = tp.msg_off[src];
                const int wl = (int)(tp.msg_off[src + 1] - wo);
                for (int i = threadIdx.x; i < wl; i += blockDim.x) sm.ov[i] = W[wo + i];
                __syncthreads();
            }
            // the ket operator of this term: a matrix, or `op`
            const double *om = W ? (opq ? opq + (long long)(base + t) * ld * ld : nullptr)
                                 : (opv ? opv + (long long)u * ld * ld : nullptr);
            // ket derivative: op^T . (A x_all M^T)
            const T *e1 = absorb_ports(A, vs, sm.mats, moff, ovp, sm.ov, -1, 1, sm.b0, sm.b1);
            for (long long i = threadIdx.x; i < vs.n; i += blockDim.x) {
                const int s = (int)(i / rest);
                const long long x = i - s * rest;
                double acc = 0.0;
                for (int sp = 0; sp < d; ++sp)
                    acc += (om ? om[sp * ld + s] : phys_op(op, sp, s)) * (double)e1[sp * rest + x];
                out[i] += (T)(c * acc);
            }
            __syncthreads();
            // bra derivative: (op.A) x_all M
            const T *ket = om ? ket_tensor_m(A, vs, om, ld, sm.b2) : ket_tensor(A, vs, op, sm.b2);
            const T *e2 = absorb_ports(ket, vs, sm.mats, moff, ovp, sm.ov, -1, 0, sm.b0, sm.b1);
            for (long long i = threadIdx.x; i < vs.n; i += blockDim.x)
                out[i] += (T)(c * (double)e2[i]);
            __syncthreads();
        }
    }
}

// ---- per-edge dot products: nz[e] = <GZ[q], GZ[r]>, dd[e] = <G[q], G[r]>
template <typename T>
__global__ void k_adj_edge_dots(DevTopo tp, const T *G, const T *GZ, double *nz, double *dd) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= tp.nports) return;
    const int r = tp.port_rev[q];
    if (q > r) return;
    const long long a = tp.msg_off[q], b = tp.msg_off[r];
    const int len = (int)(tp.msg_off[q + 1] - a);
    double sn = 0.0, sd = 0.0;
    for (int i = 0; i < len; ++i) {
        sn += (double)GZ[a + i] * (double)GZ[b + i];
        sd += (double)G[a + i] * (double)G[b + i];
    }
    nz[tp.port_edge[q]] = sn;
    dd[tp.port_edge[q]] = sd;
}

// ---- TFIM coefficients.  vertex v: E_v = -h zx/z;  edge e: E_e = -J nz/dd.
// cv1 = -h/z (op X), cv0 = -(-h/z)(zx/z) (op I); per slot q of edge e:
// cq1 = -J/dd (W = GZ of the reverse slot, op Z), cq0 = +J nz/dd^2 (W = G reverse, op I).
__global__ void k_adj_tfim_coefs(DevTopo tp, double J, double h, const double *z,
                                 const double *zx, const double *nz, const double *dd,
                                 double *cv1, double *cv0, double *cq1, double *cq0,
                                 double *ev, double *eq) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < tp.nv) {
        const double c = -h / z[i];
        const double ex = zx[i] / z[i];
        cv1[i] = c;
        cv0[i] = -c * ex;
        ev[i] = -h * ex;
    }
    if (i < tp.nports) {
        const int e = tp.port_edge[i];
        const double c = -J / dd[e];
        const double ezz = nz[e] / dd[e];
        cq1[i] = c;
        cq0[i] = -c * ezz;
        eq[i] = i < tp.port_rev[i] ? -J * ezz : 0.0;  // each edge once
    }
}

// ---- local-energy coefficients (operators carry their own coefficients).
// vertex v: E_v = zo/z; cv1 = 1/z (ket op O_v), cv0 = -(zo/z)/z (identity).
// edge e: E_e = nz/dd; per slot q of e: cq1 = 1/dd (W = the other side's inserted
// message, ket op = q's operator), cq0 = -nz/dd^2 (W = the other side's plain message).
__global__ void k_local_coefs(DevTopo tp, int has_v, int has_e, const double *z,
                              const double *zo, const double *nz, const double *dd,
                              double *cv1, double *cv0, double *cq1, double *cq0, double *ev,
                              double *eq) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < tp.nv) {
        const double ex = has_v ? zo[i] / z[i] : 0.0;
        cv1[i] = has_v ? 1.0 / z[i] : 0.0;
        cv0[i] = has_v ? -ex / z[i] : 0.0;
        ev[i] = ex;
    }
    if (i < tp.nports) {
        const int e = tp.port_edge[i];
        const double ee = has_e ? nz[e] / dd[e] : 0.0;
        cq1[i] = has_e ? 1.0 / dd[e] : 0.0;
        cq0[i] = has_e ? -ee / dd[e] : 0.0;
        eq[i] = i < tp.port_rev[i] ? ee : 0.0;  // each edge once
    }
}

// per-slot operators from per-edge (first, second) pairs: slot q takes ops_first[e] when it
// is the lower slot of its edge e, else ops_second[e]
__global__ void k_slot_ops(DevTopo tp, const double *ops_first, const double *ops_second,
                           int ld, double *opq) {
    const int q = blockIdx.x;
    if (q >= tp.nports) return;
    const int e = tp.port_edge[q];
    const double *src = (q < tp.port_rev[q] ? ops_first : ops_second) + (long long)e * ld * ld;
    for (int i = threadIdx.x; i < ld * ld; i += blockDim.x) opq[(long long)q * ld * ld + i] = src[i];
}

// compensated sum of a + b arrays into out[0] (one workgroup)
__global__ __launch_bounds__(BP_BLOCK) void k_adj_sum2(const double *a, int na, const double *b,
                                                        int nb, double *out) {
    __shared__ double ss[BP_BLOCK], cc[BP_BLOCK];
    double s = 0.0, c = 0.0;
    auto add = [&](double x) {
        const double t = s + x;
        if (fabs(s) >= fabs(x)) c += (s - t) + x; else c += (x - t) + s;
        s = t;
    };
    for (int i = threadIdx.x; i < na; i += blockDim.x) add(a[i]);
    for (int i = threadIdx.x; i < nb; i += blockDim.x) add(b[i]);
    ss[threadIdx.x] = s;
    cc[threadIdx.x] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        double ts = 0.0, tc = 0.0;
        for (int i = 0; i < (int)blockDim.x; ++i) {
            const double x = ss[i];
            const double t = ts + x;
            if (fabs(ts) >= fabs(x)) tc += (ts - t) + x; else tc += (x - t) + ts;
            ts = t;
            tc += cc[i];
        }
        out[0] = ts + tc;
    }
}

template <typename T>
__global__ void k_adj_axpy_copy(const T *src, T *dst, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        dst[i] = src[i];
}

// relative change max|a - b| / max|a| of the adjoint iterate
template <typename T>
__global__ __launch_bounds__(BP_BLOCK) void k_adj_maxdiff(const T *a, const T *b, long long n,
                                                           double *out) {
    __shared__ double red[BP_BLOCK / 64];
    double m = 0.0, s = 0.0;
    for (long long i = threadIdx.x; i < n; i += blockDim.x) {
        const double d = fabs((double)a[i] - (double)b[i]);
        m = (d > m || d != d) ? d : m;
        s = fmax(s, fabs((double)a[i]));
    }
    m = block_max(m, red);
    s = block_max(s, red);
    if (threadIdx.x == 0) out[0] = s > 0.0 ? m / s : m;
}

template <typename Tin, typename Tout>
__global__ void k_adj_cast(const Tin *in, Tout *out, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        out[i] = (Tout)in[i];
}

// Where a gradient goes: FP64 host memory, or a device buffer of FP64 or FP32 entries.
struct GradOut {
    double *host = nullptr;
    void *dev = nullptr;
    int dev_f64 = 0;
    bool any() const { return host != nullptr || dev != nullptr; }
};

// Writes the site cotangent `da` to its destination (g64: FP64 staging for host output).
// The caller synchronises the stream.
template <typename T>
static int adj_emit_grad(hipStream_t st, const T *da, double *g64, long long nsite,
                         const GradOut &go) {
    if (go.dev) {
        if (go.dev_f64)
            hipLaunchKernelGGL((k_adj_cast<T, double>), dim3(1024), dim3(256), 0, st, da,
                               (double *)go.dev, nsite);
        else
            hipLaunchKernelGGL((k_adj_cast<T, float>), dim3(1024), dim3(256), 0, st, da,
                               (float *)go.dev, nsite);
        return hipGetLastError() == hipSuccess ? 0 : 2;
    }
    hipLaunchKernelGGL((k_adj_cast<T, double>), dim3(1024), dim3(256), 0, st, da, g64, nsite);
    return hipMemcpyAsync(go.host, g64, sizeof(double) * nsite, hipMemcpyDeviceToHost, st) ==
                   hipSuccess ? 0 : 2;
}

// ------------------------------------------------------------------ host side

// Device buffers of the adjoint.  alloc() hands out the buffers of the previous call in the
// same order (growing one when a call needs more), so repeated gradients on one topology
// allocate nothing after the first; free_temps() only rewinds.  Every call ends with a
// stream synchronisation, so a rewound buffer is idle when the next call reuses it.
struct AdjointScratch {
    void *ws = nullptr;
    size_t ws_bytes = 0;
    std::vector<std::pair<void *, size_t>> pool;
    size_t next = 0;
    void *alloc(size_t bytes) {
        bytes = bytes ? (bytes + 255) & ~size_t(255) : 256;
        if (next < pool.size()) {
            auto &e = pool[next];
            if (e.second < bytes) {
                (void)hipFree(e.first);
                e.first = nullptr;
                e.second = 0;
                if (hipMalloc(&e.first, bytes) != hipSuccess) return nullptr;
                e.second = bytes;
            }
            ++next;
            return e.first;
        }
        void *p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return nullptr;
        pool.push_back({p, bytes});
        ++next;
        return p;
    }
    void free_temps() { next = 0; }
    void release() {
        for (auto &e : pool)
            if (e.first) (void)hipFree(e.first);
        pool.clear();
        next = 0;
        if (ws) (void)hipFree(ws);
        ws = nullptr;
        ws_bytes = 0;
    }
};

struct AdjLaunch {
    int mats_elems, ov_elems, red_elems;
    long long buf_elems;
    int use_lds;
    size_t lds;
    int grid;
    long long ws_stride;
};

static int adj_launch_cfg(AdjointScratch *s, const DevTopo &tp, long long max_site, int max_mats,
                          int maxD, size_t esz, AdjLaunch *L) {
    L->mats_elems = std::max(max_mats, 1);
    L->ov_elems = std::max(maxD * maxD, 1);
    L->red_elems = std::max(maxD * maxD, BP_BLOCK);
    L->buf_elems = std::max(max_site, 1LL);
    const size_t base = sizeof(double) * L->red_elems + esz * (L->mats_elems + 2 * L->ov_elems) + 16;
    const size_t with = base + 3 * esz * L->buf_elems;
    L->use_lds = with <= 64 * 1024;
    L->lds = L->use_lds ? with : base;
    L->grid = std::max(1, std::min(std::max(tp.nports, tp.nv), L->use_lds ? 65535 : 1024));
    L->ws_stride = 3 * L->buf_elems;
    if (!L->use_lds) {
        const size_t need = esz * (size_t)L->ws_stride * L->grid;
        if (need > s->ws_bytes) {
            if (s->ws) (void)hipFree(s->ws);
            s->ws = nullptr;
            if (hipMalloc(&s->ws, need) != hipSuccess) return 5;
            s->ws_bytes = need;
        }
    }
    return 0;
}

#define ADJ_ARGS(L, s) L.mats_elems, L.ov_elems, L.red_elems, L.buf_elems, L.use_lds, (T *)s->ws, L.ws_stride

// Everything both objectives need at the fixed point: G = open(A, A, m, p), S = sums.
template <typename T>
static int adj_forward_opens(AdjointScratch *s, hipStream_t st, const DevTopo &tp,
                             const AdjLaunch &L, const T *sites, const T *msgs, int op, T *G,
                             double *S) {
    hipLaunchKernelGGL(k_adj_open<T>, dim3(std::min(L.grid, std::max(tp.nports, 1))),
                       dim3(BP_BLOCK), L.lds, st, tp, sites, msgs, op, G, S, ADJ_ARGS(L, s));
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

template <typename T>
static int adj_grad_lognorm_t(AdjointScratch *s, hipStream_t st, const DevTopo &tp,
                              long long max_site, int max_mats, int maxD, const T *sites,
                              const T *msgs, long long nsite, const GradOut &go) {
    AdjLaunch L;
    if (adj_launch_cfg(s, tp, max_site, max_mats, maxD, sizeof(T), &L)) return 5;
    double *z = (double *)s->alloc(sizeof(double) * tp.nv);
    double *cv = (double *)s->alloc(sizeof(double) * tp.nv);
    T *abar = (T *)s->alloc(sizeof(T) * nsite);
    double *g64 = (double *)s->alloc(sizeof(double) * nsite);
    if (!z || !cv || !abar || !g64) { s->free_temps(); return 5; }
    hipLaunchKernelGGL(k_adj_vscalar<T>, dim3(std::min(L.grid, tp.nv)), dim3(BP_BLOCK), L.lds,
                       st, tp, sites, msgs, (int)OP_I, z, ADJ_ARGS(L, s));
    std::vector<double> hz(tp.nv);
    (void)hipMemcpyAsync(hz.data(), z, sizeof(double) * tp.nv, hipMemcpyDeviceToHost, st);
    (void)hipStreamSynchronize(st);
    for (double &x : hz) x = 1.0 / x;  // d log Z_v / d A_v = (dZ_v / dA_v) / Z_v
    (void)hipMemcpyAsync(cv, hz.data(), sizeof(double) * tp.nv, hipMemcpyHostToDevice, st);
    (void)hipMemsetAsync(abar, 0, sizeof(T) * nsite, st);
    hipLaunchKernelGGL(k_adj_sitevjp<T>, dim3(std::min(L.grid, tp.nv)), dim3(BP_BLOCK), L.lds, st,
                       tp, sites, msgs, (const T *)nullptr, 0, (const double *)nullptr,
                       (const double *)cv, (int)OP_I, abar, ADJ_ARGS(L, s), nullptr, nullptr, 0);
    const int rc = adj_emit_grad(st, (const T *)abar, g64, nsite, go);
    const hipError_t e = hipStreamSynchronize(st);
    s->free_temps();
    return rc ? rc : (e == hipSuccess ? 0 : 2);
}

template <typename T>
static int adj_tfim_t(AdjointScratch *s, hipStream_t st, const DevTopo &tp, int ne,
                      long long max_site, int max_mats, int maxD, const T *sites, const T *msgs,
                      long long nsite, long long nmsg, double J, double h, double *energy,
                      const GradOut &go, double adj_tol, int adj_maxiter, int *out_iters,
                      double *out_res) {
    AdjLaunch L;
    if (adj_launch_cfg(s, tp, max_site, max_mats, maxD, sizeof(T), &L)) return 5;
    const int nv = tp.nv, np = tp.nports;
    double *z = (double *)s->alloc(sizeof(double) * nv);
    double *zx = (double *)s->alloc(sizeof(double) * nv);
    T *G = (T *)s->alloc(sizeof(T) * nmsg);
    T *GZ = (T *)s->alloc(sizeof(T) * nmsg);
    double *S = (double *)s->alloc(sizeof(double) * np);
    double *SZ = (double *)s->alloc(sizeof(double) * np);
    double *nz = (double *)s->alloc(sizeof(double) * std::max(ne, 1));
    double *dd = (double *)s->alloc(sizeof(double) * std::max(ne, 1));
    double *cv1 = (double *)s->alloc(sizeof(double) * nv);
    double *cv0 = (double *)s->alloc(sizeof(double) * nv);
    double *cq1 = (double *)s->alloc(sizeof(double) * np);
    double *cq0 = (double *)s->alloc(sizeof(double) * np);
    double *ev = (double *)s->alloc(sizeof(double) * nv);
    double *eq = (double *)s->alloc(sizeof(double) * np);
    double *esum = (double *)s->alloc(sizeof(double));
    if (!z || !zx || !G || !GZ || !S || !SZ || !nz || !dd || !cv1 || !cv0 || !cq1 || !cq0 ||
        !ev || !eq || !esum) {
        s->free_temps();
        return 5;
    }
    const int gv = std::min(L.grid, nv), gp = std::min(L.grid, std::max(np, 1));
    hipLaunchKernelGGL(k_adj_vscalar<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                       (int)OP_I, z, ADJ_ARGS(L, s));
    hipLaunchKernelGGL(k_adj_vscalar<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                       (int)OP_X, zx, ADJ_ARGS(L, s));
    if (np > 0) {
        hipLaunchKernelGGL(k_adj_open<T>, dim3(gp), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                           (int)OP_I, G, S, ADJ_ARGS(L, s));
        hipLaunchKernelGGL(k_adj_open<T>, dim3(gp), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                           (int)OP_Z, GZ, SZ, ADJ_ARGS(L, s));
        hipLaunchKernelGGL(k_adj_edge_dots<T>, dim3((np + 255) / 256), dim3(256), 0, st, tp, G, GZ,
                           nz, dd);
    }
    const int nmax = std::max(nv, np);
    hipLaunchKernelGGL(k_adj_tfim_coefs, dim3((nmax + 255) / 256), dim3(256), 0, st, tp, J, h, z,
                       zx, nz, dd, cv1, cv0, cq1, cq0, ev, eq);
    hipLaunchKernelGGL(k_adj_sum2, dim3(1), dim3(BP_BLOCK), 0, st, ev, nv, eq, np, esum);
    if (hipMemcpyAsync(energy, esum, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        s->free_temps();
        return 2;
    }
    if (!go.any()) {
        s->free_temps();
        return 0;
    }
    // explicit partials dE/dm and dE/dA
    T *dm = (T *)s->alloc(sizeof(T) * nmsg);
    T *da = (T *)s->alloc(sizeof(T) * nsite);
    T *lam = (T *)s->alloc(sizeof(T) * nmsg);
    T *nxt = (T *)s->alloc(sizeof(T) * nmsg);
    T *gg = (T *)s->alloc(sizeof(T) * nmsg);
    double *res = (double *)s->alloc(sizeof(double));
    double *g64 = (double *)s->alloc(sizeof(double) * nsite);
    if (!dm || !da || !lam || !nxt || !gg || !res || !g64) {
        s->free_temps();
        return 5;
    }
    (void)hipMemsetAsync(dm, 0, sizeof(T) * nmsg, st);
    (void)hipMemsetAsync(da, 0, sizeof(T) * nsite, st);
    const T *nullW = nullptr;
    const double *nullc = nullptr;
    // vertex terms
    hipLaunchKernelGGL(k_adj_msgvjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                       nullW, 0, nullc, (const double *)cv1, (int)OP_X, dm, ADJ_ARGS(L, s), nullptr, nullptr, 0);
    hipLaunchKernelGGL(k_adj_msgvjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                       nullW, 0, nullc, (const double *)cv0, (int)OP_I, dm, ADJ_ARGS(L, s), nullptr, nullptr, 0);
    hipLaunchKernelGGL(k_adj_sitevjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                       nullW, 0, nullc, (const double *)cv1, (int)OP_X, da, ADJ_ARGS(L, s), nullptr, nullptr, 0);
    hipLaunchKernelGGL(k_adj_sitevjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                       nullW, 0, nullc, (const double *)cv0, (int)OP_I, da, ADJ_ARGS(L, s), nullptr, nullptr, 0);
    // edge terms: port p of u carries the other side's environment
    if (np > 0) {
        hipLaunchKernelGGL(k_adj_msgvjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                           (const T *)GZ, 1, (const double *)cq1, nullc, (int)OP_Z, dm,
                           ADJ_ARGS(L, s), nullptr, nullptr, 0);
        hipLaunchKernelGGL(k_adj_msgvjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                           (const T *)G, 1, (const double *)cq0, nullc, (int)OP_I, dm,
                           ADJ_ARGS(L, s), nullptr, nullptr, 0);
        hipLaunchKernelGGL(k_adj_sitevjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                           (const T *)GZ, 1, (const double *)cq1, nullc, (int)OP_Z, da,
                           ADJ_ARGS(L, s), nullptr, nullptr, 0);
        hipLaunchKernelGGL(k_adj_sitevjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                           (const T *)G, 1, (const double *)cq0, nullc, (int)OP_I, da,
                           ADJ_ARGS(L, s), nullptr, nullptr, 0);
    }
    // adjoint sweeps: lambda <- dE/dm + J_m^T lambda   (J_m^T: through the normalisation,
    // then port p of u carries gg[q], q = base + p)
    double *ones = (double *)s->alloc(sizeof(double) * std::max(np, 1));
    if (!ones) { s->free_temps(); return 5; }
    {
        std::vector<double> h1(std::max(np, 1), 1.0);
        (void)hipMemcpyAsync(ones, h1.data(), sizeof(double) * h1.size(), hipMemcpyHostToDevice,
                             st);
        (void)hipStreamSynchronize(st);
    }
    const int cblocks = (int)std::min<long long>((nmsg + 255) / 256, 4096);
    hipLaunchKernelGGL(k_adj_axpy_copy<T>, dim3(std::max(cblocks, 1)), dim3(256), 0, st,
                       (const T *)dm, lam, nmsg);
    int it = 0;
    double r = INFINITY;
    AdjFloor floor_;
    while (np > 0 && it < adj_maxiter) {
        hipLaunchKernelGGL(k_adj_gg<T>, dim3(np), dim3(64), 0, st, tp, (const T *)lam,
                           (const T *)G, (const double *)S, gg);
        hipLaunchKernelGGL(k_adj_axpy_copy<T>, dim3(std::max(cblocks, 1)), dim3(256), 0, st,
                           (const T *)dm, nxt, nmsg);
        hipLaunchKernelGGL(k_adj_msgvjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                           (const T *)gg, 0, (const double *)ones, nullc, (int)OP_I, nxt,
                           ADJ_ARGS(L, s), nullptr, nullptr, 0);
        hipLaunchKernelGGL(k_adj_maxdiff<T>, dim3(1), dim3(BP_BLOCK), 0, st, (const T *)nxt,
                           (const T *)lam, nmsg, res);
        if (hipMemcpyAsync(&r, res, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            s->free_temps();
            return 2;
        }
        std::swap(lam, nxt);
        ++it;
        if (r < adj_tol) break;
        if (std::is_same<T, float>::value && floor_.stalled(r, it)) break;
    }
    // grad = dE/dA + J_A^T lambda
    if (np > 0) {
        hipLaunchKernelGGL(k_adj_gg<T>, dim3(np), dim3(64), 0, st, tp, (const T *)lam,
                           (const T *)G, (const double *)S, gg);
        hipLaunchKernelGGL(k_adj_sitevjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                           (const T *)gg, 0, (const double *)ones, nullc, (int)OP_I, da,
                           ADJ_ARGS(L, s), nullptr, nullptr, 0);
    }
    const int rc = adj_emit_grad(st, (const T *)da, g64, nsite, go);
    const hipError_t e = hipStreamSynchronize(st);
    if (out_iters) *out_iters = it;
    if (out_res) *out_res = r;
    s->free_temps();
    return rc ? rc : (e == hipSuccess ? 0 : 2);
}


static int adjoint_grad_lognorm(AdjointScratch *s, hipStream_t st, DevTopo tp, long long max_site,
                                int max_mats, int maxD, int dtype, const void *sites,
                                const void *msgs, long long nsite, const GradOut &go) {
    if (dtype == 0)
        return adj_grad_lognorm_t<double>(s, st, tp, max_site, max_mats, maxD,
                                          (const double *)sites, (const double *)msgs, nsite, go);
    return adj_grad_lognorm_t<float>(s, st, tp, max_site, max_mats, maxD, (const float *)sites,
                                     (const float *)msgs, nsite, go);
}

// BP expectation values: <O_v> per vertex (ops_v != nullptr) or <A_e B_e> per edge
// (ops_a, ops_b != nullptr), operators as host arrays of ld x ld blocks; results (FP64)
// into out (host, nv or ne entries).  nmsg: entries of the message buffer.
template <typename T>
static int expect_t(AdjointScratch *s, hipStream_t st, const DevTopo &tp, int ne,
                    long long nmsg, long long max_site, int max_mats, int maxD, const T *sites,
                    const T *msgs, const double *ops_v, const double *ops_a,
                    const double *ops_b, int ld, double *out) {
    AdjLaunch L;
    if (adj_launch_cfg(s, tp, max_site, max_mats, maxD, sizeof(T), &L)) return 5;
    const size_t blk = sizeof(double) * (size_t)ld * ld;
    int rc = 0;
    if (ops_v) {
        const int nv = tp.nv;
        double *dops = (double *)s->alloc(blk * std::max(nv, 1));
        double *z = (double *)s->alloc(sizeof(double) * std::max(nv, 1));
        double *zo = (double *)s->alloc(sizeof(double) * std::max(nv, 1));
        double *r = (double *)s->alloc(sizeof(double) * std::max(nv, 1));
        if (!dops || !z || !zo || !r) { s->free_temps(); return 5; }
        if (hipMemcpyAsync(dops, ops_v, blk * nv, hipMemcpyHostToDevice, st) != hipSuccess) rc = 2;
        const int grid = std::min(L.grid, std::max(nv, 1));
        if (!rc) {
            hipLaunchKernelGGL(k_adj_vscalar<T>, dim3(grid), dim3(BP_BLOCK), L.lds, st, tp, sites,
                               msgs, (int)OP_I, z, ADJ_ARGS(L, s));
            hipLaunchKernelGGL(k_expect_vscalar<T>, dim3(grid), dim3(BP_BLOCK), L.lds, st, tp,
                               sites, msgs, (const double *)dops, ld, zo, ADJ_ARGS(L, s));
            hipLaunchKernelGGL(k_expect_ratio, dim3((nv + 255) / 256), dim3(256), 0, st,
                               (const double *)zo, (const double *)z, nv, r);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(out, r, sizeof(double) * nv, hipMemcpyDeviceToHost, st) !=
                    hipSuccess)
                rc = 2;
        }
    } else {
        const int np = tp.nports;
        const long long msz = std::max(nmsg, 1LL);
        double *da = (double *)s->alloc(blk * std::max(ne, 1));
        double *db = (double *)s->alloc(blk * std::max(ne, 1));
        T *G = (T *)s->alloc(sizeof(T) * msz);
        T *GO = (T *)s->alloc(sizeof(T) * msz);
        double *S = (double *)s->alloc(sizeof(double) * std::max(np, 1));
        double *nz = (double *)s->alloc(sizeof(double) * std::max(ne, 1));
        double *dd = (double *)s->alloc(sizeof(double) * std::max(ne, 1));
        double *r = (double *)s->alloc(sizeof(double) * std::max(ne, 1));
        if (!da || !db || !G || !GO || !S || !nz || !dd || !r) { s->free_temps(); return 5; }
        if (hipMemcpyAsync(da, ops_a, blk * ne, hipMemcpyHostToDevice, st) != hipSuccess ||
            hipMemcpyAsync(db, ops_b, blk * ne, hipMemcpyHostToDevice, st) != hipSuccess)
            rc = 2;
        if (!rc && np > 0) {
            const int grid = std::min(L.grid, np);
            hipLaunchKernelGGL(k_adj_open<T>, dim3(grid), dim3(BP_BLOCK), L.lds, st, tp, sites,
                               msgs, (int)OP_I, G, S, ADJ_ARGS(L, s));
            hipLaunchKernelGGL(k_expect_open<T>, dim3(grid), dim3(BP_BLOCK), L.lds, st, tp,
                               sites, msgs, (const double *)da, (const double *)db, ld, GO,
                               ADJ_ARGS(L, s));
            hipLaunchKernelGGL(k_adj_edge_dots<T>, dim3((np + 255) / 256), dim3(256), 0, st, tp,
                               (const T *)G, (const T *)GO, nz, dd);
            hipLaunchKernelGGL(k_expect_ratio, dim3((ne + 255) / 256), dim3(256), 0, st,
                               (const double *)nz, (const double *)dd, ne, r);
            if (hipGetLastError() != hipSuccess ||
                hipMemcpyAsync(out, r, sizeof(double) * ne, hipMemcpyDeviceToHost, st) !=
                    hipSuccess)
                rc = 2;
        }
    }
    const hipError_t e = hipStreamSynchronize(st);
    s->free_temps();
    return rc ? rc : (e == hipSuccess ? 0 : 2);
}

static int bp_expect(AdjointScratch *s, hipStream_t st, DevTopo tp, int ne, long long nmsg,
                     long long max_site, int max_mats, int maxD, int dtype, const void *sites,
                     const void *msgs, const double *ops_v, const double *ops_a,
                     const double *ops_b, int ld, double *out) {
    if (dtype == 0)
        return expect_t<double>(s, st, tp, ne, nmsg, max_site, max_mats, maxD,
                                (const double *)sites, (const double *)msgs, ops_v, ops_a, ops_b,
                                ld, out);
    return expect_t<float>(s, st, tp, ne, nmsg, max_site, max_mats, maxD, (const float *)sites,
                           (const float *)msgs, ops_v, ops_a, ops_b, ld, out);
}

// Local energy E = sum_v <O_v> + sum_e <A_e B_e> (host operator arrays, ld x ld blocks;
// ops_v or the edge pair may be null) and its gradient by fixed-point adjoint sweeps: the
// TFIM driver's structure with operator matrices.
template <typename T>
static int adj_local_t(AdjointScratch *s, hipStream_t st, const DevTopo &tp, int ne,
                       long long max_site, int max_mats, int maxD, const T *sites, const T *msgs,
                       long long nsite, long long nmsg, const double *ops_v,
                       const double *ops_first, const double *ops_second, int ld,
                       double *energy, const GradOut &go, double adj_tol, int adj_maxiter,
                       int *out_iters, double *out_res) {
    AdjLaunch L;
    if (adj_launch_cfg(s, tp, max_site, max_mats, maxD, sizeof(T), &L)) return 5;
    const int nv = tp.nv, np = tp.nports;
    const bool hv = ops_v != nullptr, he = ops_first && ops_second && np > 0;
    const size_t blk = sizeof(double) * (size_t)ld * ld;
    double *dopv = (double *)s->alloc(blk * std::max(nv, 1));
    double *dopa = (double *)s->alloc(blk * std::max(ne, 1));
    double *dopb = (double *)s->alloc(blk * std::max(ne, 1));
    double *dopq = (double *)s->alloc(blk * std::max(np, 1));
    double *z = (double *)s->alloc(sizeof(double) * nv);
    double *zo = (double *)s->alloc(sizeof(double) * nv);
    T *G = (T *)s->alloc(sizeof(T) * std::max(nmsg, 1LL));
    T *GO = (T *)s->alloc(sizeof(T) * std::max(nmsg, 1LL));
    double *S = (double *)s->alloc(sizeof(double) * std::max(np, 1));
    double *nz = (double *)s->alloc(sizeof(double) * std::max(ne, 1));
    double *dd = (double *)s->alloc(sizeof(double) * std::max(ne, 1));
    double *cv1 = (double *)s->alloc(sizeof(double) * nv);
    double *cv0 = (double *)s->alloc(sizeof(double) * nv);
    double *cq1 = (double *)s->alloc(sizeof(double) * std::max(np, 1));
    double *cq0 = (double *)s->alloc(sizeof(double) * std::max(np, 1));
    double *ev = (double *)s->alloc(sizeof(double) * nv);
    double *eq = (double *)s->alloc(sizeof(double) * std::max(np, 1));
    double *esum = (double *)s->alloc(sizeof(double));
    if (!dopv || !dopa || !dopb || !dopq || !z || !zo || !G || !GO || !S || !nz || !dd || !cv1 ||
        !cv0 || !cq1 || !cq0 || !ev || !eq || !esum) {
        s->free_temps();
        return 5;
    }
    if ((hv && hipMemcpyAsync(dopv, ops_v, blk * nv, hipMemcpyHostToDevice, st) != hipSuccess) ||
        (he && (hipMemcpyAsync(dopa, ops_first, blk * ne, hipMemcpyHostToDevice, st) !=
                    hipSuccess ||
                hipMemcpyAsync(dopb, ops_second, blk * ne, hipMemcpyHostToDevice, st) !=
                    hipSuccess))) {
        s->free_temps();
        return 2;
    }
    const int gv = std::min(L.grid, nv), gp = std::min(L.grid, std::max(np, 1));
    hipLaunchKernelGGL(k_adj_vscalar<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                       (int)OP_I, z, ADJ_ARGS(L, s));
    if (hv)
        hipLaunchKernelGGL(k_expect_vscalar<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites,
                           msgs, (const double *)dopv, ld, zo, ADJ_ARGS(L, s));
    if (np > 0)
        hipLaunchKernelGGL(k_adj_open<T>, dim3(gp), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                           (int)OP_I, G, S, ADJ_ARGS(L, s));
    if (he) {
        hipLaunchKernelGGL(k_expect_open<T>, dim3(gp), dim3(BP_BLOCK), L.lds, st, tp, sites,
                           msgs, (const double *)dopa, (const double *)dopb, ld, GO,
                           ADJ_ARGS(L, s));
        hipLaunchKernelGGL(k_adj_edge_dots<T>, dim3((np + 255) / 256), dim3(256), 0, st, tp,
                           (const T *)G, (const T *)GO, nz, dd);
        hipLaunchKernelGGL(k_slot_ops, dim3(np), dim3(64), 0, st, tp, (const double *)dopa,
                           (const double *)dopb, ld, dopq);
    }
    const int nmax = std::max(nv, np);
    hipLaunchKernelGGL(k_local_coefs, dim3((nmax + 255) / 256), dim3(256), 0, st, tp, hv ? 1 : 0,
                       he ? 1 : 0, (const double *)z, (const double *)zo, (const double *)nz,
                       (const double *)dd, cv1, cv0, cq1, cq0, ev, eq);
    hipLaunchKernelGGL(k_adj_sum2, dim3(1), dim3(BP_BLOCK), 0, st, ev, nv, eq, np, esum);
    if (hipMemcpyAsync(energy, esum, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        s->free_temps();
        return 2;
    }
    if (!go.any()) {
        s->free_temps();
        return 0;
    }
    T *dm = (T *)s->alloc(sizeof(T) * std::max(nmsg, 1LL));
    T *da = (T *)s->alloc(sizeof(T) * nsite);
    T *lam = (T *)s->alloc(sizeof(T) * std::max(nmsg, 1LL));
    T *nxt = (T *)s->alloc(sizeof(T) * std::max(nmsg, 1LL));
    T *gg = (T *)s->alloc(sizeof(T) * std::max(nmsg, 1LL));
    double *res = (double *)s->alloc(sizeof(double));
    double *g64 = (double *)s->alloc(sizeof(double) * nsite);
    double *ones = (double *)s->alloc(sizeof(double) * std::max(np, 1));
    if (!dm || !da || !lam || !nxt || !gg || !res || !g64 || !ones) {
        s->free_temps();
        return 5;
    }
    (void)hipMemsetAsync(dm, 0, sizeof(T) * nmsg, st);
    (void)hipMemsetAsync(da, 0, sizeof(T) * nsite, st);
    const T *nullW = nullptr;
    const double *nullc = nullptr, *nullo = nullptr;
    // explicit partials: vertex terms (ket op O_v, then the identity), edge terms (port p of
    // u carries the other side's message: inserted with u's operator, then plain)
    if (hv) {
        hipLaunchKernelGGL(k_adj_msgvjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                           nullW, 0, nullc, (const double *)cv1, (int)OP_I, dm, ADJ_ARGS(L, s),
                           (const double *)dopv, nullo, ld);
        hipLaunchKernelGGL(k_adj_msgvjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                           nullW, 0, nullc, (const double *)cv0, (int)OP_I, dm, ADJ_ARGS(L, s),
                           nullo, nullo, 0);
        hipLaunchKernelGGL(k_adj_sitevjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites,
                           msgs, nullW, 0, nullc, (const double *)cv1, (int)OP_I, da,
                           ADJ_ARGS(L, s), (const double *)dopv, nullo, ld);
        hipLaunchKernelGGL(k_adj_sitevjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites,
                           msgs, nullW, 0, nullc, (const double *)cv0, (int)OP_I, da,
                           ADJ_ARGS(L, s), nullo, nullo, 0);
    }
    if (he) {
        hipLaunchKernelGGL(k_adj_msgvjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                           (const T *)GO, 1, (const double *)cq1, nullc, (int)OP_I, dm,
                           ADJ_ARGS(L, s), nullo, (const double *)dopq, ld);
        hipLaunchKernelGGL(k_adj_msgvjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                           (const T *)G, 1, (const double *)cq0, nullc, (int)OP_I, dm,
                           ADJ_ARGS(L, s), nullo, nullo, 0);
        hipLaunchKernelGGL(k_adj_sitevjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites,
                           msgs, (const T *)GO, 1, (const double *)cq1, nullc, (int)OP_I, da,
                           ADJ_ARGS(L, s), nullo, (const double *)dopq, ld);
        hipLaunchKernelGGL(k_adj_sitevjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites,
                           msgs, (const T *)G, 1, (const double *)cq0, nullc, (int)OP_I, da,
                           ADJ_ARGS(L, s), nullo, nullo, 0);
    }
    {
        std::vector<double> h1(std::max(np, 1), 1.0);
        (void)hipMemcpyAsync(ones, h1.data(), sizeof(double) * h1.size(), hipMemcpyHostToDevice,
                             st);
        (void)hipStreamSynchronize(st);
    }
    // adjoint sweeps lambda <- dE/dm + J_m^T lambda, then grad = dE/dA + J_A^T lambda
    const int cblocks = (int)std::min<long long>((nmsg + 255) / 256, 4096);
    hipLaunchKernelGGL(k_adj_axpy_copy<T>, dim3(std::max(cblocks, 1)), dim3(256), 0, st,
                       (const T *)dm, lam, nmsg);
    int it = 0;
    double r = INFINITY;
    AdjFloor floor_;
    while (np > 0 && it < adj_maxiter) {
        hipLaunchKernelGGL(k_adj_gg<T>, dim3(np), dim3(64), 0, st, tp, (const T *)lam,
                           (const T *)G, (const double *)S, gg);
        hipLaunchKernelGGL(k_adj_axpy_copy<T>, dim3(std::max(cblocks, 1)), dim3(256), 0, st,
                           (const T *)dm, nxt, nmsg);
        hipLaunchKernelGGL(k_adj_msgvjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites, msgs,
                           (const T *)gg, 0, (const double *)ones, nullc, (int)OP_I, nxt,
                           ADJ_ARGS(L, s), nullo, nullo, 0);
        hipLaunchKernelGGL(k_adj_maxdiff<T>, dim3(1), dim3(BP_BLOCK), 0, st, (const T *)nxt,
                           (const T *)lam, nmsg, res);
        if (hipMemcpyAsync(&r, res, sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            s->free_temps();
            return 2;
        }
        std::swap(lam, nxt);
        ++it;
        if (r < adj_tol) break;
        if (std::is_same<T, float>::value && floor_.stalled(r, it)) break;
    }
    if (np > 0) {
        hipLaunchKernelGGL(k_adj_gg<T>, dim3(np), dim3(64), 0, st, tp, (const T *)lam,
                           (const T *)G, (const double *)S, gg);
        hipLaunchKernelGGL(k_adj_sitevjp<T>, dim3(gv), dim3(BP_BLOCK), L.lds, st, tp, sites,
                           msgs, (const T *)gg, 0, (const double *)ones, nullc, (int)OP_I, da,
                           ADJ_ARGS(L, s), nullo, nullo, 0);
    }
    const int rc = adj_emit_grad(st, (const T *)da, g64, nsite, go);
    const hipError_t e = hipStreamSynchronize(st);
    if (out_iters) *out_iters = it;
    if (out_res) *out_res = r;
    s->free_temps();
    return rc ? rc : (e == hipSuccess ? 0 : 2);
}

static int adjoint_local(AdjointScratch *s, hipStream_t st, DevTopo tp, int ne,
                         long long max_site, int max_mats, int maxD, int dtype, const void *sites,
                         const void *msgs, long long nsite, long long nmsg, const double *ops_v,
                         const double *ops_first, const double *ops_second, int ld,
                         double *energy, const GradOut &go, double adj_tol, int adj_maxiter,
                         int *out_iters, double *out_res) {
    if (dtype == 0)
        return adj_local_t<double>(s, st, tp, ne, max_site, max_mats, maxD, (const double *)sites,
                                   (const double *)msgs, nsite, nmsg, ops_v, ops_first,
                                   ops_second, ld, energy, go, adj_tol, adj_maxiter, out_iters,
                                   out_res);
    return adj_local_t<float>(s, st, tp, ne, max_site, max_mats, maxD, (const float *)sites,
                              (const float *)msgs, nsite, nmsg, ops_v, ops_first, ops_second, ld,
                              energy, go, adj_tol, adj_maxiter, out_iters, out_res);
}

#undef ADJ_ARGS

static int adjoint_tfim(AdjointScratch *s, hipStream_t st, DevTopo tp, int ne, long long max_site,
                        int max_mats, int maxD, int dtype, const void *sites, const void *msgs,
                        long long nsite, long long nmsg, double J, double h, double *energy,
                        const GradOut &go, double adj_tol, int adj_maxiter, int *out_iters,
                        double *out_res) {
    if (dtype == 0)
        return adj_tfim_t<double>(s, st, tp, ne, max_site, max_mats, maxD, (const double *)sites,
                                  (const double *)msgs, nsite, nmsg, J, h, energy, go, adj_tol,
                                  adj_maxiter, out_iters, out_res);
    return adj_tfim_t<float>(s, st, tp, ne, max_site, max_mats, maxD, (const float *)sites,
                             (const float *)msgs, nsite, nmsg, J, h, energy, go, adj_tol,
                             adj_maxiter, out_iters, out_res);
}
