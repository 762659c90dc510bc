// complex.h -- complex (ComplexF64) BP on the generic kernel's building blocks.
//
// The reference runs BP for complex element types (test/test_beliefpropagation.jl:154 with
// T = ComplexF64; vertex_scalar of ComplexF32 tensors, :84-99).  For a norm network the
// factor of vertex u is A_u (ket) times conj(A_u) (bra, normnetwork.jl:49-54), so
//   m_{u->v}[a,a'] = sum A_u[s,..a..,X] conj(A_u[s,..a'..,X']) prod_{w != v} m_{w->u}[x_w,x'_w]
// with messages [ket][bra]; a plain network contracts A_u with its incoming vectors and no
// conjugate.  Normalisation divides by the complex sum when it is nonzero
// (beliefpropagation.jl:248-253); iterate_diff is 1 - |<m1^, m2^>|^2 with the first argument
// conjugated (beliefpropagation.jl:261-267); scalars are complex; the log-norm takes the
// principal logarithm of every term (messagecache.jl:185-201).
//
// Storage: interleaved (re, im) doubles, the same element offsets as the real contexts.
// One workgroup per directed message, as k_update_generic; products and sums in FP64.
#pragma once

#include "generic.h"

struct cplx {
    double re, im;
    cplx() = default;
    __host__ __device__ constexpr cplx(double r, double i = 0.0) : re(r), im(i) {}
    __host__ __device__ cplx &operator+=(const cplx &o) {
        re += o.re;
        im += o.im;
        return *this;
    }
};

__host__ __device__ __forceinline__ cplx operator+(cplx a, cplx b) { return {a.re + b.re, a.im + b.im}; }
__host__ __device__ __forceinline__ cplx operator-(cplx a, cplx b) { return {a.re - b.re, a.im - b.im}; }
__host__ __device__ __forceinline__ cplx operator*(cplx a, cplx b) {
    return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
__host__ __device__ __forceinline__ cplx cconj(cplx a) { return {a.re, -a.im}; }
__host__ __device__ __forceinline__ double cabs2(cplx a) { return a.re * a.re + a.im * a.im; }
// a / b (b != 0), scaled so that |b|^2 cannot overflow or underflow
__device__ __forceinline__ cplx cdiv(cplx a, cplx b) {
    const double s = fmax(fabs(b.re), fabs(b.im));
    const cplx bs{b.re / s, b.im / s};
    const double d = bs.re * bs.re + bs.im * bs.im;
    const cplx n = a * cconj(bs);
    return {n.re / (d * s), n.im / (d * s)};
}

// out[a, a'] = sum_{pre,post} cur[pre, a, post] * conj(A[pre, a', post]) on axis ax (the
// open contraction of a norm-network message: the bra layer is the conjugated site).
// red: LDS scratch of 2 * max(n*n, blockDim) doubles (interleaved partial sums); out: n*n.
__device__ void block_open_contract_c(const cplx *cur, const cplx *A, const int *dims, int r,
                                      int ax, double *red, cplx *out) {
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
        cplx s(0.0);
        for (long long c = g; c < ncol; c += G) {
            const long long p = c / post, t = c - p * post;
            s += cur[(p * n + a) * post + t] * cconj(A[(p * n + b) * post + t]);
        }
        red[2 * item] = s.re;
        red[2 * item + 1] = s.im;
    }
    __syncthreads();
    for (int pr = threadIdx.x; pr < pairs; pr += blockDim.x) {
        cplx s(0.0);
        for (int g = 0; g < G; ++g) s += cplx(red[2 * (g * pairs + pr)], red[2 * (g * pairs + pr) + 1]);
        out[pr] = s;
    }
    __syncthreads();
}

// One directed message per workgroup iteration: update, normalise, residual (the complex
// counterpart of k_update_generic; same work items, control and LDS carve-up).
template <int BS = BP_BLOCK>
__global__ __launch_bounds__(BS) void k_update_cplx(
    DevTopo tp, const cplx *__restrict__ sites, const cplx *mold, cplx *mnew, NetCtl ctl,
    double *__restrict__ res_fid, double *__restrict__ res_abs, int item_begin, int item_end,
    const int *__restrict__ slot_list, int mats_elems, int out_elems, int red_elems,
    long long buf_elems, int use_lds, cplx *ws, long long ws_stride, int flags) {
    extern __shared__ __align__(16) char smem[];
    GenSmem<cplx> sm = carve<cplx>(smem, mats_elems, out_elems, red_elems, buf_elems, use_lds,
                                   ws, ws_stride);
    const bool normalize = !(flags & GEN_FLAG_NO_NORMALIZE);
    for (int it = item_begin + blockIdx.x; it < item_end; it += gridDim.x) {
        const int q = slot_list ? slot_list[it] : it;
        const int u = tp.owner[q];
        const int p = q - tp.port_off[u];
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
        const cplx *A = sites + tp.site_off[u];
        const cplx *cur = A;
        cplx *bufs[2] = {sm.buf0, sm.buf1};
        int which = 0;
        for (int j = 0; j < vs.k; ++j) {
            if (j == p) continue;
            const int nb = tp.plain ? 1 : vs.dims[1 + j];
            block_mode_product(cur, bufs[which], vs.dims, vs.r, 1 + j, sm.mats + moff[j], nb);
            if (tp.plain) vs.dims[1 + j] = 1;
            __syncthreads();
            cur = bufs[which];
            which ^= 1;
        }
        if (tp.plain) {
            for (int i = threadIdx.x; i < len; i += blockDim.x) sm.out[i] = cur[i];
            __syncthreads();
        } else {
            block_open_contract_c(cur, A, vs.dims, vs.r, 1 + p, sm.red, sm.out);
        }
        // complex sum normalisation (beliefpropagation.jl:248-253) and residuals
        double lr = 0.0, li = 0.0;
        for (int i = threadIdx.x; i < len; i += blockDim.x) {
            lr += sm.out[i].re;
            li += sm.out[i].im;
        }
        const cplx S(block_sum(lr, sm.red), block_sum(li, sm.red));
        const bool scale = normalize && (S.re != 0.0 || S.im != 0.0);
        const cplx *old = mold + mo;
        double dr = 0.0, di = 0.0, dnn = 0.0, doo = 0.0, dmx = 0.0;
        for (int i = threadIdx.x; i < len; i += blockDim.x) {
            cplx x = sm.out[i];
            if (scale) x = cdiv(x, S);
            const cplx xo = old[i];  // read before an in-place write
            mnew[mo + i] = x;
            const cplx ov = cconj(xo) * x;  // <old, new>, the first argument conjugated
            dr += ov.re;
            di += ov.im;
            dnn += cabs2(x);
            doo += cabs2(xo);
            const double d = sqrt(cabs2(x - xo));
            dmx = (d > dmx || d != d) ? d : dmx;
        }
        dr = block_sum(dr, sm.red);
        di = block_sum(di, sm.red);
        dnn = block_sum(dnn, sm.red);
        doo = block_sum(doo, sm.red);
        dmx = block_max(dmx, sm.red);
        if (threadIdx.x == 0) {
            res_fid[q] = 1.0 - (dr * dr + di * di) / (dnn * doo);
            res_abs[q] = dmx;
        }
        __syncthreads();
    }
}

// Z_v = full contraction of vertex v with all incoming messages: A absorbs every message on
// the ket side, then meets conj(A) (norm networks); plain networks end with a scalar.
__global__ __launch_bounds__(BP_BLOCK) void k_vertex_scalar_cplx(
    DevTopo tp, const cplx *__restrict__ sites, const cplx *__restrict__ msgs,
    double *__restrict__ zv, int mats_elems, int out_elems, int red_elems, long long buf_elems,
    int use_lds, cplx *ws, long long ws_stride) {
    extern __shared__ __align__(16) char smem[];
    GenSmem<cplx> sm = carve<cplx>(smem, mats_elems, out_elems, red_elems, buf_elems, use_lds,
                                   ws, ws_stride);
    for (int u = blockIdx.x; u < tp.nv; u += gridDim.x) {
        VertexShape vs = vertex_shape(tp, u);
        int moff[BP_MAXR];
        load_incoming(tp, u, vs, msgs, sm.mats, moff, -1);
        __syncthreads();
        const cplx *A = sites + tp.site_off[u];
        const cplx *cur = A;
        cplx *bufs[2] = {sm.buf0, sm.buf1};
        int which = 0;
        for (int j = 0; j < vs.k; ++j) {
            const int nb = tp.plain ? 1 : vs.dims[1 + j];
            block_mode_product(cur, bufs[which], vs.dims, vs.r, 1 + j, sm.mats + moff[j], nb);
            if (tp.plain) vs.dims[1 + j] = 1;
            __syncthreads();
            cur = bufs[which];
            which ^= 1;
        }
        cplx acc(0.0);
        if (tp.plain) {
            if (threadIdx.x == 0) acc = cur[0];
        } else {
            for (long long i = threadIdx.x; i < vs.n; i += blockDim.x) acc += cur[i] * cconj(A[i]);
        }
        const double re = block_sum(acc.re, sm.red);
        const double im = block_sum(acc.im, sm.red);
        if (threadIdx.x == 0) {
            zv[2 * u] = re;
            zv[2 * u + 1] = im;
        }
        __syncthreads();
    }
}

// Z_e = sum_i m_uv[i] m_vu[i] (both shared names contracted, no conjugate,
// messagecache.jl:153-157), one thread per port pair; interleaved complex output.
__global__ void k_edge_scalars_cplx(DevTopo tp, const cplx *__restrict__ msgs,
                                    double *__restrict__ ze) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= tp.nports) return;
    const int r = tp.port_rev[q];
    if (q > r) return;
    const long long a = tp.msg_off[q], b = tp.msg_off[r];
    const int len = (int)(tp.msg_off[q + 1] - a);
    cplx s(0.0);
    for (int i = 0; i < len; ++i) s += msgs[a + i] * msgs[b + i];
    const int e = tp.port_edge[q];
    ze[2 * e] = s.re;
    ze[2 * e + 1] = s.im;
}

// Per-slot distance between two complex message buffers (one wave per slot):
// 1 - |<a^, b^>|^2 with a conjugated, and max |a - b|.
__global__ void k_diff_slots_cplx(DevTopo tp, const cplx *a, const cplx *b, double *res_fid,
                                  double *res_abs) {
    const int q = blockIdx.x;
    if (q >= tp.nports) return;
    const long long mo = tp.msg_off[q];
    const int len = (int)(tp.msg_off[q + 1] - mo);
    double dr = 0.0, di = 0.0, aa = 0.0, bb = 0.0, mx = 0.0;
    for (int i = threadIdx.x; i < len; i += blockDim.x) {
        const cplx x = a[mo + i], y = b[mo + i];
        const cplx ov = cconj(x) * y;
        dr += ov.re;
        di += ov.im;
        aa += cabs2(x);
        bb += cabs2(y);
        const double d = sqrt(cabs2(x - y));
        mx = (d > mx || d != d) ? d : mx;
    }
    dr = wave_sum(dr);
    di = wave_sum(di);
    aa = wave_sum(aa);
    bb = wave_sum(bb);
    mx = wave_max(mx);
    if (threadIdx.x == 0) {
        res_fid[q] = 1.0 - (dr * dr + di * di) / (aa * bb);
        res_abs[q] = mx;
    }
}

// BP log-norm per network with complex terms (one workgroup per network):
// sum log Z_v - sum log Z_e with the principal logarithm log|z| + i arg z of every term,
// real parts Neumaier-compensated; any Z_e = 0 gives -Inf, a Z_v = 0 gives a real part of
// -Inf (messagecache.jl:185-201).
__global__ __launch_bounds__(BP_BLOCK) void k_lognorm_nets_cplx(DevTopo tp, const int *vbeg,
                                                                 const int *pbeg,
                                                                 const double *zv,
                                                                 const double *ze,
                                                                 double *out_re,
                                                                 double *out_im) {
    __shared__ double ss[BP_BLOCK], cc[BP_BLOCK], ii[BP_BLOCK];
    __shared__ int flag[2];
    const int n = blockIdx.x;
    if (threadIdx.x < 2) flag[threadIdx.x] = 0;
    __syncthreads();
    NeuSum acc, ang;
    int zv0 = 0, ze0 = 0;
    for (int u = vbeg[n] + threadIdx.x; u < vbeg[n + 1]; u += blockDim.x) {
        const double x = zv[2 * u], y = zv[2 * u + 1];
        ang.add(atan2(y, x));
        if (x == 0.0 && y == 0.0) { zv0 = 1; continue; }
        acc.add(log(hypot(x, y)));
    }
    for (int q = pbeg[n] + threadIdx.x; q < pbeg[n + 1]; q += blockDim.x) {
        if (q > tp.port_rev[q]) continue;
        const int e = tp.port_edge[q];
        const double x = ze[2 * e], y = ze[2 * e + 1];
        if (x == 0.0 && y == 0.0) { ze0 = 1; continue; }
        ang.add(-atan2(y, x));
        acc.add(-log(hypot(x, y)));
    }
    ss[threadIdx.x] = acc.s;
    cc[threadIdx.x] = acc.c;
    ii[threadIdx.x] = ang.s + ang.c;
    atomicOr(&flag[0], zv0);
    atomicOr(&flag[1], ze0);
    __syncthreads();
    if (threadIdx.x == 0) {
        NeuSum t, a;
        double c = 0.0;
        for (int i = 0; i < (int)blockDim.x; ++i) {
            t.add(ss[i]);
            c += cc[i];
            a.add(ii[i]);
        }
        double re = t.s + (t.c + c);
        double im = a.s + a.c;
        if (flag[1]) { re = -INFINITY; im = 0.0; }
        else if (flag[0]) re = -INFINITY;
        out_re[n] = re;
        out_im[n] = im;
    }
}
