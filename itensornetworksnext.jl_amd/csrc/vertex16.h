// vertex16.h -- BF16 MFMA sweep for bond dimensions 9..16 (padded to 16) on gfx950.
//
// One workgroup of NW waves (4, 8 or 16; 16 by default, the autotuner's choice on the
// 24x24 D=16 lattice) per (vertex, physical slice s).  The slice A[s] (four ports of
// dimension 16, 16^4 entries) lives in LDS as BF16 (128 KiB), rows of 16 entries along the
// last port, row r stored at row r ^ ((r>>3 ^ r>>6 ^ r>>9) & 7) so that accesses strided
// along any port spread over the banks.
//
// Mode product on port j ("T x_j m_j", in place) runs on v_mfma_f32_16x16x16_bf16: A operand
// = m_j transposed (lane l holds m_j[4(l/16)+t][l%16]), B operand = a group of 16 fibres
// along j (lane l: entries j = 4(l/16)+t of fibre l%16), and the result D[b][fibre] lands
// in the lane that read entries j = 4(l/16)+t of that fibre: every lane reads and writes
// the same four entries, so the product is in place.  Fibres run along a row port R (the
// last port, or the third when j is the last), the other two ports index the 256 groups.
//
// Open contraction out_j[a][a'] = sum_K G[a, K] A[a', K] (K: the other three ports) runs on
// v_mfma_f32_16x16x32_bf16 with G from LDS and the bra A from the padded site array (L2 /
// Infinity Cache; a second copy with the last two ports swapped serves the last port's
// open), prefetched into registers before the mode products: 128 MFMAs per slice, split
// over the four waves, reduced through LDS.
// The per-slice 16 x 16 results go to a scratch buffer; k_v16_finish sums the slices and
// does what the D <= 8 kernels do at their end (normalisation, residuals, write-back of the
// D x D block).
//
// Per slice: E23 (ports 2, 3 absorbed) serves outgoing ports 0 and 1, E01 serves 2 and 3;
// while the first port of a pair absorbs its third message in LDS the partial waits in
// registers (8 mode products, two slice loads; with 4 waves per slice, 12 and four).  Ports of a padded (dummy) slot are
// skipped: their message is the identity.  Arithmetic: BF16 operands (site entries,
// messages scaled by 1 / max|m|, and every partial product), FP32 accumulation, as the
// D <= 8 BF16 kernel.
#pragma once

#include "vertex.h"
#include "vertex12.h"

typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef short v8s __attribute__((ext_vector_type(8)));

static constexpr int V16_ELEMS = 65536;                            // 16^4 per slice
static constexpr size_t v16_lds(int nw) { return V16_ELEMS * 2 + (size_t)nw * 256 * 4; }

__device__ __forceinline__ int v16_row(int r) { return r ^ (((r >> 3) ^ (r >> 6) ^ (r >> 9)) & 7); }

// byte offset in LDS of entry (p0, p1, p2, p3)
__device__ __forceinline__ int v16_off(int p0, int p1, int p2, int p3) {
    return v16_row((p0 * 16 + p1) * 16 + p2) * 32 + p3 * 2;
}

// v16_row is linear over xor, and a row index (p0 << 8 | p1 << 4 | p2) splits into fields
// that depend on the lane and fields that depend only on the loop iteration, so
// v16_off = v16_off(lane fields) ^ v16_off(iteration fields): one xor per access.

__device__ __forceinline__ unsigned short v16_ld(const unsigned char *S, int off) {
    return *reinterpret_cast<const unsigned short *>(S + off);
}

__device__ __forceinline__ void v16_st(unsigned char *S, int off, unsigned short v) {
    *reinterpret_cast<unsigned short *>(S + off) = v;
}

// Ports of the fibre grouping of a mode product on port J: fibres run along R, groups are
// indexed by (P, Q), P < Q.  Compile-time, so every entry offset is a few shifts and xors.
template <int J> struct V16Ports {
    static constexpr int R = J == 3 ? 2 : 3;
    static constexpr int P = J == 0 ? 1 : 0;
    static constexpr int Q = J == 0 ? 2 : (J == 1 ? 2 : 1);
};

// Mode product on port J of the slice in LDS, in place.  Am: the message as A operand.
template <int J, int NW>
__device__ __forceinline__ void v16_mode(unsigned char *S, v4s Am) {
    using PT = V16Ports<J>;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = lane & 15, kg = lane >> 4;
    const v4f zero = {0.f, 0.f, 0.f, 0.f};
    int loff[4];  // lane fields: port J = 4 kg + t, port R = n
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        int p[4] = {0, 0, 0, 0};
        p[J] = 4 * kg + t;
        p[PT::R] = n;
        loff[t] = v16_off(p[0], p[1], p[2], p[3]);
    }
#pragma unroll 4
    for (int i = 0; i < 256 / NW; ++i) {  // this wave's tiles g = w, w + NW, ...
        const int g = w + NW * i;
        int q[4] = {0, 0, 0, 0};  // iteration fields: ports P, Q (never the last port)
        q[PT::P] = g >> 4;
        q[PT::Q] = g & 15;
        const int goff = v16_off(q[0], q[1], q[2], 0);
        int off[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) off[t] = loff[t] ^ goff;
        v4s b;
        if constexpr (J == 3) {  // the four entries are adjacent in one row
            const unsigned long long x = *reinterpret_cast<const unsigned long long *>(S + off[0]);
            b = __builtin_bit_cast(v4s, x);
        } else {
#pragma unroll
            for (int t = 0; t < 4; ++t) b[t] = (short)v16_ld(S, off[t]);
        }
        const v4f c = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(Am, b, zero, 0, 0, 0);
        const unsigned lo = bf16x2(c[0], c[1]), hi = bf16x2(c[2], c[3]);
        if constexpr (J == 3) {
            const unsigned long long x = (unsigned long long)lo | ((unsigned long long)hi << 32);
            *reinterpret_cast<unsigned long long *>(S + off[0]) = x;
        } else {
            v16_st(S, off[0], (unsigned short)(lo & 0xffffu));
            v16_st(S, off[1], (unsigned short)(lo >> 16));
            v16_st(S, off[2], (unsigned short)(hi & 0xffffu));
            v16_st(S, off[3], (unsigned short)(hi >> 16));
        }
    }
}

// Bra operands of this wave's 32 MFMA steps of the open contraction on port J, loaded
// ahead of the mode products (their addresses do not depend on the slice in LDS).  K runs
// over the three other ports, 32 values per step: for J < 3, (P, Q) pairs {2t, 2t+1} x all
// 16 values of p3, read from the natural layout `bra`; for J = 3, (p0, p1) pairs x all 16
// values of p2, read from the copy `bra3` whose last two ports are swapped, so that each
// lane's 8 K-values are 16 contiguous bytes in both cases.
template <int J, int NW>
__device__ __forceinline__ void v16_bra_prefetch(const unsigned short *bra,
                                                 const unsigned short *bra3,
                                                 uint4 (&pre)[128 / NW]) {
    using PT = V16Ports<J>;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = lane & 15, kg = lane >> 4;
#pragma unroll
    for (int i = 0; i < 128 / NW; ++i) {
        const int t = w + NW * i;
        const int hi = 2 * t + (kg >> 1), lo8 = 8 * (kg & 1);
        int e;
        if constexpr (J == 3) {
            e = (hi * 16 + n) * 16 + lo8;  // [p0 p1][p3][p2]
        } else {
            int p[4];
            p[J] = n;
            p[PT::P] = hi >> 4;
            p[PT::Q] = hi & 15;
            p[3] = lo8;
            e = ((p[0] * 16 + p[1]) * 16 + p[2]) * 16 + lo8;
        }
        pre[i] = *reinterpret_cast<const uint4 *>((J == 3 ? bra3 : bra) + e);
    }
}

// This wave's share of out_J[a][a'] = sum_K G[a, K] A[a', K] with the bra operands of
// v16_bra_prefetch; lane l returns entries (4(l/16)+i, l%16).
template <int J, int NW>
__device__ __forceinline__ v4f v16_open(const unsigned char *S, const uint4 (&pre)[128 / NW]) {
    using PT = V16Ports<J>;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = lane & 15, kg = lane >> 4;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
    // step t = w + NW i covers hi = 2 t + (kg >> 1) = 2 NW i | 2 w | (kg >> 1) (disjoint
    // bits): lane fields kg, n; iteration fields i, w
    int loff;
    if constexpr (J == 3) {
        loff = v16_off(0, kg >> 1, 8 * (kg & 1), n);
    } else {
        int p[4] = {0, 0, 0, 0};
        p[J] = n;
        p[PT::Q] = kg >> 1;
        p[3] = 8 * (kg & 1);
        loff = v16_off(p[0], p[1], p[2], p[3]);
    }
#pragma unroll
    for (int i = 0; i < 128 / NW; ++i) {
        const int it = 2 * NW * i + 2 * w;  // iteration part of hi
        uint4 ga;
        if constexpr (J == 3) {
            v8s a8;
            const int ioff = v16_off(it >> 4, it & 15, 0, 0);
#pragma unroll
            for (int u = 0; u < 8; ++u)
                a8[u] = (short)v16_ld(S, (loff ^ ioff) ^ v16_off(0, 0, u, 0));
            ga = __builtin_bit_cast(uint4, a8);
        } else {
            int q[4] = {0, 0, 0, 0};
            q[PT::P] = it >> 4;
            q[PT::Q] = it & 15;
            ga = *reinterpret_cast<const uint4 *>(S + (loff ^ v16_off(q[0], q[1], q[2], 0)));
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(v8bf, ga),
                                                      __builtin_bit_cast(v8bf, pre[i]), acc, 0,
                                                      0, 0);
    }
    return acc;
}

// Copy the slice's padded BF16 entries into LDS (16-byte chunks, half a row each), eight
// loads in flight per thread.
template <int NW>
__device__ __forceinline__ void v16_load_slice(unsigned char *S, const unsigned short *src) {
    const uint4 *g = reinterpret_cast<const uint4 *>(src);
#pragma unroll 8
    for (int i = 0; i < V16_ELEMS / 8 / (64 * NW); ++i) {
        const int c = threadIdx.x + 64 * NW * i;
        *reinterpret_cast<uint4 *>(S + v16_row(c >> 1) * 32 + (c & 1) * 16) = g[c];
    }
}

template <int J, int NW>
__device__ __forceinline__ void v16_mode_if(unsigned char *S, const int (&ins)[4],
                                            const v4s (&Am)[4]) {
    if (ins[J] < 0) return;  // identity message on a padded port (uniform)
    v16_mode<J, NW>(S, Am[J]);
    __syncthreads();
}

// One outgoing port O of the slice: reload, absorb ports Q0, Q1, Q2, open, reduce the four
// waves' shares, store the unnormalised 16 x 16 result.
template <int NW, int O, int Q0, int Q1, int Q2>
__device__ __forceinline__ void v16_out(unsigned char *S, float *red, const unsigned short *src,
                                        const unsigned short *src_t, float *dst,
                                        const int (&ins)[4], const int (&outs)[4],
                                        const v4s (&Am)[4]) {
    if (outs[O] < 0) return;  // uniform over the workgroup
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = lane & 15, kg = lane >> 4;
    uint4 pre[128 / NW];
    v16_bra_prefetch<O, NW>(src, src_t, pre);
    __syncthreads();  // the previous open has finished reading the slice
    v16_load_slice<NW>(S, src);
    __syncthreads();
    v16_mode_if<Q0, NW>(S, ins, Am);
    v16_mode_if<Q1, NW>(S, ins, Am);
    v16_mode_if<Q2, NW>(S, ins, Am);
    const v4f acc = v16_open<O, NW>(S, pre);
#pragma unroll
    for (int i = 0; i < 4; ++i) red[w * 256 + (4 * kg + i) * 16 + n] = acc[i];
    __syncthreads();
    for (int e = threadIdx.x; e < 256; e += blockDim.x) {
        float x = 0.f;
#pragma unroll
        for (int v = 0; v < NW; ++v) x += red[v * 256 + e];
        dst[O * 256 + e] = x;
    }
}

// Outgoing ports O1 and O2 from one shared partial product: reload the slice, absorb X0, X1
// (E23 for O = 0, 1; E01 for O = 2, 3), keep the partial in registers (the raw LDS bytes,
// 128 / NW 16-byte chunks per thread), absorb Y1 and open O1, restore, absorb Y2 and open
// O2: 8 mode products and two slice loads per slice instead of 12 and four.  With 4 waves
// the partial would take 128 registers per thread, so that shape keeps the per-port path.
template <int NW, int O1, int O2, int X0, int X1, int Y1, int Y2>
__device__ __forceinline__ void v16_pair(unsigned char *S, float *red, const unsigned short *src,
                                         const unsigned short *src_t, float *dst,
                                         const int (&ins)[4], const int (&outs)[4],
                                         const v4s (&Am)[4]) {
    if constexpr (NW < 8) {
        v16_out<NW, O1, X0, X1, Y1>(S, red, src, src_t, dst, ins, outs, Am);
        v16_out<NW, O2, X0, X1, Y2>(S, red, src, src_t, dst, ins, outs, Am);
    } else {
        const bool first = outs[O1] >= 0, second = outs[O2] >= 0;  // uniform
        if (!first && !second) return;
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        const int n = lane & 15, kg = lane >> 4;
        constexpr int NP = V16_ELEMS * 2 / 16 / (64 * NW);
        uint4 pre[128 / NW];
        if (first) v16_bra_prefetch<O1, NW>(src, src_t, pre);
        else v16_bra_prefetch<O2, NW>(src, src_t, pre);
        __syncthreads();  // earlier opens have finished with the slice and the partials
        v16_load_slice<NW>(S, src);
        __syncthreads();
        v16_mode_if<X0, NW>(S, ins, Am);
        v16_mode_if<X1, NW>(S, ins, Am);
        unsigned pk[4 * NP];  // scalars, so the partial stays in registers
        if (first && second) {
#pragma unroll
            for (int i = 0; i < NP; ++i) {
                const uint4 v = reinterpret_cast<const uint4 *>(S)[threadIdx.x + 64 * NW * i];
                pk[4 * i] = v.x;
                pk[4 * i + 1] = v.y;
                pk[4 * i + 2] = v.z;
                pk[4 * i + 3] = v.w;
            }
            __syncthreads();
        }
        if (first) {
            v16_mode_if<Y1, NW>(S, ins, Am);
            const v4f acc = v16_open<O1, NW>(S, pre);
#pragma unroll
            for (int i = 0; i < 4; ++i) red[w * 256 + (4 * kg + i) * 16 + n] = acc[i];
            __syncthreads();
            for (int e = threadIdx.x; e < 256; e += blockDim.x) {
                float x = 0.f;
#pragma unroll
                for (int v = 0; v < NW; ++v) x += red[v * 256 + e];
                dst[O1 * 256 + e] = x;
            }
        }
        if (second) {
            if (first) {
                v16_bra_prefetch<O2, NW>(src, src_t, pre);
#pragma unroll
                for (int i = 0; i < NP; ++i)
                    reinterpret_cast<uint4 *>(S)[threadIdx.x + 64 * NW * i] =
                        make_uint4(pk[4 * i], pk[4 * i + 1], pk[4 * i + 2], pk[4 * i + 3]);
                __syncthreads();
            }
            v16_mode_if<Y2, NW>(S, ins, Am);
            const v4f acc = v16_open<O2, NW>(S, pre);
#pragma unroll
            for (int i = 0; i < 4; ++i) red[w * 256 + (4 * kg + i) * 16 + n] = acc[i];
            __syncthreads();
            for (int e = threadIdx.x; e < 256; e += blockDim.x) {
                float x = 0.f;
#pragma unroll
                for (int v = 0; v < NW; ++v) x += red[v * 256 + e];
                dst[O2 * 256 + e] = x;
            }
        }
    }
}

// grid: items x d (one workgroup per slice), 64 NW threads.  part[((item * d + s) * 4 + j) * 256
// + a * 16 + a'] receives slice s's unnormalised out_j.
template <int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(NW / 4, NW / 4)))
void k_vertex_d16(const unsigned short *__restrict__ vsites16,
                  const unsigned short *__restrict__ vsites16t, const float *__restrict__ mold,
                  const int4 *__restrict__ vin, const int4 *__restrict__ vout,
                  const int4 *__restrict__ vdim, const int *__restrict__ vnet,
                  const long long *__restrict__ msg_off, NetCtl ctl, int d,
                  float *__restrict__ part) {
    extern __shared__ __align__(16) unsigned char v16sm[];
    unsigned char *S = v16sm;
    float *red = reinterpret_cast<float *>(v16sm + V16_ELEMS * 2);
    const int item = blockIdx.x / d, s = blockIdx.x - (blockIdx.x / d) * d;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = lane & 15, kg = lane >> 4;
    if (ctl.enabled && ctl.done[vnet[item]]) return;  // k_v16_finish carries the messages
    const int4 vi = vin[item], vo = vout[item], vd = vdim[item];
    const int ins[4] = {vi.x, vi.y, vi.z, vi.w};
    const int outs[4] = {vo.x, vo.y, vo.z, vo.w};
    const int dims[4] = {vd.x, vd.y, vd.z, vd.w};
    // incoming messages as MFMA A operands: lane l holds m[4(l/16)+t][l%16], scaled by
    // 1 / max|m| (identity on padded ports, zero beyond the bond dimension)
    v4s Am[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        float x[4];
        float mx = 0.f;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int a = 4 * kg + t, b = n, D = dims[j];
            if (ins[j] < 0) x[t] = a == b ? 1.f : 0.f;
            else x[t] = (a < D && b < D) ? mold[msg_off[ins[j]] + a * D + b] : 0.f;
            mx = fmaxf(mx, fabsf(x[t]));
        }
        mx = wave_max(mx);
        const float sc = mx > 0.f ? 1.f / mx : 1.f;
        Am[j] = bf16x4(x[0] * sc, x[1] * sc, x[2] * sc, x[3] * sc);
    }
    const unsigned short *src = vsites16 + ((long long)item * d + s) * V16_ELEMS;
    const unsigned short *src_t = vsites16t + ((long long)item * d + s) * V16_ELEMS;
    float *dst = part + ((long long)item * d + s) * 4 * 256;
    // E23 serves outgoing ports 0 (absorb 1) and 1 (absorb 0); E01 serves 2 and 3
    v16_pair<NW, 0, 1, 2, 3, 1, 0>(S, red, src, src_t, dst, ins, outs, Am);
    v16_pair<NW, 2, 3, 0, 1, 3, 2>(S, red, src, src_t, dst, ins, outs, Am);
}

// grid: items, 256 threads (wave j = port j): sum the slices, normalise, residuals, write.
__global__ __launch_bounds__(256) void k_v16_finish(
    const float *__restrict__ part, const float *__restrict__ mold, float *__restrict__ mnew,
    const int4 *__restrict__ vout, const int4 *__restrict__ vdim, const int *__restrict__ vnet,
    const long long *__restrict__ msg_off, NetCtl ctl, int d, double *__restrict__ res_fid,
    double *__restrict__ res_abs) {
    const int item = blockIdx.x;
    const int j = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int4 vo = vout[item], vd = vdim[item];
    const int outs[4] = {vo.x, vo.y, vo.z, vo.w};
    const int dims[4] = {vd.x, vd.y, vd.z, vd.w};
    const int q = outs[j], D = dims[j];
    if (q < 0) return;  // uniform over the wave
    const long long mo = msg_off[q];
    if (ctl.enabled && ctl.done[vnet[item]]) {
        for (int e = lane; e < D * D; e += 64) mnew[mo + e] = mold[mo + e];
        if (lane == 0) { res_fid[q] = 0.0; res_abs[q] = 0.0; }
        return;
    }
    float x[4];
    double ls = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int e = lane + 64 * k;
        float v = 0.f;
        for (int s = 0; s < d; ++s) v += part[(((long long)item * d + s) * 4 + j) * 256 + e];
        x[k] = v;
        ls += (double)v;
    }
    const double S = wave_sum_v(ls);
    double dn = 0.0, nn = 0.0, oo = 0.0, mx = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const int e = lane + 64 * k, a = e >> 4, b = e & 15;
        if (a >= D || b >= D) continue;  // padded entries are exactly zero
        const float xf = (float)(S != 0.0 ? (double)x[k] / S : (double)x[k]);
        const long long o = mo + a * D + b;
        const double xo = (double)mold[o];
        mnew[o] = xf;
        dn += (double)xf * xo;
        nn += (double)xf * (double)xf;
        oo += xo * xo;
        const double dd = fabs((double)xf - xo);
        mx = (dd > mx || dd != dd) ? dd : mx;
    }
    dn = wave_sum_v(dn);
    nn = wave_sum_v(nn);
    oo = wave_sum_v(oo);
    mx = wave_max_v(mx);
    if (lane == 0) {
        const double c = dn / sqrt(nn * oo);
        res_fid[q] = 1.0 - c * c;
        res_abs[q] = mx;
    }
}

// Padded BF16 site entries of the bucket: vsites16[item][s][p0][p1][p2][p3] (16 each), and
// vsites16t with the last two ports swapped; missing ports keep index 0 only, bond indices
// beyond D are zero.  One workgroup per (slice, p0): the vertex's bond dimensions are read
// once (missing ports count as dimension 1), then the 4096 entries [p1][p2][p3] follow.
__global__ __launch_bounds__(256) void k_v16_pad_sites(const double *sites64,
                                                       const long long *site_off,
                                                       const int *port_off, const int *port_edge,
                                                       const int *edge_dim, const int *vlist,
                                                       int d, unsigned short *out,
                                                       unsigned short *out_t) {
    const int us = blockIdx.x >> 4, p0 = blockIdx.x & 15;
    const int s = us % d, u = vlist[us / d];
    const int k = port_off[u + 1] - port_off[u];
    int D[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) D[j] = j < k ? edge_dim[port_edge[port_off[u] + j]] : 1;
    const long long n = (long long)D[0] * D[1] * D[2] * D[3];
    const double *src = sites64 + site_off[u] + s * n;
    const size_t base = (size_t)us * V16_ELEMS + p0 * 4096;
    for (int e = threadIdx.x; e < 4096; e += blockDim.x) {
        const int p1 = e >> 8, p2 = (e >> 4) & 15, p3 = e & 15;
        const bool real = p0 < D[0] && p1 < D[1] && p2 < D[2] && p3 < D[3];
        const unsigned short v =
            real ? bf16_rne((float)src[((p0 * D[1] + p1) * D[2] + p2) * D[3] + p3])
                 : (unsigned short)0;
        out[base + e] = v;
        out_t[base + (p1 << 8) + (p3 << 4) + p2] = v;  // [p0][p1][p3][p2]
    }
}

// ---------------------------------------------------------------- host side

static int vertex16_upload_sites(VertexPlan *p, hipStream_t st, const double *d_sites64,
                                 const long long *d_site_off, const int *d_port_off,
                                 const int *d_port_edge, const int *d_edge_dim) {
    if (p->nv16 == 0) return 0;
    if (!p->bf16) {  // FP32: the bond-dimension-12 kernels' layouts (degree 4 at side 10 and
                     // 12, degree <= 3)
        const int n10 = p->nv16_k10, n12 = p->nv16_k4 - p->nv16_k10;
        const int n4 = p->nv16_k4, n3 = p->nv16 - p->nv16_k4;
        if (p->d_v12sites) (void)hipFree(p->d_v12sites);
        if (p->d_v10sites) (void)hipFree(p->d_v10sites);
        if (p->d_v3sites) (void)hipFree(p->d_v3sites);
        p->d_v12sites = p->d_v10sites = nullptr;
        p->d_v3sites = nullptr;
        if (n10 > 0) {
            const size_t n = (size_t)n10 * p->d * VQ<10>::ELEMS;
            if (hipMalloc((void **)&p->d_v10sites, 2 * n * sizeof(float)) != hipSuccess) return 5;
            hipLaunchKernelGGL(k_v12_pad_sites<10>, dim3(n10 * p->d * 10), dim3(256), 0, st,
                               d_sites64, d_site_off, d_port_off, d_port_edge, d_edge_dim,
                               p->d_v16list, p->d, p->d_v10sites, p->d_v10sites + n);
        }
        if (n12 > 0) {
            const size_t n = (size_t)n12 * p->d * VQ<12>::ELEMS;
            if (hipMalloc((void **)&p->d_v12sites, 2 * n * sizeof(float)) != hipSuccess) return 5;
            hipLaunchKernelGGL(k_v12_pad_sites<12>, dim3(n12 * p->d * 12), dim3(256), 0, st,
                               d_sites64, d_site_off, d_port_off, d_port_edge, d_edge_dim,
                               p->d_v16list + n10, p->d, p->d_v12sites, p->d_v12sites + n);
        }
        if (n3 > 0) {
            const size_t n = (size_t)n3 * p->d * 3 * V3<12>::ELEMS;
            if (hipMalloc((void **)&p->d_v3sites, n * sizeof(float)) != hipSuccess) return 5;
            hipLaunchKernelGGL(k_v3_pad_sites<12>, dim3(4096), dim3(256), 0, st, d_sites64,
                               d_site_off, d_port_off, d_port_edge, d_edge_dim,
                               p->d_v16list + n4, p->d, n3, p->d_v3sites);
        }
        return hipGetLastError() == hipSuccess ? 0 : 2;
    }
    const size_t bytes = 2 * (size_t)p->nv16 * p->d * V16_ELEMS;
    if (p->d_v16sites) (void)hipFree(p->d_v16sites);
    p->d_v16sites = nullptr;
    if (hipMalloc((void **)&p->d_v16sites, 2 * bytes) != hipSuccess) return 5;
    hipLaunchKernelGGL(k_v16_pad_sites, dim3(p->nv16 * p->d * 16), dim3(256), 0, st, d_sites64,
                       d_site_off, d_port_off, d_port_edge, d_edge_dim, p->d_v16list, p->d,
                       p->d_v16sites, p->d_v16sites + bytes / 2);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

static int vertex16_set_lds_limit() {
    static int done = 0;
    if (done) return 0;
    const void *f[3] = {(const void *)k_vertex_d16<4>, (const void *)k_vertex_d16<8>,
                        (const void *)k_vertex_d16<16>};
    const int nw[3] = {4, 8, 16};
    for (int i = 0; i < 3; ++i)
        if (hipFuncSetAttribute(f[i], hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)v16_lds(nw[i])) != hipSuccess)
            return 2;
    if (hipFuncSetAttribute((const void *)k_vertex_d12<12>,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)VQ<12>::lds()) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_vertex_d12<10>,
                            hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)VQ<10>::lds()) != hipSuccess)
        return 2;
    done = 1;
    return 0;
}

// The large-bond bucket's part of a sweep: slices, then the per-message finish.
static int vertex16_sweep(VertexPlan *p, hipStream_t st, DevTopo tp, const void *src, void *dst,
                          const NetCtl &ctl, double *res_fid, double *res_abs) {
    if (p->nv16 == 0) return 0;
    if (!p->bf16) {
        const int n10 = p->nv16_k10, n12 = p->nv16_k4 - p->nv16_k10;
        const int n4 = p->nv16_k4, n3 = p->nv16 - p->nv16_k4;
        if (n10 > 0) {
            const size_t n = (size_t)n10 * p->d * VQ<10>::ELEMS;
            hipLaunchKernelGGL((k_vertex_d12<10>), dim3(n10 * p->d),
                               dim3(64 * VQ<10>::NW), (VQ<10>::lds()), st,
                               (const float *)p->d_v10sites,
                               (const float *)(p->d_v10sites + n), (const float *)src,
                               p->d_v16in, p->d_v16out, p->d_v16dim, p->d_v16net, tp.msg_off,
                               ctl, p->d, 0, p->d_v16part);
        }
        if (n12 > 0) {
            const size_t n = (size_t)n12 * p->d * VQ<12>::ELEMS;
            hipLaunchKernelGGL((k_vertex_d12<12>), dim3(n12 * p->d),
                               dim3(64 * VQ<12>::NW), (VQ<12>::lds()), st,
                               (const float *)p->d_v12sites,
                               (const float *)(p->d_v12sites + n), (const float *)src,
                               p->d_v16in, p->d_v16out, p->d_v16dim, p->d_v16net, tp.msg_off,
                               ctl, p->d, n10, p->d_v16part);
        }
        if (n3 > 0) {
            const int nsl = n3 * p->d;
            hipLaunchKernelGGL(k_vertex_k3<12>, dim3((nsl + V3_WAVES - 1) / V3_WAVES),
                               dim3(64 * V3_WAVES), 0, st, (const float *)p->d_v3sites,
                               (const float *)src, p->d_v16in, p->d_v16out, p->d_v16dim,
                               p->d_v16net, tp.msg_off, ctl, p->d, n4, n3,
                               p->d_v16part + (size_t)n4 * p->d * 4 * 256);
        }
        hipLaunchKernelGGL(k_v16_finish, dim3(p->nv16), dim3(256), 0, st,
                           (const float *)p->d_v16part, (const float *)src, (float *)dst,
                           p->d_v16out, p->d_v16dim, p->d_v16net, tp.msg_off, ctl, p->d,
                           res_fid, res_abs);
        return hipGetLastError() == hipSuccess ? 0 : 2;
    }
    const size_t half = (size_t)p->nv16 * p->d * V16_ELEMS;
    const unsigned short *s0 = p->d_v16sites, *s1 = p->d_v16sites + half;
#define V16_LAUNCH(NW)                                                                         \
    hipLaunchKernelGGL((k_vertex_d16<NW>), dim3(p->nv16 * p->d), dim3(64 * NW), v16_lds(NW),    \
                       st, s0, s1, (const float *)src, p->d_v16in, p->d_v16out, p->d_v16dim,   \
                       p->d_v16net, tp.msg_off, ctl, p->d, p->d_v16part)
    switch (p->v16_waves) {
        case 4: V16_LAUNCH(4); break;
        case 8: V16_LAUNCH(8); break;
        default: V16_LAUNCH(16); break;
    }
#undef V16_LAUNCH
    hipLaunchKernelGGL(k_v16_finish, dim3(p->nv16), dim3(256), 0, st, (const float *)p->d_v16part,
                       (const float *)src, (float *)dst, p->d_v16out, p->d_v16dim, p->d_v16net,
                       tp.msg_off, ctl, p->d, res_fid, res_abs);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

// ---------------------------------------------------------------- host side (D <= 8 bucket)

// FP32 networks of bond dimension <= 8 whose vertices all have degree <= 3 sit at the end of
// the D <= 8 bucket (items [nv - nv_k3, nv)); they sweep on k_vertex_k3<8> and the finisher
// instead of the four-port kernel (their padded four-port layout stays, for the adjoint).
static int vertex_k3_upload_sites(VertexPlan *p, hipStream_t st, const double *d_sites64,
                                  const long long *d_site_off, const int *d_port_off,
                                  const int *d_port_edge, const int *d_edge_dim) {
    if (p->d_vk3sites) (void)hipFree(p->d_vk3sites);
    p->d_vk3sites = nullptr;
    if (p->nv_k3 == 0) return 0;
    const size_t n = (size_t)p->nv_k3 * p->d * 3 * V3<8>::ELEMS;
    if (hipMalloc((void **)&p->d_vk3sites, n * sizeof(float)) != hipSuccess) return 5;
    hipLaunchKernelGGL(k_v3_pad_sites<8>, dim3(2048), dim3(256), 0, st, d_sites64, d_site_off,
                       d_port_off, d_port_edge, d_edge_dim, p->d_vlist + (p->nv - p->nv_k3),
                       p->d, p->nv_k3, p->d_vk3sites);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

static int vertex_k3_sweep(VertexPlan *p, hipStream_t st, DevTopo tp, const void *src,
                           void *dst, const NetCtl &ctl, double *res_fid, double *res_abs) {
    if (p->nv_k3 == 0) return 0;
    const int i0 = p->nv - p->nv_k3, nsl = p->nv_k3 * p->d;
    hipLaunchKernelGGL(k_vertex_k3<8>, dim3((nsl + V3_WAVES - 1) / V3_WAVES),
                       dim3(64 * V3_WAVES), 0, st, (const float *)p->d_vk3sites,
                       (const float *)src, p->d_vin, p->d_vout, p->d_vdim, p->d_vnet, tp.msg_off,
                       ctl, p->d, i0, p->nv_k3, p->d_vk3part);
    hipLaunchKernelGGL(k_v16_finish, dim3(p->nv_k3), dim3(256), 0, st,
                       (const float *)p->d_vk3part, (const float *)src, (float *)dst,
                       p->d_vout + i0, p->d_vdim + i0, p->d_vnet + i0, tp.msg_off, ctl, p->d,
                       res_fid, res_abs);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}
