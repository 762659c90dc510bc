// vertex12.h -- FP32 MFMA sweep for bond dimensions 9..12 (padded to 12, or to 10 for
// vertices whose bond dimensions are all <= 10) on gfx950.
//
// The FP32 counterpart of vertex16.h: one workgroup of 12 waves per (vertex, physical slice
// s), the slice A[s] (four ports of dimension 12, 12^4 entries) in LDS as FP32 (82 KiB),
// with rows of 144 (p2, p3) entries padded to 145 and blocks of 12 rows padded by one entry,
// so that the stride along every port is odd or 12.  Side 10 is the same kernel with 10
// waves, 10^4 entries (40 KiB, two workgroups per compute unit instead of one), 7 fibre
// tiles (the last one partly idle) and the contracted index stopping at 10.  (Dealing the
// tiles over 16 waves at side 12, or 8 at side 10, with the tile indices computed at run
// time instead of unrolled per wave, measured 20-25% slower.)
//
// Mode product on port J ("T x_J m_J", in place) runs on v_mfma_f32_16x16x4_f32: A operand
// = m_J transposed (lane l holds m_J[4 kc + l/16][l%16], kc = 0..2 for the 12 values of the
// contracted index), B operand = 16 fibres along J (lane l: entry J = 4 kc + l/16 of fibre
// l%16).  The fibres of a tile run over the two later ports of the other three, flattened
// (144 = 9 tiles of 16); wave w takes the 9 tiles whose first other port is w.  The 16 x 16
// result holds rows J = 0..11 of the tile (rows 12..15 are the zero rows of the padded A
// operand and are dropped); every tile is read whole before it is written, and tiles of
// different waves are disjoint, so the product is in place.
//
// Open contraction out_O[a][a'] = sum_K G[a, K] A[a', K] (K: the other three ports, 1728
// values) runs on the same MFMA: G from LDS, the bra A from HBM (L2 / Infinity Cache),
// prefetched as 16-byte loads before the mode products.  Wave w takes the K values whose
// outer other port is w; for O = 0, 1 the bra rows come from the natural copy
// [p0][p1][p2][p3], for O = 2, 3 from a copy laid out [p2][p3][p0][p1], so that each lane's
// four K values are contiguous in both.  The twelve waves' 16 x 16 partials are summed
// through LDS into the same per-slice buffer as the BF16 kernel, and k_v16_finish sums the
// slices, normalises, writes and forms the residuals.
//
// Per slice: the partial product E23 (ports 2, 3 absorbed) serves outgoing ports 0 and 1,
// E01 serves 2 and 3; while the first of each pair absorbs its third message in LDS, the
// partial waits in registers (28 words per thread), so a slice costs 8 mode products and
// two loads from HBM.  Ports of a padded (dummy) slot are skipped (identity message).
// Arithmetic: FP32 operands and accumulation throughout.
#pragma once

#include "vertex.h"

// Layout of a padded degree-4 slice of side DP (12, or 10 for items whose bond dimensions
// are all <= 10) in LDS: rows of DP^2 (p2, p3) entries padded by one, blocks of DP rows
// padded by one, so that the stride along every port is odd or DP.  One wave per value of
// a tile's first other port: wave w owns the FT fibre tiles of that value.
template <int DP> struct VQ {
    static constexpr int ELEMS = DP * DP * DP * DP;  // per slice (HBM copies, unpadded)
    static constexpr int NW = DP;                    // waves per slice workgroup
    static constexpr int S3 = 1, S2 = DP, S1 = DP * DP + 1, S0 = DP * S1 + 1;
    static constexpr int LDS_ELEMS = DP * S0;
    static constexpr int FT = (DP * DP + 15) / 16;   // fibre tiles of 16 over two ports
    static constexpr int PARK = (LDS_ELEMS + 64 * NW - 1) / (64 * NW);
    static constexpr size_t lds() { return sizeof(float) * (LDS_ELEMS + NW * 256); }
    static constexpr int str(int j) { return j == 0 ? S0 : (j == 1 ? S1 : (j == 2 ? S2 : S3)); }
};

// The other three ports of J in ascending order: tiles are indexed by P (one per wave),
// fibres by (Q, R) flattened.
template <int J> struct V12Ports {
    static constexpr int P = J == 0 ? 1 : 0;
    static constexpr int Q = J <= 1 ? 2 : 1;
    static constexpr int R = J == 3 ? 2 : 3;
};

// Mode product on port J of the slice in LDS, in place.  Am: the message as A operand (three
// chunks of 4 of the contracted index; entries beyond the bond dimension are zero).  At
// DP = 10 the last fibre tile is partly idle (100 = 6 x 16 + 4) and the contracted index
// stops at 10: idle lanes and indices read zeros and store nothing.
template <int DP, int J>
__device__ __forceinline__ void v12_mode(float *S, const float (&Am)[3]) {
    using C = VQ<DP>;
    using PT = V12Ports<J>;
    constexpr int SJ = C::str(J), SP = C::str(PT::P), SQ = C::str(PT::Q), SR = C::str(PT::R);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = lane & 15, kk = lane >> 4;
    const v4f zero = {0.f, 0.f, 0.f, 0.f};
    const int pbase = w * SP;
#pragma unroll
    for (int jf = 0; jf < C::FT; ++jf) {
        const int f = 16 * jf + n;  // fibre over (Q, R): 0..DP^2 - 1
        const bool live = f < DP * DP;
        const int q = f / DP, r = f - DP * q;
        const int base = pbase + q * SQ + r * SR;
        float b[3];
#pragma unroll
        for (int kc = 0; kc < 3; ++kc) {
            const int k = 4 * kc + kk;
            b[kc] = (live && k < DP) ? S[base + k * SJ] : 0.f;
        }
        v4f c = zero;
#pragma unroll
        for (int kc = 0; kc < 3; ++kc)
            c = __builtin_amdgcn_mfma_f32_16x16x4f32(Am[kc], b[kc], c, 0, 0, 0);
        if (live) {
#pragma unroll
            for (int t = 0; t < 4; ++t)
                if (4 * kk + t < DP) S[base + (4 * kk + t) * SJ] = c[t];
        }
    }
}

// Bra operands of this wave's MFMA steps of the open contraction on port O: step (jm, t) of
// lane (n, kk) takes K = (hi = w, lo = 16 jm + 4 kk + t), a' = n (zero for lo >= DP^2).
template <int DP, int O>
__device__ __forceinline__ void v12_bra_prefetch(const float *natural, const float *swapped,
                                                 float4 (&pre)[VQ<DP>::FT]) {
    constexpr int D2 = DP * DP, D3 = DP * DP * DP;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = lane & 15, kk = lane >> 4;
    const float *src = O < 2 ? natural : swapped;
    const int row = (O == 0 || O == 2) ? n * D3 + D2 * w : w * D3 + n * D2;
#pragma unroll
    for (int jm = 0; jm < VQ<DP>::FT; ++jm) {
        const int lo = 16 * jm + 4 * kk;
        if (n < DP && lo < D2) pre[jm] = *reinterpret_cast<const float4 *>(src + row + lo);
        else pre[jm] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// This wave's share of out_O[a][a'] = sum_K G[a, K] A[a', K]; lane l returns entries
// (4(l/16)+i, l%16).  K = (hi, lo) maps to ports as: O = 0: (p1, p2, p3), O = 1: (p0, p2,
// p3), O = 2: (p3, p0, p1), O = 3: (p2, p0, p1), with lo = DP x + y over the last two.
template <int DP, int O>
__device__ __forceinline__ v4f v12_open(const float *S, const float4 (&pre)[VQ<DP>::FT]) {
    using C = VQ<DP>;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = lane & 15, kk = lane >> 4;
    const bool live = n < DP;
    const int a = live ? n : 0;
    int base;
    if constexpr (O == 0) base = a * C::S0 + w * C::S1;
    else if constexpr (O == 1) base = w * C::S0 + a * C::S1;
    else if constexpr (O == 2) base = a * C::S2 + w * C::S3;
    else base = w * C::S2 + a * C::S3;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int jm = 0; jm < C::FT; ++jm) {
        const float pv[4] = {pre[jm].x, pre[jm].y, pre[jm].z, pre[jm].w};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int lo = 16 * jm + 4 * kk + t;
            const int x = lo / DP, y = lo - DP * x;
            const int off = O < 2 ? x * C::S2 + y * C::S3 : x * C::S0 + y * C::S1;
            const float g = (live && lo < DP * DP) ? S[base + off] : 0.f;
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(g, pv[t], acc, 0, 0, 0);
        }
    }
    return acc;
}

// Copy the slice's natural-layout entries into the padded LDS layout (16-byte loads; DP^2
// is a multiple of 4, so no load straddles two rows).
template <int DP>
__device__ __forceinline__ void v12_load_slice(float *S, const float *src) {
    using C = VQ<DP>;
    const float4 *g = reinterpret_cast<const float4 *>(src);
    for (int c = threadIdx.x; c < C::ELEMS / 4; c += 64 * C::NW) {
        const float4 v = g[c];
        const int e = 4 * c;
        const int row = e / (DP * DP), col = e - DP * DP * row;  // row = DP p0 + p1
        const int p0 = row / DP, p1 = row - DP * p0;
        float *d = S + p0 * C::S0 + p1 * C::S1 + col;
        d[0] = v.x;
        d[1] = v.y;
        d[2] = v.z;
        d[3] = v.w;
    }
}

template <int DP, int J>
__device__ __forceinline__ void v12_mode_if(float *S, const int (&ins)[4],
                                            const float (&Am)[4][3]) {
    if (ins[J] < 0) return;  // identity message on a padded port (uniform)
    v12_mode<DP, J>(S, Am[J]);
    __syncthreads();
}

// Open on port O with the bra operands in `pre`, reduce the waves' shares through LDS,
// store the unnormalised 16 x 16 block.
template <int DP, int O>
__device__ __forceinline__ void v12_open_store(const float *S, float *red, float *dst,
                                               const float4 (&pre)[VQ<DP>::FT]) {
    constexpr int NW = VQ<DP>::NW;
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = lane & 15, kk = lane >> 4;
    const v4f acc = v12_open<DP, O>(S, pre);
#pragma unroll
    for (int i = 0; i < 4; ++i) red[w * 256 + (4 * kk + i) * 16 + n] = acc[i];
    __syncthreads();
    for (int e = threadIdx.x; e < 256; e += 64 * NW) {
        float x = 0.f;
#pragma unroll
        for (int v = 0; v < NW; ++v) x += red[v * 256 + e];
        dst[O * 256 + e] = x;
    }
}

// Raw LDS words of the slice buffer held by this thread while the buffer is reused (DP = 12:
// 20892 words over 768 threads, 28 each; DP = 10: 10110 over 640, 16 each).
template <int DP>
__device__ __forceinline__ void v12_park(const float *S, float (&pk)[VQ<DP>::PARK]) {
    using C = VQ<DP>;
#pragma unroll
    for (int i = 0; i < C::PARK; ++i) {
        const int a = threadIdx.x + 64 * C::NW * i;
        pk[i] = a < C::LDS_ELEMS ? S[a] : 0.f;
    }
}

template <int DP>
__device__ __forceinline__ void v12_unpark(float *S, const float (&pk)[VQ<DP>::PARK]) {
    using C = VQ<DP>;
#pragma unroll
    for (int i = 0; i < C::PARK; ++i) {
        const int a = threadIdx.x + 64 * C::NW * i;
        if (a < C::LDS_ELEMS) S[a] = pk[i];
    }
}

// Outgoing ports O1 and O2 of the slice from one shared partial product: reload the slice,
// absorb ports X0, X1 (E23 for O = 0, 1; E01 for O = 2, 3), keep the partial in registers,
// then absorb Y1 and open O1, restore the partial, absorb Y2 and open O2.
template <int DP, int O1, int O2, int X0, int X1, int Y1, int Y2>
__device__ __forceinline__ void v12_pair(float *S, float *red, const float *src,
                                         const float *src_w, float *dst, const int (&ins)[4],
                                         const int (&outs)[4], const float (&Am)[4][3]) {
    const bool first = outs[O1] >= 0, second = outs[O2] >= 0;  // uniform
    if (!first && !second) return;
    float4 pre[VQ<DP>::FT];
    if (first) v12_bra_prefetch<DP, O1>(src, src_w, pre);
    else v12_bra_prefetch<DP, O2>(src, src_w, pre);
    __syncthreads();  // earlier opens have finished with the slice and the partials
    v12_load_slice<DP>(S, src);
    __syncthreads();
    v12_mode_if<DP, X0>(S, ins, Am);
    v12_mode_if<DP, X1>(S, ins, Am);
    float pk[VQ<DP>::PARK];
    if (first && second) {
        v12_park<DP>(S, pk);
        __syncthreads();
    }
    if (first) {
        v12_mode_if<DP, Y1>(S, ins, Am);
        v12_open_store<DP, O1>(S, red, dst, pre);
    }
    if (second) {
        if (first) {
            v12_bra_prefetch<DP, O2>(src, src_w, pre);
            v12_unpark<DP>(S, pk);  // the open above has passed its barrier
            __syncthreads();
        }
        v12_mode_if<DP, Y2>(S, ins, Am);
        v12_open_store<DP, O2>(S, red, dst, pre);
    }
}

// grid: items x d (one workgroup per slice of items i0 .. i0 + grid / d - 1), 64 DP threads.
// vsites_n / vsites_w hold those items' padded slices from item i0 on.  part[((item * d + s)
// * 4 + j) * 256 + a * 16 + a'] receives slice s's unnormalised out_j (the layout
// k_v16_finish reads).
template <int DP>
__global__ __launch_bounds__(64 * DP) void k_vertex_d12(
    const float *__restrict__ vsites_n, const float *__restrict__ vsites_w,
    const float *__restrict__ mold, const int4 *__restrict__ vin, const int4 *__restrict__ vout,
    const int4 *__restrict__ vdim, const int *__restrict__ vnet,
    const long long *__restrict__ msg_off, NetCtl ctl, int d, int i0, float *__restrict__ part) {
    using C = VQ<DP>;
    extern __shared__ __align__(16) float v12sm[];
    float *S = v12sm;
    float *red = v12sm + C::LDS_ELEMS;
    const int il = blockIdx.x / d, s = blockIdx.x - il * d, item = i0 + il;
    const int lane = threadIdx.x & 63;
    const int n = lane & 15, kk = lane >> 4;
    if (ctl.enabled && ctl.done[vnet[item]]) return;  // k_v16_finish carries the messages
    const int4 vi = vin[item], vo = vout[item], vd = vdim[item];
    const int ins[4] = {vi.x, vi.y, vi.z, vi.w};
    const int outs[4] = {vo.x, vo.y, vo.z, vo.w};
    const int dims[4] = {vd.x, vd.y, vd.z, vd.w};
    // incoming messages as MFMA A operands: lane l holds m[4 kc + l/16][l%16] (zero beyond
    // the bond dimension; padded ports are never absorbed)
    float Am[4][3];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int kc = 0; kc < 3; ++kc) {
            const int k = 4 * kc + kk, D = dims[j];
            Am[j][kc] = (ins[j] >= 0 && k < D && n < D) ? mold[msg_off[ins[j]] + k * D + n]
                                                         : 0.f;
        }
    }
    const long long so = ((long long)il * d + s) * C::ELEMS;
    const float *src = vsites_n + so, *src_w = vsites_w + so;
    float *dst = part + ((long long)item * d + s) * 4 * 256;
    // E23 serves outgoing ports 0 (absorb 1) and 1 (absorb 0); E01 serves 2 and 3
    v12_pair<DP, 0, 1, 2, 3, 1, 0>(S, red, src, src_w, dst, ins, outs, Am);
    v12_pair<DP, 2, 3, 0, 1, 3, 2>(S, red, src, src_w, dst, ins, outs, Am);
}

// Padded FP32 site entries of DP-side items: vsites_n[item][s][p0][p1][p2][p3] (DP each)
// and vsites_w[item][s][p2][p3][p0][p1]; missing ports keep index 0 only, bond indices
// beyond D are zero.  One workgroup per (slice, p0): the vertex's bond dimensions are read
// once (missing ports count as dimension 1).
template <int DP>
__global__ __launch_bounds__(256) void k_v12_pad_sites(const double *sites64,
                                                       const long long *site_off,
                                                       const int *port_off, const int *port_edge,
                                                       const int *edge_dim, const int *vlist,
                                                       int d, float *out_n, float *out_w) {
    constexpr int D2 = DP * DP, D3 = DP * DP * DP;
    const int us = blockIdx.x / DP, p0 = blockIdx.x - DP * us;
    const int s = us % d, u = vlist[us / d];
    const int k = port_off[u + 1] - port_off[u];
    int D[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) D[j] = j < k ? edge_dim[port_edge[port_off[u] + j]] : 1;
    const long long cnt = (long long)D[0] * D[1] * D[2] * D[3];
    const double *src = sites64 + site_off[u] + s * cnt;
    float *dn = out_n + (size_t)us * VQ<DP>::ELEMS + p0 * D3;
    float *dw = out_w + (size_t)us * VQ<DP>::ELEMS;
    for (int e = threadIdx.x; e < D3; e += blockDim.x) {
        const int p1 = e / D2, r = e - D2 * p1, p2 = r / DP, p3 = r - DP * p2;
        const bool real = p0 < D[0] && p1 < D[1] && p2 < D[2] && p3 < D[3];
        const float v = real ? (float)src[((p0 * D[1] + p1) * D[2] + p2) * D[3] + p3] : 0.f;
        dn[e] = v;
        dw[((p2 * DP + p3) * DP + p0) * DP + p1] = v;
    }
}

// ---------------------------------------------------------------- degree <= 3 vertices
//
// A vertex with at most three ports has a DP^3 slice (DP = 12: 6.9 KiB; DP = 8: 2 KiB, for
// networks of bond dimension <= 8 whose vertices all have degree <= 3, such as heavy-hex
// lattices): one wave owns one slice (four slices per workgroup, no workgroup barriers),
// with the same MFMA tiles as above: fibres over the two other ports flattened (DP^2 / 16
// tiles of 16), contraction index in DP / 4 chunks of 4.  Per slice: E2 (port 2 absorbed)
// serves outgoing ports 0 and 1 (the partial parked in registers meanwhile), port 2 absorbs
// 0 and 1 from a reload: 4 mode products, 3 opens of K = DP^2.  The bra of the open on
// port O comes from a copy of the slice with port O outermost, so each lane's four K
// values are 16 contiguous bytes.  A degree-2 (or 1) vertex runs the same code with the
// missing ports skipped.  Results go to per-slice 16 x 16 partials that k_v16_finish
// completes.
template <int DP> struct V3 {
    static constexpr int ELEMS = DP * DP * DP;     // per slice (HBM copies, unpadded)
    static constexpr int S0 = DP * DP + 1, S1 = DP, S2 = 1;  // LDS strides of (p0, p1, p2)
    static constexpr int LDS = DP * S0;            // words per wave
    static constexpr int TILES = DP * DP / 16;     // fibre tiles (and K chunks of 16)
    static constexpr int KC = DP / 4;              // contraction chunks of 4
    static constexpr int PARK = (LDS + 63) / 64;
    static constexpr int str(int j) { return j == 0 ? S0 : (j == 1 ? S1 : S2); }
};
static constexpr int V3_WAVES = 4;  // slices per workgroup

template <int J> struct V3Ports {
    static constexpr int P = J == 0 ? 1 : 0;
    static constexpr int Q = J == 2 ? 1 : 2;
};

template <int DP, int J>
__device__ __forceinline__ void v3_mode(float *S, const float (&Am)[3]) {
    using C = V3<DP>;
    using PT = V3Ports<J>;
    constexpr int SJ = C::str(J), SP = C::str(PT::P), SQ = C::str(PT::Q);
    const int lane = threadIdx.x & 63;
    const int n = lane & 15, kk = lane >> 4;
    const v4f zero = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int jf = 0; jf < C::TILES; ++jf) {
        const int f = 16 * jf + n;  // fibre over (P, Q)
        const int p = f / DP, q = f - DP * p;
        const int base = p * SP + q * SQ;
        float b[C::KC];
#pragma unroll
        for (int kc = 0; kc < C::KC; ++kc) b[kc] = S[base + (4 * kc + kk) * SJ];
        v4f c = zero;
#pragma unroll
        for (int kc = 0; kc < C::KC; ++kc)
            c = __builtin_amdgcn_mfma_f32_16x16x4f32(Am[kc], b[kc], c, 0, 0, 0);
        if (kk < C::KC) {
#pragma unroll
            for (int t = 0; t < 4; ++t) S[base + (4 * kk + t) * SJ] = c[t];
        }
    }
    wave_lds_sync();  // the next product reads other lanes' rows
}

template <int DP, int J>
__device__ __forceinline__ void v3_mode_if(float *S, const int (&ins)[4],
                                           const float (&Am)[4][3]) {
    if (ins[J] < 0) return;  // missing or padded port (uniform over the wave)
    v3_mode<DP, J>(S, Am[J]);
}

// bra rows of the open on port O: copy O of the slice, [O][the other two ascending]
template <int DP>
__device__ __forceinline__ void v3_bra_prefetch(const float *bra, float4 (&pre)[V3<DP>::TILES]) {
    const int lane = threadIdx.x & 63;
    const int n = lane & 15, kk = lane >> 4;
#pragma unroll
    for (int j = 0; j < V3<DP>::TILES; ++j) {
        if (n < DP) pre[j] = *reinterpret_cast<const float4 *>(bra + n * DP * DP + 16 * j + 4 * kk);
        else pre[j] = make_float4(0.f, 0.f, 0.f, 0.f);
    }
}

// out_O[a][a'] over K = the two other ports (x, y) = divmod(K, DP); lane l returns entries
// (4(l/16)+i, l%16)
template <int DP, int O>
__device__ __forceinline__ v4f v3_open(const float *S, const float4 (&pre)[V3<DP>::TILES]) {
    using C = V3<DP>;
    constexpr int SO = C::str(O), SX = C::str(O == 0 ? 1 : 0), SY = C::str(O == 2 ? 1 : 2);
    const int lane = threadIdx.x & 63;
    const int n = lane & 15, kk = lane >> 4;
    const bool live = n < DP;
    const int base = (live ? n : 0) * SO;
    v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < C::TILES; ++j) {
        const float pv[4] = {pre[j].x, pre[j].y, pre[j].z, pre[j].w};
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int K = 16 * j + 4 * kk + t;
            const int x = K / DP, y = K - DP * x;
            const float g = live ? S[base + x * SX + y * SY] : 0.f;
            acc = __builtin_amdgcn_mfma_f32_16x16x4f32(g, pv[t], acc, 0, 0, 0);
        }
    }
    return acc;
}

template <int DP>
__device__ __forceinline__ void v3_load_slice(float *S, const float *src) {
    using C = V3<DP>;
    const int lane = threadIdx.x & 63;
    const float4 *g = reinterpret_cast<const float4 *>(src);
    for (int c = lane; c < C::ELEMS / 4; c += 64) {
        const float4 v = g[c];
        const int e = 4 * c;
        const int p0 = e / (DP * DP), rem = e - DP * DP * p0;  // rem = DP p1 + p2
        float *d = S + p0 * C::S0 + rem;
        d[0] = v.x;
        d[1] = v.y;
        d[2] = v.z;
        d[3] = v.w;
    }
    wave_lds_sync();
}

template <int DP, int O>
__device__ __forceinline__ void v3_open_store(const float *S, const float4 (&pre)[V3<DP>::TILES],
                                              float *dst) {
    const int lane = threadIdx.x & 63;
    const int n = lane & 15, kk = lane >> 4;
    const v4f acc = v3_open<DP, O>(S, pre);
#pragma unroll
    for (int i = 0; i < 4; ++i) dst[O * 256 + (4 * kk + i) * 16 + n] = acc[i];
}

// grid: ceil(nitems * d / 4) workgroups of 4 waves; wave w takes slice 4 blockIdx + w of the
// items [item0, item0 + nitems) (for vin / vout / vdim / vnet); vs3[(i * d + s) * 3 + O]
// is slice s of the i-th item with port O outermost, and part[(i * d + s) * 4 + j] receives
// its unnormalised out_j.
template <int DP>
__global__ __launch_bounds__(64 * V3_WAVES) void k_vertex_k3(
    const float *__restrict__ vs3, const float *__restrict__ mold, const int4 *__restrict__ vin,
    const int4 *__restrict__ vout, const int4 *__restrict__ vdim, const int *__restrict__ vnet,
    const long long *__restrict__ msg_off, NetCtl ctl, int d, int item0, int nitems,
    float *__restrict__ part) {
    using C = V3<DP>;
    __shared__ float v3sm[V3_WAVES * C::LDS];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int n = lane & 15, kk = lane >> 4;
    const int slice = blockIdx.x * V3_WAVES + w;
    if (slice >= nitems * d) return;  // uniform over the wave; no workgroup barriers below
    const int i = slice / d, s = slice - i * d;
    const int item = item0 + i;
    if (ctl.enabled && ctl.done[vnet[item]]) return;
    float *S = v3sm + w * C::LDS;
    const int4 vi = vin[item], vo = vout[item], vd = vdim[item];
    const int ins[4] = {vi.x, vi.y, vi.z, vi.w};
    const int outs[4] = {vo.x, vo.y, vo.z, vo.w};
    const int dims[4] = {vd.x, vd.y, vd.z, vd.w};
    float Am[4][3];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
#pragma unroll
        for (int kc = 0; kc < 3; ++kc) {
            const int k = 4 * kc + kk, D = dims[j];
            Am[j][kc] = (j < 3 && kc < C::KC && ins[j] >= 0 && k < D && n < D)
                            ? mold[msg_off[ins[j]] + k * D + n]
                            : 0.f;
        }
    }
    const float *src = vs3 + ((long long)i * d + s) * 3 * C::ELEMS;
    float *dst = part + ((long long)i * d + s) * 4 * 256;
    float4 pre[C::TILES];
    v3_load_slice<DP>(S, src);
    // E2 -> outgoing ports 0 (absorb 1) and 1 (absorb 0)
    const bool o0 = outs[0] >= 0, o1 = outs[1] >= 0;
    v3_mode_if<DP, 2>(S, ins, Am);
    float pk[C::PARK];
    if (o0 && o1) {
#pragma unroll
        for (int r = 0; r < C::PARK; ++r) {
            const int a = lane + 64 * r;
            pk[r] = a < C::LDS ? S[a] : 0.f;
        }
    }
    if (o0) {
        v3_bra_prefetch<DP>(src + 0 * C::ELEMS, pre);
        v3_mode_if<DP, 1>(S, ins, Am);
        v3_open_store<DP, 0>(S, pre, dst);
    }
    if (o1) {
        v3_bra_prefetch<DP>(src + 1 * C::ELEMS, pre);
        if (o0) {
            wave_lds_sync();  // the open above has finished reading the slice
#pragma unroll
            for (int r = 0; r < C::PARK; ++r) {
                const int a = lane + 64 * r;
                if (a < C::LDS) S[a] = pk[r];
            }
            wave_lds_sync();
        }
        v3_mode_if<DP, 0>(S, ins, Am);
        v3_open_store<DP, 1>(S, pre, dst);
    }
    // outgoing port 2: absorb 0 and 1 from a reload
    if (outs[2] >= 0) {
        v3_bra_prefetch<DP>(src + 2 * C::ELEMS, pre);
        wave_lds_sync();
        v3_load_slice<DP>(S, src);
        v3_mode_if<DP, 0>(S, ins, Am);
        v3_mode_if<DP, 1>(S, ins, Am);
        v3_open_store<DP, 2>(S, pre, dst);
    }
}

// Site copies of degree <= 3 items: vs3[(i * d + s) * 3 + O][a'][x][y], port O first, then
// the other two ascending; missing ports keep index 0 only, bond indices beyond D are zero.
template <int DP>
__global__ void k_v3_pad_sites(const double *sites64, const long long *site_off,
                               const int *port_off, const int *port_edge, const int *edge_dim,
                               const int *vlist, int d, int nitems, float *out) {
    constexpr int E = V3<DP>::ELEMS;
    const long long total = (long long)nitems * d * E;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int e = (int)(i % E);
        const long long us = i / E;
        const int s = (int)(us % d);
        const int u = vlist[us / d];
        const int k = port_off[u + 1] - port_off[u];
        const int idx[3] = {e / (DP * DP), (e / DP) % DP, e % DP};
        long long off = 0, cnt = 1;
        bool real = true;
        for (int j = 0; j < 3; ++j) {
            if (j < k) {
                const int D = edge_dim[port_edge[port_off[u] + j]];
                if (idx[j] >= D) real = false;
                off = off * D + idx[j];
                cnt *= D;
            } else if (idx[j] != 0) {
                real = false;
            }
        }
        const float v = real ? (float)sites64[site_off[u] + s * cnt + off] : 0.f;
        float *o = out + us * 3 * E;
        o[e] = v;                                                 // [p0][p1][p2]
        o[E + (idx[1] * DP + idx[0]) * DP + idx[2]] = v;          // [p1][p0][p2]
        o[2 * E + (idx[2] * DP + idx[0]) * DP + idx[1]] = v;      // [p2][p0][p1]
    }
}

