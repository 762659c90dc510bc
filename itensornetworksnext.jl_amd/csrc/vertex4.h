// vertex4.h -- BP sweep for networks whose bond dimensions are all <= 4 (FP32, VALU).
//
// Below D = 8 the padded MFMA kernel of vertex.h does (8 / D)^5 times the work (32x at
// D = 4) and the generic kernel pays a workgroup and its barriers per message.  Here a
// physical slice A_u[s] (4 ports padded to DP = 2, 3 or 4) lives in one quad of lanes:
// lane p0 of the quad holds the DP^3 entries A[s][p0][p1][p2][p3] (register
// r = (p1 * DP + p2) * DP + p3); lanes p0 >= DP hold zeros.  A wave holds 16 slices:
// W = 16 / d vertices of d slices each, quad (vi * d + s).  Mode products on ports 1..3
// are in-lane; on port 0 they read the other lanes of the quad through DPP quad
// permutations (no LDS).  The sweep shares the partial products as the D <= 8 kernel does:
//     phase A: E = A x0 m0 x1 m1;  out2 = open_2(E x3 m3, A),  out3 = open_3(E x2 m2, A)
//     phase B: E = A x2 m2 x3 m3;  out1 = open_1(E x0 m0, A),  out0 = open_0(E x1 m1, A)
// (the last mode product of each open is fused into it, block by block).  Per-slice
// results go through LDS; one lane per message sums the d slices, normalises by the sum
// (when nonzero) and forms the residuals in FP64 (beliefpropagation.jl:242-257, 261-267).
//
// Degree and bond padding as in vertex.h: a missing port has dimension 1 (entries vanish
// beyond index 0, its message is e0 e0^T), bonds below DP are zero-padded.
//
// Padded site layout of a DP bucket: float[group][RP / 4][64 lanes][4], RP = DP^3 rounded
// up to a multiple of 4, so each lane's float4 load is 16 contiguous bytes and a wave's load
// one contiguous 1 KiB line.
#pragma once

#include "vertex.h"

template <int DP>
struct VsShape {
    static constexpr int R = DP * DP * DP;        // entries of a slice in one lane
    static constexpr int RP = (R + 3) & ~3;       // padded to float4
    static constexpr int M2 = DP * DP;            // entries of a padded message
};

// sum over the 4 lanes of the quad (quad_perm xor 1, then xor 2): every lane gets it
__device__ __forceinline__ float quad_sum(float v) {
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));
    return v;
}

// v_mfma_f32_4x4x1f32 works on 16 blocks of 4 lanes, one block per quad here: register i of
// lane j of a quad accumulates a(lane i) * b(lane j).

// In-lane mode product on port P (1, 2 or 3), in place, on the matrix cores:
// X[.., b, ..] = sum_p X[.., p, ..] m[p][b].  With a = mc[p] = m[p][lane j] and b = the lane's
// own X[.., p, ..], register i of lane j collects sum_p m[p][i] X_j[.., p, ..].
template <int DP, int P>
__device__ __forceinline__ void vs_mode_in(float (&X)[VsShape<DP>::RP], const float (&mc)[DP]) {
    constexpr int st = P == 1 ? DP * DP : P == 2 ? DP : 1;  // register stride of port P
#pragma unroll
    for (int o = 0; o < DP * DP; ++o) {
        const int hi = o / DP, lo = o % DP;  // the other two in-lane digits
        const int base = P == 1 ? hi * DP + lo : P == 2 ? hi * DP * DP + lo : (hi * DP + lo) * DP;
        v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int p = 0; p < DP; ++p)
            acc = __builtin_amdgcn_mfma_f32_4x4x1f32(mc[p], X[base + p * st], acc, 0, 0, 0);
#pragma unroll
        for (int i = 0; i < DP; ++i) X[base + i * st] = acc[i];
    }
}

// The value x of every lane of the quad (a = x, b = 1): g[i] = x of lane i.
__device__ __forceinline__ v4f quad_gather(float x) {
    const v4f z = {0.f, 0.f, 0.f, 0.f};
    return __builtin_amdgcn_mfma_f32_4x4x1f32(x, 1.f, z, 0, 0, 0);
}

// Port-0 mode product, in place: lane b of the quad gets X_b[r] = sum_p X_p[r] m0[p][b]
// (c[p] = m0[p][this lane's b], zero for lanes >= DP): a gather, then DP FMAs.
template <int DP>
__device__ __forceinline__ void vs_mode_0(float (&X)[VsShape<DP>::RP], const float (&c)[DP]) {
#pragma unroll
    for (int r = 0; r < VsShape<DP>::R; ++r) {
        const v4f gx = quad_gather(X[r]);
        float acc = 0.f;
#pragma unroll
        for (int p = 0; p < DP; ++p) acc = fmaf(gx[p], c[p], acc);
        X[r] = acc;
    }
}

template <int DP>
__global__ __launch_bounds__(64) void k_vertex_s(
    const float *__restrict__ vs_sites, const float *__restrict__ mold, float *__restrict__ mnew,
    const int *__restrict__ vs_ino, const int *__restrict__ vs_outo,
    const int *__restrict__ vs_out, const int *__restrict__ vs_dim, const int *__restrict__ vnet,
    NetCtl ctl, double *__restrict__ res_fid, double *__restrict__ res_abs, int nitems, int d) {
    using Shp = VsShape<DP>;
    constexpr int R = Shp::R, RP = Shp::RP, M2 = Shp::M2;
    __shared__ float msm[16 * 4 * M2];   // [vi][port][p * DP + b], padded incoming messages
    __shared__ float part[16 * 4 * M2];  // [quad][port][a * DP + a'], per-slice results
    const int g = blockIdx.x, lane = threadIdx.x;
    const int quad = lane >> 2, j = lane & 3;
    const int W = 16 / d;

    // every independent load first: this lane's message (vertex vi2, port), its slice
    const int vi2 = lane >> 2, port = lane & 3, b2 = g * W + vi2;
    const bool mine = vi2 < W && b2 < nitems;
    long long ino = -2, outo = 0;
    int qout = -1, Dm = 0, net2 = 0;
    if (mine) {
        ino = vs_ino[4 * b2 + port];
        outo = vs_outo[4 * b2 + port];
        qout = vs_out[4 * b2 + port];
        Dm = (vs_dim[b2] >> (8 * port)) & 255;
        net2 = vnet[b2];
    }
    const int vi = quad / d, b = g * W + vi;
    const bool inr = vi < W && b < nitems;  // uniform over the quad
    float A[RP];
    if (inr) {
        const float4 *src = reinterpret_cast<const float4 *>(vs_sites) + (size_t)g * (RP / 4) * 64;
#pragma unroll
        for (int k = 0; k < RP / 4; ++k) {
            const float4 v = src[k * 64 + lane];
            A[4 * k] = v.x;
            A[4 * k + 1] = v.y;
            A[4 * k + 2] = v.z;
            A[4 * k + 3] = v.w;
        }
    } else {
#pragma unroll
        for (int r = 0; r < RP; ++r) A[r] = 0.f;
    }
    const bool live = inr && !(ctl.enabled && ctl.done[vnet[b]]);
    const bool done2 = mine && ctl.enabled && ctl.done[net2];
    // the incoming message of (vi2, port), zero-padded to DP x DP (padded port: e0 e0^T)
    {
        float *dst = msm + (vi2 * 4 + port) * M2;
#pragma unroll
        for (int e = 0; e < M2; ++e) {
            const int p = e / DP, q = e % DP;
            float v = 0.f;
            if (ino == -1) v = (p == 0 && q == 0) ? 1.f : 0.f;
            else if (ino >= 0 && p < Dm && q < Dm) v = mold[ino + p * Dm + q];
            dst[e] = v;
        }
    }
    __syncthreads();

    float *pq = part + quad * 4 * M2;
    // an in-lane open's partial sums, summed over the quad (port 0) and stored by every lane
    // of the quad (the same values)
    auto store = [&](float (&o)[M2], int prt) {
#pragma unroll
        for (int e = 0; e < M2; ++e) pq[prt * M2 + e] = quad_sum(o[e]);
    };
    if (!live) {  // uniform over the quad; an idle or stopped slice contributes zeros
#pragma unroll
        for (int e = 0; e < 4 * M2; ++e) pq[e] = 0.f;
    } else {
        const float *m0 = msm + (vi * 4 + 0) * M2, *m1 = msm + (vi * 4 + 1) * M2;
        const float *m2 = msm + (vi * 4 + 2) * M2, *m3 = msm + (vi * 4 + 3) * M2;
        float c0[DP], c1[DP], c2[DP], c3[DP];  // m_k[p][this lane]
#pragma unroll
        for (int p = 0; p < DP; ++p) {
            c0[p] = j < DP ? m0[p * DP + j] : 0.f;
            c1[p] = j < DP ? m1[p * DP + j] : 0.f;
            c2[p] = j < DP ? m2[p * DP + j] : 0.f;
            c3[p] = j < DP ? m3[p * DP + j] : 0.f;
        }
        float E[RP], o[M2];

        // ---- phase A: E = A x0 m0 x1 m1
#pragma unroll
        for (int r = 0; r < RP; ++r) E[r] = A[r];
        vs_mode_0<DP>(E, c0);
        vs_mode_in<DP, 1>(E, c1);
        // out2[a][a'] = sum_{p1,b3} G[p1,a,b3] A[p1,a',b3], G = E x3 m3 block by block (p1)
#pragma unroll
        for (int e = 0; e < M2; ++e) o[e] = 0.f;
#pragma unroll
        for (int p1 = 0; p1 < DP; ++p1) {
            float G[M2];  // [a = p2][b3]
#pragma unroll
            for (int x = 0; x < DP; ++x) {
                v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int p = 0; p < DP; ++p)
                    acc = __builtin_amdgcn_mfma_f32_4x4x1f32(c3[p], E[(p1 * DP + x) * DP + p], acc, 0, 0, 0);
#pragma unroll
                for (int y = 0; y < DP; ++y) G[x * DP + y] = acc[y];
            }
#pragma unroll
            for (int a = 0; a < DP; ++a)
#pragma unroll
                for (int ap = 0; ap < DP; ++ap)
#pragma unroll
                    for (int t = 0; t < DP; ++t)
                        o[a * DP + ap] = fmaf(G[a * DP + t], A[(p1 * DP + ap) * DP + t], o[a * DP + ap]);
        }
        store(o, 2);
        // out3[a][a'] = sum_{p1,b2} G[p1,b2,a] A[p1,b2,a'], G = E x2 m2 block by block (p1)
#pragma unroll
        for (int e = 0; e < M2; ++e) o[e] = 0.f;
#pragma unroll
        for (int p1 = 0; p1 < DP; ++p1) {
            float G[M2];  // [b2][a = p3]
#pragma unroll
            for (int y = 0; y < DP; ++y) {
                v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
                for (int p = 0; p < DP; ++p)
                    acc = __builtin_amdgcn_mfma_f32_4x4x1f32(c2[p], E[(p1 * DP + p) * DP + y], acc, 0, 0, 0);
#pragma unroll
                for (int x = 0; x < DP; ++x) G[x * DP + y] = acc[x];
            }
#pragma unroll
            for (int a = 0; a < DP; ++a)
#pragma unroll
                for (int ap = 0; ap < DP; ++ap)
#pragma unroll
                    for (int t = 0; t < DP; ++t)
                        o[a * DP + ap] = fmaf(G[t * DP + a], A[(p1 * DP + t) * DP + ap], o[a * DP + ap]);
        }
        store(o, 3);

        // ---- phase B: E = A x2 m2 x3 m3
#pragma unroll
        for (int r = 0; r < RP; ++r) E[r] = A[r];
        vs_mode_in<DP, 2>(E, c2);
        vs_mode_in<DP, 3>(E, c3);
        // out1 = open_1(E x0 m0, A), the port-0 product fused entry by entry
#pragma unroll
        for (int e = 0; e < M2; ++e) o[e] = 0.f;
#pragma unroll
        for (int a = 0; a < DP; ++a)
#pragma unroll
            for (int t = 0; t < DP * DP; ++t) {
                const v4f gx = quad_gather(E[a * DP * DP + t]);
                float gv = 0.f;
#pragma unroll
                for (int p = 0; p < DP; ++p) gv = fmaf(gx[p], c0[p], gv);
#pragma unroll
                for (int ap = 0; ap < DP; ++ap) o[a * DP + ap] = fmaf(gv, A[ap * DP * DP + t], o[a * DP + ap]);
            }
        store(o, 1);
        // out0 = open_0(E x1 m1, A) on the matrix cores: with a = G (row a = lane i) and
        // b = A (column a' = lane j), register i of lane j collects out0[i][j]
        vs_mode_in<DP, 1>(E, c1);
        v4f o0 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int r = 0; r < R; ++r) o0 = __builtin_amdgcn_mfma_f32_4x4x1f32(E[r], A[r], o0, 0, 0, 0);
        if (j < DP) {
#pragma unroll
            for (int i = 0; i < DP; ++i) pq[i * DP + j] = o0[i];
        }
    }
    __syncthreads();

    // ---- one lane per message: sum over slices, normalise, residuals, write
    if (qout < 0) return;
    if (done2) {
        for (int e = 0; e < Dm * Dm; ++e) mnew[outo + e] = mold[outo + e];
        res_fid[qout] = 0.0;
        res_abs[qout] = 0.0;
        return;
    }
    float x[M2];
    double S = 0.0;
#pragma unroll
    for (int e = 0; e < M2; ++e) {
        float v = 0.f;
        for (int s = 0; s < d; ++s) v += part[((vi2 * d + s) * 4 + port) * M2 + e];
        x[e] = v;
        S += (double)v;  // entries beyond the bond dimension are exactly zero
    }
    double dn = 0.0, nn = 0.0, oo = 0.0, mx = 0.0;
#pragma unroll
    for (int e = 0; e < M2; ++e) {
        const int a = e / DP, ap = e % DP;
        if (a >= Dm || ap >= Dm) continue;
        const float xf = (float)(S != 0.0 ? (double)x[e] / S : (double)x[e]);
        const long long o = outo + a * Dm + ap;
        const double xo = (double)mold[o];
        mnew[o] = xf;
        dn += (double)xf * xo;
        nn += (double)xf * (double)xf;
        oo += xo * xo;
        const double df = fabs((double)xf - xo);
        mx = (df > mx || df != df) ? df : mx;
    }
    const double c = dn / sqrt(nn * oo);
    res_fid[qout] = 1.0 - c * c;
    res_abs[qout] = mx;
}

// ---------------------------------------------------------------- bond dimension <= 2
//
// At DP = 2 a slice is 16 values: one lane holds a whole slice A_u[s] (register
// r = p0*8 + p1*4 + p2*2 + p3) and every product is in-lane, so a wave holds W2 = 64 / d
// vertices (lane = vi * d + s).  The d slices of a vertex meet in LDS; lane s of the vertex
// then finishes the messages of ports j with j % d == s.  Site layout per group:
// float[group][4][64 lanes][4] (four float4 loads per lane, each a contiguous 1 KiB line).

__device__ __forceinline__ int t2_idx(int p0, int p1, int p2, int p3) {
    return p0 * 8 + p1 * 4 + p2 * 2 + p3;
}

// X[.., b at port P, ..] = sum_p X[.., p, ..] m[p][b], in place (m: 2 x 2, [p][b])
template <int P>
__device__ __forceinline__ void t2_mode(float (&X)[16], const float (&m)[4]) {
    constexpr int st = 8 >> P;  // register stride of port P
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        if (r & st) continue;  // r has digit 0 at port P
        const float x0 = X[r], x1 = X[r + st];
        X[r] = fmaf(x0, m[0], x1 * m[2]);
        X[r + st] = fmaf(x0, m[1], x1 * m[3]);
    }
}

// out[a][a'] = sum over the other ports of G[.., a at P, ..] A[.., a' at P, ..]
template <int P>
__device__ __forceinline__ void t2_open(const float (&G)[16], const float (&A)[16], float (&out)[4]) {
    constexpr int st = 8 >> P;
#pragma unroll
    for (int e = 0; e < 4; ++e) out[e] = 0.f;
#pragma unroll
    for (int r = 0; r < 16; ++r) {
        if (r & st) continue;
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int ap = 0; ap < 2; ++ap)
                out[a * 2 + ap] = fmaf(G[r + a * st], A[r + ap * st], out[a * 2 + ap]);
    }
}

// One message of the bond-dimension <= 2 bucket: 2 x 2 zero-padded (a padded port gives
// e0 e0^T); a full 2 x 2 message at a 16-byte aligned offset is one float4 access.
__device__ __forceinline__ float4 t2_load_msg(const float *m, long long off, int D) {
    if (off == -1) return make_float4(1.f, 0.f, 0.f, 0.f);
    if (D == 2 && (off & 3) == 0) return *reinterpret_cast<const float4 *>(m + off);
    if (D == 2) return make_float4(m[off], m[off + 1], m[off + 2], m[off + 3]);
    return make_float4(D == 1 ? m[off] : 0.f, 0.f, 0.f, 0.f);
}

__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_vertex_t2(
    const float *__restrict__ vs_sites, const float *__restrict__ mold, float *__restrict__ mnew,
    const int *__restrict__ vs_ino, const int *__restrict__ vs_outo,
    const int *__restrict__ vs_out, const int *__restrict__ vs_dim, const int *__restrict__ vnet,
    NetCtl ctl, double *__restrict__ res_fid, double *__restrict__ res_abs, int nitems, int d) {
    // 4 KiB of LDS: the incoming messages [vi * 4 + port], then, once every lane holds its
    // four, the per-slice results [lane * 4 + port], so that eight one-wave workgroups share
    // a SIMD (64 registers).
    __shared__ float4 msm[256];
    float4 *part = msm;
    const int g = blockIdx.x, lane = threadIdx.x;
    const int W2 = 64 / d, nm = W2 * 4;  // vertices and messages of this group
    const int vi = lane / d, s = lane % d, b = g * W2 + vi;
    const bool inr = vi < W2 && b < nitems;
    float A[16];
    if (inr) {
        const float4 *src = reinterpret_cast<const float4 *>(vs_sites) + (size_t)g * 4 * 64;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const float4 v = src[k * 64 + lane];
            A[4 * k] = v.x;
            A[4 * k + 1] = v.y;
            A[4 * k + 2] = v.z;
            A[4 * k + 3] = v.w;
        }
    }
    // incoming messages: one lane per message, 16-byte loads
    for (int mi = lane; mi < nm; mi += 64) {
        const int bb = g * W2 + mi / 4;
        float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
        if (bb < nitems)
            v = t2_load_msg(mold, vs_ino[4 * bb + (mi & 3)], (vs_dim[bb] >> (8 * (mi & 3))) & 255);
        msm[mi] = v;
    }
    wave_lds_sync();  // one wave per workgroup
    const bool live = inr && !(ctl.enabled && ctl.done[vnet[b]]);
    float m[4][4];  // vi * 4 + j < 256 for every lane
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const float4 v = msm[vi * 4 + j];
        m[j][0] = v.x;
        m[j][1] = v.y;
        m[j][2] = v.z;
        m[j][3] = v.w;
    }
    wave_lds_sync();  // every lane holds its messages: the buffer takes the results
    // each result goes to LDS as soon as it is made (four registers live, not sixteen)
    auto put = [&](int j, const float (&o)[4]) {
        part[lane * 4 + j] = make_float4(o[0], o[1], o[2], o[3]);
    };
    if (live) {
        float E[16], G[16], o[4];
        // phase A: E = A x0 m0 x1 m1; out2 = open_2(E x3 m3, A), out3 = open_3(E x2 m2, A)
#pragma unroll
        for (int r = 0; r < 16; ++r) E[r] = A[r];
        t2_mode<0>(E, m[0]);
        t2_mode<1>(E, m[1]);
#pragma unroll
        for (int r = 0; r < 16; ++r) G[r] = E[r];
        t2_mode<3>(G, m[3]);
        t2_open<2>(G, A, o);
        put(2, o);
        t2_mode<2>(E, m[2]);
        t2_open<3>(E, A, o);
        put(3, o);
        // phase B: E = A x2 m2 x3 m3; out0 = open_0(E x1 m1, A), out1 = open_1(E x0 m0, A)
#pragma unroll
        for (int r = 0; r < 16; ++r) E[r] = A[r];
        t2_mode<2>(E, m[2]);
        t2_mode<3>(E, m[3]);
#pragma unroll
        for (int r = 0; r < 16; ++r) G[r] = E[r];
        t2_mode<1>(G, m[1]);
        t2_open<0>(G, A, o);
        put(0, o);
        t2_mode<0>(E, m[0]);
        t2_open<1>(E, A, o);
        put(1, o);
    } else {
        const float z[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int j = 0; j < 4; ++j) put(j, z);
    }
    wave_lds_sync();
    // one lane per outgoing message: sum over the vertex's slices, normalise, residuals
    for (int mi = lane; mi < nm; mi += 64) {
        const int v2 = mi / 4, j = mi & 3, bb = g * W2 + v2;
        if (bb >= nitems) continue;
        const int qout = vs_out[4 * bb + j];
        if (qout < 0) continue;
        const int D = (vs_dim[bb] >> (8 * j)) & 255;
        const long long mo = vs_outo[4 * bb + j];
        const float4 xo4 = t2_load_msg(mold, mo, D);
        const float xo[4] = {xo4.x, xo4.y, xo4.z, xo4.w};
        if (ctl.enabled && ctl.done[vnet[bb]]) {
            for (int e = 0; e < D * D; ++e) mnew[mo + e] = xo[e];
            res_fid[qout] = 0.0;
            res_abs[qout] = 0.0;
            continue;
        }
        float x[4] = {0.f, 0.f, 0.f, 0.f};
        for (int t = 0; t < d; ++t) {
            const float4 pv = part[(v2 * d + t) * 4 + j];
            x[0] += pv.x;
            x[1] += pv.y;
            x[2] += pv.z;
            x[3] += pv.w;
        }
        const double S = (double)x[0] + (double)x[1] + (double)x[2] + (double)x[3];
        float xf[4];
        double dn = 0.0, nn = 0.0, oo = 0.0, mx = 0.0;
#pragma unroll
        for (int e = 0; e < 4; ++e) {
            // entries beyond the bond dimension are exactly zero; the message is stored D x D
            const int a = e >> 1, ap = e & 1;
            xf[e] = (float)(S != 0.0 ? (double)x[e] / S : (double)x[e]);
            if (a >= D || ap >= D) continue;
            const double xv = (double)xf[e], xw = (double)xo[a * D + ap];
            dn += xv * xw;
            nn += xv * xv;
            oo += xw * xw;
            const double df = fabs(xv - xw);
            mx = (df > mx || df != df) ? df : mx;
        }
        if (D == 2 && (mo & 3) == 0) {
            *reinterpret_cast<float4 *>(mnew + mo) = make_float4(xf[0], xf[1], xf[2], xf[3]);
        } else if (D == 2) {
            for (int e = 0; e < 4; ++e) mnew[mo + e] = xf[e];
        } else {
            mnew[mo] = xf[0];
        }
        const double c = dn / sqrt(nn * oo);
        res_fid[qout] = 1.0 - c * c;
        res_abs[qout] = mx;
    }
}

// Sites of the bond-dimension <= 2 bucket in the lane-per-slice layout.
__global__ void k_t2_pad_sites(const double *__restrict__ sites, const long long *__restrict__ site_off,
                               const int *__restrict__ port_off, const int *__restrict__ port_edge,
                               const int *__restrict__ edge_dim, const int *__restrict__ vlist,
                               int nitems, int d, int ngroup, float *__restrict__ out) {
    const int W2 = 64 / d;
    const long long total = (long long)ngroup * 16 * 64;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(((i / 256) % 4) * 4 + (i % 4));
        const int lane = (int)((i / 4) % 64);
        const int g = (int)(i / 1024);
        const int vi = lane / d, s = lane % d, b = g * W2 + vi;
        float v = 0.f;
        if (vi < W2 && b < nitems) {
            const int u = vlist[b];
            const int x[4] = {(r >> 3) & 1, (r >> 2) & 1, (r >> 1) & 1, r & 1};
            const int k = port_off[u + 1] - port_off[u];
            long long idx = s;
            bool ok = true;
            for (int t = 0; t < 4; ++t) {
                const int D = t < k ? edge_dim[port_edge[port_off[u] + t]] : 1;
                if (x[t] >= D) { ok = false; break; }
                if (t < k) idx = idx * D + x[t];
            }
            if (ok) v = (float)sites[site_off[u] + idx];
        }
        out[i] = v;
    }
}

// Pad the sites of one DP bucket into its layout (grid-stride over every float).
template <int DP>
__global__ void k_vs_pad_sites(const double *__restrict__ sites, const long long *__restrict__ site_off,
                               const int *__restrict__ port_off, const int *__restrict__ port_edge,
                               const int *__restrict__ edge_dim, const int *__restrict__ vlist,
                               int nitems, int d, int ngroup, float *__restrict__ out) {
    using Shp = VsShape<DP>;
    const int W = 16 / d;
    const long long total = (long long)ngroup * Shp::RP * 64;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int r = (int)(((i / 256) % (Shp::RP / 4)) * 4 + (i % 4));
        const int lane = (int)((i / 4) % 64);
        const int g = (int)(i / ((long long)Shp::RP * 64));
        const int quad = lane >> 2, p0 = lane & 3;
        const int vi = quad / d, s = quad % d, b = g * W + vi;
        float v = 0.f;
        if (vi < W && b < nitems && r < Shp::R && p0 < DP) {
            const int u = vlist[b];
            const int x[4] = {p0, r / (DP * DP), (r / DP) % DP, r % DP};
            const int k = port_off[u + 1] - port_off[u];
            long long idx = s;
            bool ok = true;
            for (int t = 0; t < 4; ++t) {
                const int D = t < k ? edge_dim[port_edge[port_off[u] + t]] : 1;
                if (x[t] >= D) { ok = false; break; }
                if (t < k) idx = idx * D + x[t];
            }
            if (ok) v = (float)sites[site_off[u] + idx];
        }
        out[i] = v;
    }
}

static int vertex_s_groups(const VertexPlan *p, int k) {
    const int W = k == 0 ? 64 / p->d : 16 / p->d;  // DP = 2: a slice per lane
    return (p->ns[k] + W - 1) / W;
}

static int vertex_s_upload_sites(VertexPlan *p, hipStream_t st, const double *d_sites64,
                                 const long long *d_site_off, const int *d_port_off,
                                 const int *d_port_edge, const int *d_edge_dim) {
    for (int k = 0; k < 3; ++k) {
        if (p->d_vs_sites[k]) (void)hipFree(p->d_vs_sites[k]);
        p->d_vs_sites[k] = nullptr;
        if (p->ns[k] == 0) continue;
        const int DP = k + 2, RP = DP == 2 ? 16 : (DP * DP * DP + 3) & ~3;
        const int ng = vertex_s_groups(p, k);
        if (hipMalloc((void **)&p->d_vs_sites[k], sizeof(float) * (size_t)ng * RP * 64) != hipSuccess)
            return 5;
        const long long total = (long long)ng * RP * 64;
        const int blocks = (int)std::min<long long>((total + 255) / 256, 4096);
        if (DP == 2) {
            hipLaunchKernelGGL(k_t2_pad_sites, dim3(std::max(blocks, 1)), dim3(256), 0, st,
                               d_sites64, d_site_off, d_port_off, d_port_edge, d_edge_dim,
                               p->d_vs_list[k], p->ns[k], p->d, ng, p->d_vs_sites[k]);
            if (hipGetLastError() != hipSuccess) return 2;
            continue;
        }
#define VS_PAD(DPV)                                                                              \
    hipLaunchKernelGGL((k_vs_pad_sites<DPV>), dim3(std::max(blocks, 1)), dim3(256), 0, st,         \
                       d_sites64, d_site_off, d_port_off, d_port_edge, d_edge_dim,               \
                       p->d_vs_list[k], p->ns[k], p->d, ng, p->d_vs_sites[k])
        if (DP == 3) VS_PAD(3);
        else VS_PAD(4);
#undef VS_PAD
        if (hipGetLastError() != hipSuccess) return 2;
    }
    return 0;
}

static int vertex_s_sweep(const VertexPlan *p, hipStream_t st, DevTopo tp, const void *src,
                          void *dst, const NetCtl &ctl, double *res_fid, double *res_abs) {
    for (int k = 0; k < 3; ++k) {
        if (p->ns[k] == 0) continue;
        const int ng = vertex_s_groups(p, k);
#define VS_GO(DPV)                                                                               \
    hipLaunchKernelGGL((k_vertex_s<DPV>), dim3(ng), dim3(64), 0, st, p->d_vs_sites[k],             \
                       (const float *)src, (float *)dst, p->d_vs_ino[k], p->d_vs_outo[k],        \
                       p->d_vs_out[k], p->d_vs_dim[k], p->d_vs_net[k], ctl, res_fid, res_abs,    \
                       p->ns[k], p->d)
        if (k == 0)
            hipLaunchKernelGGL(k_vertex_t2, dim3(ng), dim3(64), 0, st, p->d_vs_sites[k],
                               (const float *)src, (float *)dst, p->d_vs_ino[k], p->d_vs_outo[k],
                               p->d_vs_out[k], p->d_vs_dim[k], p->d_vs_net[k], ctl, res_fid,
                               res_abs, p->ns[k], p->d);
        else if (k == 1) VS_GO(3);
        else VS_GO(4);
#undef VS_GO
        if (hipGetLastError() != hipSuccess) return 2;
    }
    return 0;
}
