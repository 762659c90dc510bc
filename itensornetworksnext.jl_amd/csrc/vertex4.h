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

__device__ __forceinline__ float quad_bcast(float v, int p) {
    // quad_perm(p, p, p, p): every lane of the quad reads lane p of the quad
    switch (p) {
        case 0: return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x00, 0xF, 0xF, false));
        case 1: return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x55, 0xF, 0xF, false));
        case 2: return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xAA, 0xF, 0xF, false));
        default: return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xFF, 0xF, 0xF, false));
    }
}

// sum over the 4 lanes of the quad (quad_perm xor 1, then xor 2): every lane gets it
__device__ __forceinline__ float quad_sum(float v) {
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
    v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, false));
    return v;
}

// In-lane mode product on port P (1, 2 or 3), in place: X[.., b, ..] = sum_p X[.., p, ..] m[p][b]
template <int DP, int P>
__device__ __forceinline__ void vs_mode_in(float (&X)[VsShape<DP>::RP], const float *m) {
    float mm[DP * DP];
#pragma unroll
    for (int e = 0; e < DP * DP; ++e) mm[e] = m[e];
    constexpr int st = P == 1 ? DP * DP : P == 2 ? DP : 1;  // register stride of port P
#pragma unroll
    for (int o = 0; o < DP * DP; ++o) {
        // the other two in-lane digits: o = (hi, lo) in the order of the remaining ports
        const int hi = o / DP, lo = o % DP;
        const int base = P == 1 ? hi * DP + lo : P == 2 ? hi * DP * DP + lo : (hi * DP + lo) * DP;
        float v[DP];
#pragma unroll
        for (int p = 0; p < DP; ++p) v[p] = X[base + p * st];
#pragma unroll
        for (int b = 0; b < DP; ++b) {
            float acc = 0.f;
#pragma unroll
            for (int p = 0; p < DP; ++p) acc = fmaf(v[p], mm[p * DP + b], acc);
            X[base + b * st] = acc;
        }
    }
}

// Cross-lane mode product on port 0, in place: lane b of the quad gets
// X_b[r] = sum_p X_p[r] m0[p][b] (c[p] = m0[p][this lane's b], zero for lanes >= DP).
template <int DP>
__device__ __forceinline__ void vs_mode_0(float (&X)[VsShape<DP>::RP], const float (&c)[DP]) {
#pragma unroll
    for (int r = 0; r < VsShape<DP>::R; ++r) {
        const float x = X[r];
        float acc = 0.f;
#pragma unroll
        for (int p = 0; p < DP; ++p) acc = fmaf(quad_bcast(x, p), c[p], acc);
        X[r] = acc;
    }
}

template <int DP>
__global__ __launch_bounds__(64) void k_vertex_s(
    const float *__restrict__ vs_sites, const float *__restrict__ mold, float *__restrict__ mnew,
    const int4 *__restrict__ vin, const int4 *__restrict__ vout, const int4 *__restrict__ vdim,
    const int *__restrict__ vnet, const long long *__restrict__ msg_off, NetCtl ctl,
    double *__restrict__ res_fid, double *__restrict__ res_abs, int nitems, int d) {
    using Shp = VsShape<DP>;
    constexpr int R = Shp::R, RP = Shp::RP, M2 = Shp::M2;
    __shared__ float msm[16 * 4 * M2];   // [vi][port][p * DP + b], padded incoming messages
    __shared__ float part[16 * 4 * M2];  // [quad][port][a * DP + a'], per-slice results
    const int g = blockIdx.x, lane = threadIdx.x;
    const int quad = lane >> 2, j = lane & 3;
    const int W = 16 / d;

    // incoming messages: lane (vi2, port) loads one, zero-padded to DP x DP
    {
        const int vi2 = lane >> 2, port = lane & 3, b2 = g * W + vi2;
        float *dst = msm + (vi2 * 4 + port) * M2;
        int q = -2, D = 0;
        if (vi2 < W && b2 < nitems) {
            const int4 i4 = vin[b2], d4 = vdim[b2];
            q = port == 0 ? i4.x : port == 1 ? i4.y : port == 2 ? i4.z : i4.w;
            D = port == 0 ? d4.x : port == 1 ? d4.y : port == 2 ? d4.z : d4.w;
        }
        const long long o = q >= 0 ? msg_off[q] : 0;
#pragma unroll
        for (int e = 0; e < M2; ++e) {
            const int p = e / DP, b = e % DP;
            float v = 0.f;
            if (q == -1) v = (p == 0 && b == 0) ? 1.f : 0.f;  // padded port
            else if (q >= 0 && p < D && b < D) v = mold[o + p * D + b];
            dst[e] = v;
        }
    }
    __syncthreads();

    const int vi = quad / d, b = g * W + vi;
    const bool live = vi < W && b < nitems && !(ctl.enabled && ctl.done[vnet[b]]);
    float *pq = part + quad * 4 * M2;
    // one open's in-lane partial sums: summed over the quad (port 0), stored to part
    auto store = [&](float (&o)[M2], int port) {
#pragma unroll
        for (int e = 0; e < M2; ++e) o[e] = quad_sum(o[e]);
#pragma unroll
        for (int e = 0; e < M2; ++e)
            if ((e & 3) == j) pq[port * M2 + e] = o[e];
    };
    float o[M2];
    if (!live) {  // uniform over the quad; an idle or stopped slice contributes zeros
#pragma unroll
        for (int e = 0; e < M2; ++e) o[e] = 0.f;
        store(o, 0);
        store(o, 1);
        store(o, 2);
        store(o, 3);
    } else {  // uniform over the quad (DPP reads stay inside it)
        const float *m0 = msm + (vi * 4 + 0) * M2, *m1 = msm + (vi * 4 + 1) * M2;
        const float *m2 = msm + (vi * 4 + 2) * M2, *m3 = msm + (vi * 4 + 3) * M2;
        float A[RP], E[RP];
        const float4 *src = reinterpret_cast<const float4 *>(vs_sites) + (size_t)g * (RP / 4) * 64;
#pragma unroll
        for (int k = 0; k < RP / 4; ++k) {
            const float4 v = src[k * 64 + lane];
            A[4 * k] = v.x;
            A[4 * k + 1] = v.y;
            A[4 * k + 2] = v.z;
            A[4 * k + 3] = v.w;
        }
        float c0[DP];  // m0[p][this lane] for the port-0 products
#pragma unroll
        for (int p = 0; p < DP; ++p) c0[p] = j < DP ? m0[p * DP + j] : 0.f;

        // ---- phase A: E = A x0 m0 x1 m1
#pragma unroll
        for (int r = 0; r < RP; ++r) E[r] = A[r];
        vs_mode_0<DP>(E, c0);
        vs_mode_in<DP, 1>(E, m1);
        // out2[a][a'] = sum_{p1,b3} (sum_p3 E[p1,a,p3] m3[p3][b3]) A[p1,a',b3]
#pragma unroll
        for (int e = 0; e < M2; ++e) o[e] = 0.f;
#pragma unroll
        for (int p1 = 0; p1 < DP; ++p1) {
            float G[M2];  // [a = p2][b3]
#pragma unroll
            for (int x = 0; x < DP; ++x)
#pragma unroll
                for (int y = 0; y < DP; ++y) {
                    float t = 0.f;
#pragma unroll
                    for (int p = 0; p < DP; ++p) t = fmaf(E[(p1 * DP + x) * DP + p], m3[p * DP + y], t);
                    G[x * DP + y] = t;
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
        // out3[a][a'] = sum_{p1,b2} (sum_p2 E[p1,p2,a] m2[p2][b2]) A[p1,b2,a']
#pragma unroll
        for (int e = 0; e < M2; ++e) o[e] = 0.f;
#pragma unroll
        for (int p1 = 0; p1 < DP; ++p1) {
            float G[M2];  // [b2][a = p3]
#pragma unroll
            for (int x = 0; x < DP; ++x)
#pragma unroll
                for (int y = 0; y < DP; ++y) {
                    float t = 0.f;
#pragma unroll
                    for (int p = 0; p < DP; ++p) t = fmaf(E[(p1 * DP + p) * DP + y], m2[p * DP + x], t);
                    G[x * DP + y] = t;
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
        vs_mode_in<DP, 2>(E, m2);
        vs_mode_in<DP, 3>(E, m3);
        // out1 = open_1(E x0 m0, A), the port-0 product fused entry by entry
#pragma unroll
        for (int e = 0; e < M2; ++e) o[e] = 0.f;
#pragma unroll
        for (int a = 0; a < DP; ++a)
#pragma unroll
            for (int t = 0; t < DP * DP; ++t) {
                const float x = E[a * DP * DP + t];
                float gv = 0.f;
#pragma unroll
                for (int p = 0; p < DP; ++p) gv = fmaf(quad_bcast(x, p), c0[p], gv);
#pragma unroll
                for (int ap = 0; ap < DP; ++ap) o[a * DP + ap] = fmaf(gv, A[ap * DP * DP + t], o[a * DP + ap]);
            }
        store(o, 1);
        // out0 = open_0(E x1 m1, A): this lane holds row a = j of out0
        vs_mode_in<DP, 1>(E, m1);
        float o0[DP];
#pragma unroll
        for (int ap = 0; ap < DP; ++ap) o0[ap] = 0.f;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const float x = E[r], y = A[r];
#pragma unroll
            for (int ap = 0; ap < DP; ++ap) o0[ap] = fmaf(x, quad_bcast(y, ap), o0[ap]);
        }
        if (j < DP) {
#pragma unroll
            for (int ap = 0; ap < DP; ++ap) pq[j * DP + ap] = o0[ap];
        }
    }
    __syncthreads();

    // ---- one lane per message: sum over slices, normalise, residuals, write
    const int vi2 = lane >> 2, port = lane & 3, b2 = g * W + vi2;
    if (vi2 >= W || b2 >= nitems) return;
    const int4 o4 = vout[b2], d4 = vdim[b2];
    const int q = port == 0 ? o4.x : port == 1 ? o4.y : port == 2 ? o4.z : o4.w;
    if (q < 0) return;
    const int D = port == 0 ? d4.x : port == 1 ? d4.y : port == 2 ? d4.z : d4.w;
    const long long mo = msg_off[q];
    if (ctl.enabled && ctl.done[vnet[b2]]) {
        for (int e = 0; e < D * D; ++e) mnew[mo + e] = mold[mo + e];
        res_fid[q] = 0.0;
        res_abs[q] = 0.0;
        return;
    }
    float x[M2];
    double S = 0.0;
#pragma unroll
    for (int e = 0; e < M2; ++e) {
        float v = 0.f;
        for (int s = 0; s < d; ++s) v += part[((vi2 * d + s) * 4 + port) * M2 + e];
        x[e] = v;
        if (e / DP < D && e % DP < D) S += (double)v;
    }
    double dn = 0.0, nn = 0.0, oo = 0.0, mx = 0.0;
#pragma unroll
    for (int e = 0; e < M2; ++e) {
        const int a = e / DP, ap = e % DP;
        if (a >= D || ap >= D) continue;
        const float xf = (float)(S != 0.0 ? (double)x[e] / S : (double)x[e]);
        const long long o = mo + a * D + ap;
        const double xo = (double)mold[o];
        mnew[o] = xf;
        dn += (double)xf * xo;
        nn += (double)xf * (double)xf;
        oo += xo * xo;
        const double df = fabs((double)xf - xo);
        mx = (df > mx || df != df) ? df : mx;
    }
    const double c = dn / sqrt(nn * oo);
    res_fid[q] = 1.0 - c * c;
    res_abs[q] = mx;
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
    const int W = 16 / p->d;
    return (p->ns[k] + W - 1) / W;
}

static int vertex_s_upload_sites(VertexPlan *p, hipStream_t st, const double *d_sites64,
                                 const long long *d_site_off, const int *d_port_off,
                                 const int *d_port_edge, const int *d_edge_dim) {
    for (int k = 0; k < 3; ++k) {
        if (p->d_vs_sites[k]) (void)hipFree(p->d_vs_sites[k]);
        p->d_vs_sites[k] = nullptr;
        if (p->ns[k] == 0) continue;
        const int DP = k + 2, RP = (DP * DP * DP + 3) & ~3;
        const int ng = vertex_s_groups(p, k);
        if (hipMalloc((void **)&p->d_vs_sites[k], sizeof(float) * (size_t)ng * RP * 64) != hipSuccess)
            return 5;
        const long long total = (long long)ng * RP * 64;
        const int blocks = (int)std::min<long long>((total + 255) / 256, 4096);
#define VS_PAD(DPV)                                                                              \
    hipLaunchKernelGGL((k_vs_pad_sites<DPV>), dim3(std::max(blocks, 1)), dim3(256), 0, st,         \
                       d_sites64, d_site_off, d_port_off, d_port_edge, d_edge_dim,               \
                       p->d_vs_list[k], p->ns[k], p->d, ng, p->d_vs_sites[k])
        if (DP == 2) VS_PAD(2);
        else if (DP == 3) VS_PAD(3);
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
                       (const float *)src, (float *)dst, p->d_vs_in[k], p->d_vs_out[k],          \
                       p->d_vs_dim[k], p->d_vs_net[k], tp.msg_off, ctl, res_fid, res_abs,        \
                       p->ns[k], p->d)
        if (k == 0) VS_GO(2);
        else if (k == 1) VS_GO(3);
        else VS_GO(4);
#undef VS_GO
        if (hipGetLastError() != hipSuccess) return 2;
    }
    return 0;
}
