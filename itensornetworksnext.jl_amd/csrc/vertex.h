// vertex.h -- vertex-centric BP sweep for bond dimension 8 on gfx950 (FP32, MFMA).
//
// One workgroup per vertex u computes all of u's outgoing messages (the reference
// computes each message separately, beliefpropagation.jl:242-257; the result is the
// same contraction).  One wave per physical index value s: the wave holds the slice
// A_u[s] (4 ports of dimension 8 = 4096 values, 64 per lane) in registers in one of two
// layouts
//     X: register r = (p0, p1), lane l = (p2, p3)      Y: register r = (p2, p3), lane = (p0, p1)
// A ket-side mode product on a register port ("T x_j m_j") runs on
// v_mfma_f32_4x4x1_16b_f32: register i of lane 4k+j accumulates A(lane 4k+i) * B(lane 4k+j),
// so with A = m_j[a][4h + lane%4] (the same for every block) and B = the lane's own
// tensor column, each lane gets its output column without any data movement.
// Partial products are shared between the four outgoing messages:
//     E01 = A x0 m0 x1 m1 (layout X)  -> transposed to Y -> out2, out3
//     E23 = A x2 m2 x3 m3 (layout Y)  -> transposed to X -> out0, out1
// (8 mode products + 4 open contractions per slice instead of 16).  The open contraction
// out[a][a'] = sum E~[a..] A[a'..] runs per lane on the VALU (an 8x8x8 product) and is
// reduced across the wave by a halving exchange, so lane l ends with entry l = (a, a').
// Vertices of degree < 4 are padded with dimension-8 ports whose tensor entries vanish
// beyond index 0 and whose message is the identity; the result is unchanged.
// The two LDS buffers per wave hold A (transposable, row pitch 65 floats: conflict-free
// row and column access) and the transposed partial product.
#pragma once

#include <string>
#include <vector>

#include "common.h"

typedef float v4f __attribute__((ext_vector_type(4)));

struct VertexPlan {
    std::string why;
    int d = 0;                  // physical dimension (waves per workgroup)
    int nv = 0;
    int4 *d_vin = nullptr;      // per vertex: incoming message slot per padded port (-1: dummy)
    int4 *d_vout = nullptr;     // per vertex: outgoing message slot per padded port (-1: dummy)
    int *d_vnet = nullptr;      // network of each vertex
    size_t lds_bytes = 0;
};

static constexpr int VX_D = 8;
static constexpr int VX_PITCH = 65;                 // LDS row pitch (floats)
static constexpr int VX_BUF = 64 * VX_PITCH;        // one transposable 64x64 buffer

static void vertex_plan_release(VertexPlan *p) {
    if (p->d_vin) (void)hipFree(p->d_vin);
    if (p->d_vout) (void)hipFree(p->d_vout);
    if (p->d_vnet) (void)hipFree(p->d_vnet);
    p->d_vin = p->d_vout = nullptr;
    p->d_vnet = nullptr;
}

static int vertex_plan_build(VertexPlan *p, int dtype, int plain, int nv, const int *port_off,
                             const int *port_edge, const int *edge_dim, const int *phys_dim,
                             const long long *site_off, const long long *msg_off,
                             const int *port_rev, const int *vnet) {
    (void)site_off;
    (void)msg_off;
    p->why.clear();
    if (plain) { p->why = "plain (vector-message) networks use the generic path"; return 4; }
    if (dtype != 1) { p->why = "the vertex path computes in FP32"; return 4; }
    const int d = phys_dim[0];
    for (int u = 0; u < nv; ++u) {
        const int k = port_off[u + 1] - port_off[u];
        if (k > 4) { p->why = "degree > 4"; return 4; }
        if (phys_dim[u] != d) { p->why = "physical dimension differs between vertices"; return 4; }
        for (int q = port_off[u]; q < port_off[u + 1]; ++q)
            if (edge_dim[port_edge[q]] != VX_D) { p->why = "bond dimension is not 8"; return 4; }
    }
    if (d < 1 || d > 4) { p->why = "physical dimension outside 1..4"; return 4; }
    std::vector<int4> vin(nv), vout(nv);
    for (int u = 0; u < nv; ++u) {
        int in[4] = {-1, -1, -1, -1}, out[4] = {-1, -1, -1, -1};
        for (int j = 0; j < port_off[u + 1] - port_off[u]; ++j) {
            out[j] = port_off[u] + j;
            in[j] = port_rev[port_off[u] + j];
        }
        vin[u] = make_int4(in[0], in[1], in[2], in[3]);
        vout[u] = make_int4(out[0], out[1], out[2], out[3]);
    }
    if (hipMalloc((void **)&p->d_vin, sizeof(int4) * nv) != hipSuccess ||
        hipMalloc((void **)&p->d_vout, sizeof(int4) * nv) != hipSuccess ||
        hipMalloc((void **)&p->d_vnet, sizeof(int) * nv) != hipSuccess) {
        p->why = "allocation failed";
        vertex_plan_release(p);
        return 5;
    }
    if (hipMemcpy(p->d_vin, vin.data(), sizeof(int4) * nv, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->d_vout, vout.data(), sizeof(int4) * nv, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(p->d_vnet, vnet, sizeof(int) * nv, hipMemcpyHostToDevice) != hipSuccess) {
        p->why = "copy failed";
        vertex_plan_release(p);
        return 2;
    }
    p->d = d;
    p->nv = nv;
    // per wave: two transposable buffers; per workgroup: 4 messages + d x 4 x 64 partials
    p->lds_bytes = sizeof(float) * ((size_t)d * 2 * VX_BUF + 4 * 64 + (size_t)d * 4 * 64);
    return 0;
}

// vsites[u][s][4096]: the site tensor of u zero-padded to 4 ports of dimension 8, row-major
// over (p0, p1, p2, p3) = layout X order (element r * 64 + l).
__global__ void k_vertex_pad_sites(const double *sites64, const long long *site_off,
                                   const int *port_off, int d, int nv, float *vsites) {
    const long long total = (long long)nv * d * 4096;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int e = (int)(i & 4095);
        const long long us = i >> 12;
        const int s = (int)(us % d);
        const int u = (int)(us / d);
        const int k = port_off[u + 1] - port_off[u];
        const int idx[4] = {(e >> 9) & 7, (e >> 6) & 7, (e >> 3) & 7, e & 7};
        bool zero = false;
        long long off = 0;
        for (int j = 0; j < 4; ++j) {
            if (j < k) off = off * 8 + idx[j];
            else if (idx[j] != 0) zero = true;
        }
        long long n = 1;
        for (int j = 0; j < k; ++j) n *= 8;
        vsites[i] = zero ? 0.f : (float)sites64[site_off[u] + s * n + off];
    }
}

static int vertex_upload_sites(VertexPlan *p, hipStream_t st, const double *d_sites64,
                               void **d_vsites, const long long *d_site_off,
                               const int *d_port_off) {
    const size_t bytes = sizeof(float) * (size_t)p->nv * p->d * 4096;
    if (*d_vsites) (void)hipFree(*d_vsites);
    *d_vsites = nullptr;
    if (hipMalloc(d_vsites, bytes) != hipSuccess) return 5;
    hipLaunchKernelGGL(k_vertex_pad_sites, dim3(2048), dim3(256), 0, st, d_sites64, d_site_off,
                       d_port_off, p->d, p->nv, (float *)*d_vsites);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

// ---------------------------------------------------------------- device helpers

// Mode product on a register port of an 8x8 register block T[hi*8 + lo]:
//   HI: O[b*8 + lo] = sum_a T[a*8 + lo] m[a][b]     LO: O[hi*8 + b] = sum_a T[hi*8 + a] m[a][b]
// M[a*2 + h] holds m[a][4h + lane%4] (the A operand of the 16 4x4 blocks).
template <bool HI>
__device__ __forceinline__ void vx_mp(const float (&T)[64], float (&O)[64], const float (&M)[16]) {
#pragma unroll
    for (int x = 0; x < 8; ++x) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int a = 0; a < 8; ++a) {
                const float b = HI ? T[a * 8 + x] : T[x * 8 + a];
                acc = __builtin_amdgcn_mfma_f32_4x4x1f32(M[a * 2 + h], b, acc, 0, 0, 0);
            }
#pragma unroll
            for (int i = 0; i < 4; ++i) {
                const int bb = 4 * h + i;
                if (HI) O[bb * 8 + x] = acc[i]; else O[x * 8 + bb] = acc[i];
            }
        }
    }
}

__device__ __forceinline__ void vx_load_m(const float *msg, float (&M)[16]) {
    const int c = threadIdx.x & 3;
#pragma unroll
    for (int a = 0; a < 8; ++a)
#pragma unroll
        for (int h = 0; h < 2; ++h) M[a * 2 + h] = msg[a * 8 + 4 * h + c];
}

// Sum v[i] over the 64 lanes for every i; lane l returns entry l (halving exchange).
// Step 1 pairs lane l with l^32 through v_permlane32_swap (lanes 32-63 of the first
// operand trade places with lanes 0-31 of the second, so x' + y' holds entry i on the
// low half and entry i+32 on the high half); step 2 does the same for 16-lane rows with
// v_permlane16_swap; the last four steps exchange through ds_bpermute.  Values are
// selected after they are loaded so the array stays in registers.
__device__ __forceinline__ float vx_reduce64(float (&v)[64]) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < 32; ++i) {
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[i]),
                                                        __float_as_uint(v[i + 32]), false, false);
        v[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v[i]),
                                                        __float_as_uint(v[i + 16]), false, false);
        v[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
#pragma unroll
    for (int s = 8; s >= 1; s >>= 1) {
        const bool up = (lane & s) != 0;
#pragma unroll
        for (int i = 0; i < s; ++i) {
            const float lo = v[i], hi = v[i + s];
            const float send = up ? lo : hi;
            const float keep = up ? hi : lo;
            v[i] = keep + __shfl_xor(send, s, 64);
        }
    }
    return v[0];
}

// Open contraction with the bra tensor read from the wave's A buffer.
//   HI (open port = register hi index): out[a][a'] = sum_lo G[a*8+lo] A[a'*8+lo]
//   LO (open port = register lo index): out[a][a'] = sum_hi G[hi*8+a] A[hi*8+a']
// XL: the A buffer is read in layout X (else Y).  Returns entry (lane/8, lane%8).
template <bool HI, bool XL>
__device__ __forceinline__ float vx_final(const float (&G)[64], const float *bufA) {
    const int lane = threadIdx.x & 63;
    float acc[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) acc[i] = 0.f;
#pragma unroll
    for (int r = 0; r < 64; ++r) {
        const float av = XL ? bufA[r * VX_PITCH + lane] : bufA[lane * VX_PITCH + r];
        const int hi = r >> 3, lo = r & 7;
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            if (HI) acc[a * 8 + hi] += G[a * 8 + lo] * av;   // a' = hi
            else acc[a * 8 + lo] += G[hi * 8 + a] * av;      // a' = lo
        }
    }
    return vx_reduce64(acc);
}

// ---------------------------------------------------------------- the sweep kernel

template <int NS>
__global__ __launch_bounds__(64 * NS) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_vertex_d8(const float *__restrict__ vsites, const float *__restrict__ mold,
                 float *__restrict__ mnew, const int4 *__restrict__ vin,
                 const int4 *__restrict__ vout, const int *__restrict__ vnet,
                 NetCtl ctl, double *__restrict__ res_fid, double *__restrict__ res_abs) {
    extern __shared__ __align__(16) float vsm[];
    const int u = blockIdx.x;
    const int w = threadIdx.x >> 6;     // wave = physical index s
    const int lane = threadIdx.x & 63;
    float *bufA = vsm + w * 2 * VX_BUF;
    float *bufB = bufA + VX_BUF;
    float *msm = vsm + NS * 2 * VX_BUF;  // 4 x 64 incoming messages
    float *part = msm + 4 * 64;          // [s][port][64] per-slice results
    const int4 vi = vin[u], vo = vout[u];
    const int ins[4] = {vi.x, vi.y, vi.z, vi.w};
    const int outs[4] = {vo.x, vo.y, vo.z, vo.w};

    const int net = vnet[u];
    if (ctl.enabled && ctl.done[net]) {  // uniform over the workgroup
        for (int j = w; j < 4; j += NS) {
            if (outs[j] < 0) continue;
            mnew[(long long)outs[j] * 64 + lane] = mold[(long long)outs[j] * 64 + lane];
            if (lane == 0) { res_fid[outs[j]] = 0.0; res_abs[outs[j]] = 0.0; }
        }
        return;
    }
    // incoming messages (identity on padded ports)
    for (int i = threadIdx.x; i < 4 * 64; i += 64 * NS) {
        const int j = i >> 6, e = i & 63;
        msm[i] = ins[j] >= 0 ? mold[(long long)ins[j] * 64 + e] : ((e >> 3) == (e & 7) ? 1.f : 0.f);
    }
    // slice A[s] in layout X
    float T[64], E[64], G[64];
    const float *src = vsites + ((long long)u * NS + w) * 4096;
#pragma unroll
    for (int r = 0; r < 64; ++r) T[r] = src[r * 64 + lane];
#pragma unroll
    for (int r = 0; r < 64; ++r) bufA[r * VX_PITCH + lane] = T[r];
    __syncthreads();
    float M[16], out[4];

    // ---- phase A: E01 = A x0 m0 x1 m1 (X) -> Y; out2 (open p2), out3 (open p3)
    vx_load_m(msm + 0 * 64, M);
    vx_mp<true>(T, E, M);
    vx_load_m(msm + 1 * 64, M);
    vx_mp<false>(E, G, M);
#pragma unroll
    for (int r = 0; r < 64; ++r) bufB[r * VX_PITCH + lane] = G[r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 64; ++r) E[r] = bufB[lane * VX_PITCH + r];  // E01 in Y
    vx_load_m(msm + 3 * 64, M);
    vx_mp<false>(E, G, M);
    out[2] = vx_final<true, false>(G, bufA);
    vx_load_m(msm + 2 * 64, M);
    vx_mp<true>(E, G, M);
    out[3] = vx_final<false, false>(G, bufA);

    // ---- phase B: E23 = A x2 m2 x3 m3 (Y) -> X; out0 (open p0), out1 (open p1)
#pragma unroll
    for (int r = 0; r < 64; ++r) T[r] = bufA[lane * VX_PITCH + r];  // A in Y
    vx_load_m(msm + 2 * 64, M);
    vx_mp<true>(T, E, M);
    vx_load_m(msm + 3 * 64, M);
    vx_mp<false>(E, G, M);
    __syncthreads();  // every lane has finished reading bufB (phase A)
#pragma unroll
    for (int r = 0; r < 64; ++r) bufB[lane * VX_PITCH + r] = G[r];  // write from Y
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 64; ++r) E[r] = bufB[r * VX_PITCH + lane];  // E23 in X
    vx_load_m(msm + 1 * 64, M);
    vx_mp<false>(E, G, M);
    out[0] = vx_final<true, true>(G, bufA);
    vx_load_m(msm + 0 * 64, M);
    vx_mp<true>(E, G, M);
    out[1] = vx_final<false, true>(G, bufA);

    // ---- sum over s, normalise, residual, write
#pragma unroll
    for (int j = 0; j < 4; ++j) part[(w * 4 + j) * 64 + lane] = out[j];
    __syncthreads();
    for (int j = w; j < 4; j += NS) {
        const int q = outs[j];
        if (q < 0) continue;
        float x = 0.f;
#pragma unroll
        for (int s = 0; s < NS; ++s) x += part[(s * 4 + j) * 64 + lane];
        const double S = wave_sum((double)x);
        const double xn = S != 0.0 ? (double)x / S : (double)x;
        const float xf = (float)xn;
        const double xo = (double)mold[(long long)q * 64 + lane];
        mnew[(long long)q * 64 + lane] = xf;
        const double dn = wave_sum((double)xf * xo), nn = wave_sum((double)xf * (double)xf),
                     oo = wave_sum(xo * xo);
        const double mx = wave_max(fabs((double)xf - xo));
        if (lane == 0) {
            const double c = dn / sqrt(nn * oo);
            const double fid = 1.0 - c * c;
            res_fid[q] = fid;
            res_abs[q] = mx;
        }
    }
}

static int vertex_sweep(VertexPlan *p, hipStream_t st, DevTopo tp, const void *vsites,
                        const void *src, void *dst, const NetCtl &ctl, double *res_fid,
                        double *res_abs) {
    (void)tp;
    switch (p->d) {
#define VX_CASE(NS)                                                                         \
    case NS:                                                                                \
        hipLaunchKernelGGL(k_vertex_d8<NS>, dim3(p->nv), dim3(64 * NS), p->lds_bytes, st,    \
                           (const float *)vsites, (const float *)src, (float *)dst, p->d_vin, \
                           p->d_vout, p->d_vnet, ctl, res_fid, res_abs);                     \
        break;
        VX_CASE(1)
        VX_CASE(2)
        VX_CASE(3)
        VX_CASE(4)
#undef VX_CASE
        default:
            return 4;
    }
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

static int vertex_set_lds_limits(const VertexPlan *p) {
    const void *fns[4] = {(const void *)k_vertex_d8<1>, (const void *)k_vertex_d8<2>,
                          (const void *)k_vertex_d8<3>, (const void *)k_vertex_d8<4>};
    if (p->d < 1 || p->d > 4) return 4;
    return hipFuncSetAttribute(fns[p->d - 1], hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)p->lds_bytes) == hipSuccess ? 0 : 2;
}
