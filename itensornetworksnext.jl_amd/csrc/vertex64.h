// vertex64.h -- vertex-centric BP sweep in FP64 for bond dimensions <= 8 (gfx950 VALU).
//
// The FP64 counterpart of k_vertex_d8 (vertex.h): one workgroup per vertex, one wave per
// physical index value s, the wave holding the slice A_u[s] (4 ports padded to dimension 8,
// 4096 values, 64 per lane) in registers in layout X (register r = (p0, p1), lane =
// (p2, p3)) or Y (register = (p2, p3), lane = (p0, p1)).  On MI355X the FP64 vector rate
// equals the FP64 matrix rate (78.6 TFLOP/s), and the FP64 MFMAs have no K = 1 shape that
// would let a lane contract its own register column, so the products run on the VALU as
// FP64 FMAs.
//
// Schedule per slice (the same sharing of partial products as the FP32 kernel: 8 mode
// products and 4 open contractions instead of 16 and 4):
//     phase A: R = A (X) x0 m0 x1 m1 = E01 (X) -> LDS transpose -> E01 (Y)
//              out2 = open_p2(E01 x3 m3, A (Y)),  out3 = open_p3(E01 x2 m2, A (Y))
//     phase B: R = A (Y) x2 m2 x3 m3 = E23 (Y) -> LDS transpose -> E23 (X)
//              out0 = open_p0(E23 x1 m1, A (X)),  out1 = open_p1(E23 x0 m0, A (X))
// A mode product on a register port is done in place (one register column, or row, at a
// time), so a lane needs one 64-double array; the partial products E01 / E23 are kept in
// the wave's LDS buffer between their two uses.  The bra of every open contraction is read
// from one of two padded FP64 copies of the sites in HBM (layout X and layout Y), so each
// register row is one coalesced 512-byte load per wave.
//
// Open contraction (open port = register hi index, HI; or lo index, LO):
//     HI: out[a][a'] = sum_{lo, lanes} G[a*8+lo] A[a'*8+lo]
//     LO: out[a][a'] = sum_{hi, lanes} G[hi*8+a] A[hi*8+a']
// in two halves of a' (32 accumulators each), the bra rows streamed once; each half is
// summed over the wave by a halving exchange so lane l ends with entry (l % 32) of the half.
// Results are sum-normalised and written as in the FP32 kernel; residuals in FP64.
#pragma once

#include "vertex.h"

static constexpr int V64_PITCH = 65;             // LDS row pitch (doubles)
static constexpr int V64_BUF = 64 * V64_PITCH;   // one transposable 64 x 64 buffer

// FP64 plan: every network of the context qualifies (degree <= 4, physical dimension
// d <= 4 shared by all, bond dimensions <= 8 with the largest one enabled in `mask`).
struct Vertex64Plan {
    std::string why;
    int d = 0;
    int nv = 0;
    int4 *d_vin = nullptr, *d_vout = nullptr, *d_vdim = nullptr;
    int *d_vnet = nullptr, *d_vlist = nullptr;
    double *d_sx = nullptr;   // [item][s][4096] padded sites, layout X order
    double *d_sy = nullptr;   // the same slices in layout Y order
    size_t lds_bytes = 0;
};

static void vertex64_release(Vertex64Plan *p) {
    void *ptrs[7] = {p->d_vin, p->d_vout, p->d_vdim, p->d_vnet, p->d_vlist, p->d_sx, p->d_sy};
    for (void *x : ptrs)
        if (x) (void)hipFree(x);
    p->d_vin = p->d_vout = p->d_vdim = nullptr;
    p->d_vnet = p->d_vlist = nullptr;
    p->d_sx = p->d_sy = nullptr;
    p->nv = p->d = 0;
}

// 0 on success; otherwise the reason is in p->why.
static int vertex64_plan_build(Vertex64Plan *p, int plain, int nv, const int *port_off,
                               const int *port_edge, const int *edge_dim, const int *phys_dim,
                               const int *port_rev, const int *vnet, unsigned mask) {
    vertex64_release(p);
    p->why.clear();
    if (plain) { p->why = "plain (vector-message) networks use the generic kernel"; return 4; }
    if (nv == 0) { p->why = "empty network"; return 4; }
    const int d = phys_dim[0];
    if (d < 1 || d > 4) { p->why = "physical dimension outside 1..4"; return 4; }
    for (int u0 = 0; u0 < nv;) {
        int u1 = u0;
        while (u1 < nv && vnet[u1] == vnet[u0]) ++u1;
        int dmax = 0;
        for (int u = u0; u < u1; ++u) {
            if (port_off[u + 1] - port_off[u] > 4) { p->why = "degree > 4"; return 4; }
            if (phys_dim[u] != d) { p->why = "physical dimensions differ"; return 4; }
            for (int q = port_off[u]; q < port_off[u + 1]; ++q) dmax = std::max(dmax, edge_dim[port_edge[q]]);
        }
        if (dmax > VX_D) { p->why = "bond dimension > 8"; return 4; }
        if (dmax > 0 && !((mask >> dmax) & 1u)) {
            p->why = "bond dimension " + std::to_string(dmax) + " runs on the generic kernel";
            return 4;
        }
        u0 = u1;
    }
    std::vector<int4> vin(nv), vout(nv), vdim(nv);
    std::vector<int> vn(nv), vl(nv);
    for (int u = 0; u < nv; ++u) {
        int in[4] = {-1, -1, -1, -1}, out[4] = {-1, -1, -1, -1}, dm[4] = {0, 0, 0, 0};
        for (int j = 0; j < port_off[u + 1] - port_off[u]; ++j) {
            out[j] = port_off[u] + j;
            in[j] = port_rev[port_off[u] + j];
            dm[j] = edge_dim[port_edge[port_off[u] + j]];
        }
        vin[u] = make_int4(in[0], in[1], in[2], in[3]);
        vout[u] = make_int4(out[0], out[1], out[2], out[3]);
        vdim[u] = make_int4(dm[0], dm[1], dm[2], dm[3]);
        vn[u] = vnet[u];
        vl[u] = u;
    }
    int rc = vx_upload(&p->d_vin, vin);
    if (!rc) rc = vx_upload(&p->d_vout, vout);
    if (!rc) rc = vx_upload(&p->d_vdim, vdim);
    if (!rc) rc = vx_upload(&p->d_vnet, vn);
    if (!rc) rc = vx_upload(&p->d_vlist, vl);
    if (rc) {
        p->why = "allocation failed";
        vertex64_release(p);
        return rc;
    }
    p->d = d;
    p->nv = nv;
    p->lds_bytes = sizeof(double) * ((size_t)d * V64_BUF + 4 * 64 + (size_t)d * 4 * 64);
    return 0;
}

// Padded FP64 slices in layout X (element e = r * 64 + l, r = (p0, p1), l = (p2, p3)) and
// layout Y (element e' = r' * 64 + l' with r' = (p2, p3), l' = (p0, p1)).
__global__ void k_vertex64_pad_sites(const double *sites64, const long long *site_off,
                                     const int *port_off, const int *port_edge,
                                     const int *edge_dim, const int *vlist, int d, int nv,
                                     double *sx, double *sy) {
    const long long total = (long long)nv * d * 4096;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int e = (int)(i & 4095);
        const long long us = i >> 12;
        const int s = (int)(us % d);
        const int u = vlist[us / d];
        const int k = port_off[u + 1] - port_off[u];
        const int idx[4] = {(e >> 9) & 7, (e >> 6) & 7, (e >> 3) & 7, e & 7};
        long long off = 0, n = 1;
        bool real = true;
        for (int j = 0; j < 4 && real; ++j) {
            if (j < k) {
                const int D = edge_dim[port_edge[port_off[u] + j]];
                if (idx[j] >= D) real = false;
                off = off * D + idx[j];
                n *= D;
            } else if (idx[j] != 0) {
                real = false;
            }
        }
        const double v = real ? sites64[site_off[u] + s * n + off] : 0.0;
        sx[i] = v;
        // X element (r, l) sits at Y position (l, r)
        sy[(us << 12) + (long long)(e & 63) * 64 + (e >> 6)] = v;
    }
}

static int vertex64_upload_sites(Vertex64Plan *p, hipStream_t st, const double *d_sites64,
                                 const long long *d_site_off, const int *d_port_off,
                                 const int *d_port_edge, const int *d_edge_dim) {
    const size_t n = (size_t)p->nv * p->d * 4096;
    if (p->d_sx) (void)hipFree(p->d_sx);
    if (p->d_sy) (void)hipFree(p->d_sy);
    p->d_sx = p->d_sy = nullptr;
    if (n == 0) return 0;
    if (hipMalloc((void **)&p->d_sx, sizeof(double) * n) != hipSuccess ||
        hipMalloc((void **)&p->d_sy, sizeof(double) * n) != hipSuccess)
        return 5;
    hipLaunchKernelGGL(k_vertex64_pad_sites, dim3(2048), dim3(256), 0, st, d_sites64, d_site_off,
                       d_port_off, d_port_edge, d_edge_dim, p->d_vlist, p->d, p->nv, p->d_sx,
                       p->d_sy);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

// ---------------------------------------------------------------- device helpers

// In-place mode product on a register port of R[hi * 8 + lo] with the 8 x 8 message m
// (LDS, broadcast reads): HI: R[b*8+x] = sum_a R[a*8+x] m[a][b]; LO: R[x*8+b] = sum_a
// R[x*8+a] m[a][b].  Column (row) x is read whole before it is written.
template <bool HI>
__device__ __forceinline__ void v64_mode(double (&R)[64], const double *m) {
    double M[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) M[i] = m[i];
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        double t[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) t[a] = HI ? R[a * 8 + x] : R[x * 8 + a];
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            double acc = t[0] * M[b];
#pragma unroll
            for (int a = 1; a < 8; ++a) acc = fma(t[a], M[a * 8 + b], acc);
            if (HI) R[b * 8 + x] = acc; else R[x * 8 + b] = acc;
        }
    }
}

// One step of a halving exchange inside 16-lane rows (doubles): lane l keeps entries
// i + S * (bit S of l) for i < S and adds the same entries from its DPP partner.
template <int S, int N>
__device__ __forceinline__ void v64_halve(double (&v)[N], int lane) {
    const bool up = (lane & S) != 0;
#pragma unroll
    for (int i = 0; i < S; ++i) {
        const double lo = v[i], hi = v[i + S];
        const double send = up ? lo : hi;
        const double keep = up ? hi : lo;
        v[i] = keep + row_partner<S>(send);
    }
}

// Sum of 32 values over the wave: every lane ends with entry (lane % 32).
__device__ __forceinline__ double v64_reduce32(double (&acc)[32]) {
    const int lane = threadIdx.x & 63;
    (void)lane;
#pragma unroll
    for (int i = 0; i < 32; ++i) {
        const DPair h = halves32_pair(acc[i]);
        acc[i] = h.first + h.second;  // sum over lane bit 5, on both halves
    }
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        // v_permlane16_swap(x, y): rows 1 and 3 of x trade places with rows 0 and 2 of y,
        // so x' + y' holds entry i (summed over the row pair) on rows 0 and 2 and entry
        // i + 16 on rows 1 and 3 -- per 32-bit half of the double
        const double x = acc[i], y = acc[i + 16];
        const auto l = __builtin_amdgcn_permlane16_swap((unsigned)__double2loint(x),
                                                        (unsigned)__double2loint(y), false, false);
        const auto h = __builtin_amdgcn_permlane16_swap((unsigned)__double2hiint(x),
                                                        (unsigned)__double2hiint(y), false, false);
        acc[i] = __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
    }
    v64_halve<8>(acc, lane);
    v64_halve<4>(acc, lane);
    v64_halve<2>(acc, lane);
    v64_halve<1>(acc, lane);
    return acc[0];
}

// Open contraction of the partial product G (registers) with the bra rows of `bra` (the
// slice in the layout of G, row r of lane l at bra[r * 64 + l]).  Entry (a, a') of the
// result is written to out[a * 8 + a'] (LDS) by lanes 0..31.
template <bool HI>
__device__ __forceinline__ void v64_open(const double (&G)[64], const double *__restrict__ bra,
                                         double *out) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        double acc[32];
#pragma unroll
        for (int i = 0; i < 32; ++i) acc[i] = 0.0;
#pragma unroll
        for (int ap = 0; ap < 4; ++ap) {
            const int a2 = 4 * h + ap;  // a'
            double row[8];
#pragma unroll
            for (int y = 0; y < 8; ++y)
                row[y] = HI ? bra[(a2 * 8 + y) * 64 + lane] : bra[(y * 8 + a2) * 64 + lane];
#pragma unroll
            for (int a = 0; a < 8; ++a)
#pragma unroll
                for (int y = 0; y < 8; ++y) {
                    const double g = HI ? G[a * 8 + y] : G[y * 8 + a];
                    acc[a * 4 + ap] = fma(g, row[y], acc[a * 4 + ap]);
                }
        }
        const double v = v64_reduce32(acc);
        const int i = lane & 31;
        if (lane < 32) out[(i >> 2) * 8 + 4 * h + (i & 3)] = v;
    }
}

// ---------------------------------------------------------------- the sweep kernel

template <int NS>
__global__ __launch_bounds__(64 * NS) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_vertex_d8_f64(const double *__restrict__ sx, const double *__restrict__ sy,
                     const double *__restrict__ mold, double *__restrict__ mnew,
                     const int4 *__restrict__ vin, const int4 *__restrict__ vout,
                     const int4 *__restrict__ vdim, const int *__restrict__ vnet,
                     const long long *__restrict__ msg_off, NetCtl ctl,
                     double *__restrict__ res_fid, double *__restrict__ res_abs) {
    extern __shared__ __align__(16) double v64sm[];
    const int u = blockIdx.x;
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    double *buf = v64sm + w * V64_BUF;
    double *msm = v64sm + NS * V64_BUF;  // 4 padded incoming messages
    double *part = msm + 4 * 64;         // [s][port][64] per-slice results
    const int4 vi = vin[u], vo = vout[u], vdm = vdim[u];
    const int ins[4] = {vi.x, vi.y, vi.z, vi.w};
    const int outs[4] = {vo.x, vo.y, vo.z, vo.w};
    const int dims[4] = {vdm.x, vdm.y, vdm.z, vdm.w};
    const int la = lane >> 3, lb = lane & 7;
    const int net = vnet[u];
    if (ctl.enabled && ctl.done[net]) {  // uniform over the workgroup
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j % NS != w || outs[j] < 0) continue;
            const long long o = msg_off[outs[j]];
            if (lane < dims[j] * dims[j]) mnew[o + lane] = mold[o + lane];
            if (lane == 0) { res_fid[outs[j]] = 0.0; res_abs[outs[j]] = 0.0; }
        }
        return;
    }
    // incoming messages zero-padded to 8 x 8 (identity on padded ports)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j % NS != w) continue;
        const int D = dims[j];
        double x;
        if (ins[j] < 0) x = la == lb ? 1.0 : 0.0;
        else x = (la < D && lb < D) ? mold[msg_off[ins[j]] + la * D + lb] : 0.0;
        msm[j * 64 + lane] = x;
    }
    const long long so = ((long long)u * NS + w) * 4096;
    const double *ax = sx + so, *ay = sy + so;
    double *po = part + w * 4 * 64;
    double R[64];
#pragma unroll
    for (int r = 0; r < 64; ++r) R[r] = ax[r * 64 + lane];
    __syncthreads();  // messages in LDS

    // ---- phase A: E01 = A x0 m0 x1 m1 (X), transposed to Y through the wave's buffer
    v64_mode<true>(R, msm + 0 * 64);
    v64_mode<false>(R, msm + 1 * 64);
#pragma unroll
    for (int r = 0; r < 64; ++r) buf[r * V64_PITCH + lane] = R[r];
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 64; ++r) R[r] = buf[lane * V64_PITCH + r];  // E01 in Y
    v64_mode<false>(R, msm + 3 * 64);
    v64_open<true>(R, ay, po + 2 * 64);   // open port 2 (Y register hi)
#pragma unroll
    for (int r = 0; r < 64; ++r) R[r] = buf[lane * V64_PITCH + r];  // E01 in Y again
    v64_mode<true>(R, msm + 2 * 64);
    v64_open<false>(R, ay, po + 3 * 64);  // open port 3 (Y register lo)

    // ---- phase B: E23 = A x2 m2 x3 m3 (Y), transposed to X
#pragma unroll
    for (int r = 0; r < 64; ++r) R[r] = ay[r * 64 + lane];
    v64_mode<true>(R, msm + 2 * 64);
    v64_mode<false>(R, msm + 3 * 64);
    wave_lds_sync();  // every lane has finished reading E01 from the buffer
#pragma unroll
    for (int r = 0; r < 64; ++r) buf[lane * V64_PITCH + r] = R[r];
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 64; ++r) R[r] = buf[r * V64_PITCH + lane];  // E23 in X
    v64_mode<false>(R, msm + 1 * 64);
    v64_open<true>(R, ax, po + 0 * 64);   // open port 0 (X register hi)
#pragma unroll
    for (int r = 0; r < 64; ++r) R[r] = buf[r * V64_PITCH + lane];  // E23 in X again
    v64_mode<true>(R, msm + 0 * 64);
    v64_open<false>(R, ax, po + 1 * 64);  // open port 1 (X register lo)

    // ---- sum over s, normalise, residual, write
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int q = outs[j];
        if (j % NS != w || q < 0) continue;
        double x = 0.0;
#pragma unroll
        for (int s = 0; s < NS; ++s) x += part[(s * 4 + j) * 64 + lane];
        const double S = wave_sum_v(x);
        const double xn = S != 0.0 ? x / S : x;  // exactly 0 on padded entries
        const int D = dims[j];
        const bool real = la < D && lb < D;
        const long long o = msg_off[q] + la * D + lb;
        const double xo = real ? mold[o] : 0.0;
        if (real) mnew[o] = xn;
        const double dn = wave_sum_v(xn * xo), nn = wave_sum_v(xn * xn), oo = wave_sum_v(xo * xo);
        const double mx = wave_max_v(fabs(xn - xo));
        if (lane == 0) {
            const double c = dn / sqrt(nn * oo);
            res_fid[q] = 1.0 - c * c;
            res_abs[q] = mx;
        }
    }
}

static int vertex64_set_lds_limit(const Vertex64Plan *p) {
    const void *fns[4] = {(const void *)k_vertex_d8_f64<1>, (const void *)k_vertex_d8_f64<2>,
                          (const void *)k_vertex_d8_f64<3>, (const void *)k_vertex_d8_f64<4>};
    if (p->d < 1 || p->d > 4) return 4;
    return hipFuncSetAttribute(fns[p->d - 1], hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)p->lds_bytes) == hipSuccess ? 0 : 2;
}

static int vertex64_sweep(const Vertex64Plan *p, hipStream_t st, const long long *msg_off,
                          const double *src, double *dst, const NetCtl &ctl, double *res_fid,
                          double *res_abs) {
    if (p->nv == 0) return 0;
    switch (p->d) {
#define V64_CASE(NS)                                                                         \
    case NS:                                                                                 \
        hipLaunchKernelGGL((k_vertex_d8_f64<NS>), dim3(p->nv), dim3(64 * NS), p->lds_bytes,  \
                           st, p->d_sx, p->d_sy, src, dst, p->d_vin, p->d_vout, p->d_vdim,   \
                           p->d_vnet, msg_off, ctl, res_fid, res_abs);                       \
        break;
        V64_CASE(1)
        V64_CASE(2)
        V64_CASE(3)
        V64_CASE(4)
#undef V64_CASE
        default:
            return 4;
    }
    return hipGetLastError() == hipSuccess ? 0 : 2;
}
