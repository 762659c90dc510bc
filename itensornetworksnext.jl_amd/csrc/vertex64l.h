// vertex64l.h -- FP64 vertex sweep for bond dimensions 9..16 (padded to 16), on
// v_mfma_f64_16x16x4f64.
//
// A 16^4 FP64 slice is 512 KiB, more than LDS, so the slice stays in HBM and the sweep is
// cut into 16 x 16 tiles that every product handles whole.  With A[s][p0][p1][p2][p3] and the
// shared partial products of the other vertex kernels:
//   pass 1, per column (p2, p3):  E01[:, :, p2, p3] = m0^T A[:, :, p2, p3] m1
//   pass 2, per block (p0, p1):   out2 += open over p3 of (E01 blk m3) with A blk,
//                                 out3 += open over p2 of (m2^T E01 blk) with A blk,
//                                 E23 blk = m2^T A blk m3
//   pass 3, per column (p2, p3):  out0 += open of (E23 col m1) with A col,
//                                 out1 += open of (m0^T E23 col) with A col
// which is 8 mode products and 4 open contractions per slice (25 MFLOP at D = 16), A read
// three times and E01, E23 written and read once.  Each pass is one launch.  Passes 1 and 3
// run one workgroup of 8 waves per (slice, p2) and read their column tiles from a slab
// [p0][p1][p3] staged in LDS, so that HBM sees whole 128-byte rows; pass 2 runs one
// workgroup of 4 waves per (slice, quarter of p0), a wave per 16 blocks.  The per-wave
// partial sums of the opens meet in LDS and go to a per-(slice, chunk) buffer that the
// finishing kernel sums in a fixed order (no atomics: the result does not depend on
// scheduling), normalises (beliefpropagation.jl:248-253) and compares with the previous
// message (fidelity and max-abs residuals).
//
// MFMA operand layouts (16x16x4 FP64, lane l: li = l % 16, lk = l / 16): the A operand of
// step t is A[li][4t + lk], the B operand B[4t + lk][li], and register r of the result is
// C[lk + 4r][li].  The result of one product is therefore the B operand of the next one
// contracting its row index, so a chain m_f^T (X m_l) needs no data movement; a message
// m[p][b] sits in registers as m[4t + lk][li] for both uses (B operand of X m, A operand of
// m^T X).
#pragma once

#include "vertex.h"

typedef double v4d __attribute__((ext_vector_type(4)));

static constexpr int V64L_ELEMS = 65536;  // 16^4 per slice
static constexpr int V64L_CHUNKS = 16;      // partial-message chunks per slice and port

struct Vertex64LPlan {
    std::string why;
    int d = 0, nv = 0;
    int4 *d_vin = nullptr, *d_vout = nullptr, *d_vdim = nullptr;
    int *d_vnet = nullptr, *d_vlist = nullptr;
    double *d_sites = nullptr;  // [item][s][16^4] padded slices, natural port order
    double *d_x = nullptr;      // E01 of every slice
    double *d_y = nullptr;      // E23 of every slice
    double *d_part = nullptr;   // [item][s][chunk 16][port][256] partial messages (ports 0, 1:
                                // one chunk per p2; ports 2, 3: chunks 0..3, one per quarter
                                // of p0)
};

static void vertex64l_release(Vertex64LPlan *p) {
    void *ptrs[9] = {p->d_vin, p->d_vout, p->d_vdim, p->d_vnet, p->d_vlist,
                     p->d_sites, p->d_x, p->d_y, p->d_part};
    for (void *x : ptrs)
        if (x) (void)hipFree(x);
    p->d_vin = p->d_vout = p->d_vdim = nullptr;
    p->d_vnet = p->d_vlist = nullptr;
    p->d_sites = p->d_x = p->d_y = p->d_part = nullptr;
    p->nv = p->d = 0;
}

// 0 on success (every network qualifies: degree <= 4, one physical dimension 1..4, largest
// bond dimension in 9..16 and enabled in `mask`); otherwise the reason is in p->why.
static int vertex64l_plan_build(Vertex64LPlan *p, int plain, int nv, const int *port_off,
                                const int *port_edge, const int *edge_dim, const int *phys_dim,
                                const int *port_rev, const int *vnet, unsigned mask) {
    vertex64l_release(p);
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
        if (dmax < 9 || dmax > 16) { p->why = "largest bond dimension outside 9..16"; return 4; }
        if (!((mask >> dmax) & 1u)) {
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
    const size_t ns = (size_t)nv * d;
    if (!rc && (hipMalloc((void **)&p->d_x, sizeof(double) * ns * V64L_ELEMS) != hipSuccess ||
                hipMalloc((void **)&p->d_y, sizeof(double) * ns * V64L_ELEMS) != hipSuccess ||
                hipMalloc((void **)&p->d_part, sizeof(double) * ns * V64L_CHUNKS * 4 * 256) != hipSuccess))
        rc = 5;
    if (rc) {
        p->why = "allocation failed";
        vertex64l_release(p);
        return rc;
    }
    p->d = d;
    p->nv = nv;
    return 0;
}

// Padded FP64 slices [item][s][16^4], natural port order (missing ports: index 0 only).
// One workgroup per (slice, p0): the vertex's bond dimensions are read once, then the 4096
// entries [p1][p2][p3] are written contiguously (missing ports count as dimension 1).
__global__ __launch_bounds__(256) void k_vertex64l_pad_sites(const double *sites64,
                                                             const long long *site_off,
                                                             const int *port_off,
                                                             const int *port_edge,
                                                             const int *edge_dim,
                                                             const int *vlist, int d,
                                                             double *out) {
    const int us = blockIdx.x >> 4, p0 = blockIdx.x & 15;
    const int s = us % d, u = vlist[us / d];
    const int k = port_off[u + 1] - port_off[u];
    int D[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) D[j] = j < k ? edge_dim[port_edge[port_off[u] + j]] : 1;
    const long long n = (long long)D[0] * D[1] * D[2] * D[3];
    const double *src = sites64 + site_off[u] + s * n;
    double *dst = out + (size_t)us * V64L_ELEMS + p0 * 4096;
    for (int e = threadIdx.x; e < 4096; e += blockDim.x) {
        const int p1 = e >> 8, p2 = (e >> 4) & 15, p3 = e & 15;
        const bool real = p0 < D[0] && p1 < D[1] && p2 < D[2] && p3 < D[3];
        dst[e] = real ? src[((p0 * D[1] + p1) * D[2] + p2) * D[3] + p3] : 0.0;
    }
}

static int vertex64l_upload_sites(Vertex64LPlan *p, hipStream_t st, const double *d_sites64,
                                  const long long *d_site_off, const int *d_port_off,
                                  const int *d_port_edge, const int *d_edge_dim) {
    const size_t n = (size_t)p->nv * p->d * V64L_ELEMS;
    if (p->d_sites) (void)hipFree(p->d_sites);
    p->d_sites = nullptr;
    if (n == 0) return 0;
    if (hipMalloc((void **)&p->d_sites, sizeof(double) * n) != hipSuccess) return 5;
    hipLaunchKernelGGL(k_vertex64l_pad_sites, dim3(p->nv * p->d * 16), dim3(256), 0, st,
                       d_sites64, d_site_off, d_port_off, d_port_edge, d_edge_dim, p->d_vlist,
                       p->d, p->d_sites);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

// ---------------------------------------------------------------- device helpers

__device__ __forceinline__ v4d v64l_mf(double a, double b, v4d c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// The incoming message of a port in operand registers: mr[t] = m[4t + lk][li] (zero-padded to
// 16 x 16; a padded port's message is e0 e0^T).
__device__ __forceinline__ void v64l_msg(const double *msgs, long long off, int D, double (&mr)[4]) {
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int t = 0; t < 4; ++t) {
        const int k = 4 * t + lk;
        if (off == -1) mr[t] = (k == 0 && li == 0) ? 1.0 : 0.0;
        else mr[t] = (k < D && li < D) ? msgs[off + k * D + li] : 0.0;
    }
}

// A tile T[x][y] at base + x * sx + y * sy: row pattern R[t] = T[li][4t + lk], column pattern
// C[t] = T[4t + lk][li].
__device__ __forceinline__ void v64l_row(const double *base, int sx, int sy, double (&R)[4]) {
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int t = 0; t < 4; ++t) R[t] = base[li * sx + (4 * t + lk) * sy];
}
__device__ __forceinline__ void v64l_col(const double *base, int sx, int sy, double (&C)[4]) {
    const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int t = 0; t < 4; ++t) C[t] = base[(4 * t + lk) * sx + li * sy];
}

// m_f^T X m_l from the row pattern of X: register r holds E[lk + 4r][li].
__device__ __forceinline__ v4d v64l_sandwich(const double (&Xr)[4], const double (&mf)[4],
                                             const double (&ml)[4]) {
    v4d z = {0.0, 0.0, 0.0, 0.0}, e = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < 4; ++t) z = v64l_mf(Xr[t], ml[t], z);  // Z = X m_l, rows x
#pragma unroll
    for (int t = 0; t < 4; ++t) e = v64l_mf(mf[t], z[t], e);   // m_f^T Z
    return e;
}

// The two halves of an open contraction: lead = sum_t m[t] (x) X[t] (the product of a message
// with a tile, as the B operand of the next product), close: acc += T lead.
__device__ __forceinline__ v4d v64l_lead(const double (&m)[4], const double (&X)[4]) {
    v4d g = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < 4; ++t) g = v64l_mf(m[t], X[t], g);
    return g;
}
__device__ __forceinline__ void v64l_close(const double (&T)[4], const v4d &g, v4d &acc) {
#pragma unroll
    for (int t = 0; t < 4; ++t) acc = v64l_mf(T[t], g[t], acc);
}

// acc[r] += out[a = li][a' = lk + 4r] of the open over the tile's first index of (X m_l)
// against the A tile: out[a][a'] = sum_b (X m_l)[a][b] A[a'][b].
__device__ __forceinline__ void v64l_open_last(const double (&Xr)[4], const double (&Ar)[4],
                                               const double (&ml)[4], v4d &acc) {
    v4d g = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < 4; ++t) g = v64l_mf(ml[t], Xr[t], g);   // (X m_l)^T, rows b
#pragma unroll
    for (int t = 0; t < 4; ++t) acc = v64l_mf(Ar[t], g[t], acc);
}

// acc[r] += out[a = li][a' = lk + 4r] of the open over the tile's second index of
// (m_f^T X): out[a][a'] = sum_b (m_f^T X)[b][a] A[b][a'].
__device__ __forceinline__ void v64l_open_first(const double (&Xc)[4], const double (&Ac)[4],
                                                const double (&mf)[4], v4d &acc) {
    v4d g = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int t = 0; t < 4; ++t) g = v64l_mf(mf[t], Xc[t], g);   // m_f^T X, rows b
#pragma unroll
    for (int t = 0; t < 4; ++t) acc = v64l_mf(Ac[t], g[t], acc);
}

// Sum the workgroup's waves' accumulators in LDS (red: 2 * 256 doubles per wave) and store
// out[a][a'] of each (256 entries).
__device__ __forceinline__ void v64l_store_partials(double *red, const v4d &acc_a, const v4d &acc_b,
                                                    double *dst_a, double *dst_b) {
    const int nw = (int)(blockDim.x >> 6);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        red[(w * 2 + 0) * 256 + li * 16 + lk + 4 * r] = acc_a[r];
        red[(w * 2 + 1) * 256 + li * 16 + lk + 4 * r] = acc_b[r];
    }
    __syncthreads();
    for (int e = threadIdx.x; e < 512; e += blockDim.x) {
        const int h = e >> 8, x = e & 255;
        double s = 0.0;
        for (int v = 0; v < nw; ++v) s += red[(v * 2 + h) * 256 + x];
        (h == 0 ? dst_a : dst_b)[x] = s;
    }
}

// Workgroup (slice us, quarter c4) of the passes: 256 threads, wave w takes outer index
// 4 c4 + w.  `us` = item * d + s.
struct V64LArgs {
    const double *sites;
    double *x, *y, *part;
    const double *mold;
    const int4 *vin, *vdim;
    const int *vnet;
    const long long *msg_off;
    NetCtl ctl;
    int d;
};

__device__ __forceinline__ bool v64l_begin(const V64LArgs &a, int item, int4 &in4, int4 &dim4) {
    if (a.ctl.enabled && a.ctl.done[a.vnet[item]]) return false;  // uniform: finisher copies
    in4 = a.vin[item];
    dim4 = a.vdim[item];
    return true;
}

__device__ __forceinline__ void v64l_port_msg(const V64LArgs &a, int4 in4, int4 dim4, int j,
                                              double (&mr)[4]) {
    const int q = j == 0 ? in4.x : j == 1 ? in4.y : j == 2 ? in4.z : in4.w;
    const int D = j == 0 ? dim4.x : j == 1 ? dim4.y : j == 2 ? dim4.z : dim4.w;
    v64l_msg(a.mold, q >= 0 ? a.msg_off[q] : -1, D, mr);
}

// A slab [p0][p1][p3] of the slice at fixed p2 (4096 entries) in LDS, rows of 16 padded to
// V64L_SP (row (p0, p1) at (p0 * 16 + p1) * V64L_SP): the column tiles of passes 1 and 3 are
// read from it, while HBM sees whole 128-byte rows.
static constexpr int V64L_SP = 17;
static constexpr int V64L_SLAB = 256 * V64L_SP;

__device__ __forceinline__ void v64l_slab_load(double *slab, const double *g, int p2) {
    for (int i = threadIdx.x; i < 4096; i += blockDim.x) {
        const int row = i >> 4, p3 = i & 15;  // row = p0 * 16 + p1
        slab[row * V64L_SP + p3] = g[row * 256 + p2 * 16 + p3];
    }
}

// pass 1: E01 columns, one workgroup of 8 waves per (slice, p2), wave w takes p3 = 2w, 2w + 1.
// Column (p2, p3) of the slab: element (p0, p1) at (p0 * 16 + p1) * V64L_SP + p3.  One slab
// of LDS (34 KiB, four workgroups per compute unit): the column tiles are read into
// registers, then the slab takes the E01 columns for a coalesced store (342 us per launch
// at 24x24 D=16, against 396 with a second slab for the results).
__global__ __launch_bounds__(512) void k_v64l_pass1(V64LArgs a) {
    __shared__ double slab[V64L_SLAB];
    const int us = blockIdx.x >> 4, p2 = blockIdx.x & 15, item = us / a.d;
    int4 in4, dim4;
    if (!v64l_begin(a, item, in4, dim4)) return;
    double m0[4], m1[4];
    v64l_port_msg(a, in4, dim4, 0, m0);
    v64l_port_msg(a, in4, dim4, 1, m1);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const double *A = a.sites + (size_t)us * V64L_ELEMS;
    double *X = a.x + (size_t)us * V64L_ELEMS;
    v64l_slab_load(slab, A, p2);
    __syncthreads();
    double R[2][4];
#pragma unroll
    for (int k = 0; k < 2; ++k) v64l_row(slab + 2 * w + k, 16 * V64L_SP, V64L_SP, R[k]);
    __syncthreads();  // every tile is in registers: the slab takes the results
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int p3 = 2 * w + k;
        const v4d e = v64l_sandwich(R[k], m0, m1);
#pragma unroll
        for (int r = 0; r < 4; ++r) slab[((lk + 4 * r) * 16 + li) * V64L_SP + p3] = e[r];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < 4096; i += blockDim.x) {
        const int row = i >> 4, p3 = i & 15;
        X[row * 256 + p2 * 16 + p3] = slab[row * V64L_SP + p3];
    }
}

// pass 2: out2, out3 and E23, per block (p0, p1) = 256 contiguous entries [p2][p3].
// (Loading block p1 + 1 into registers while block p1 is contracted, at four waves per SIMD
// instead of three, measured 406 against 401 us at 24x24 D=16.)
__global__ __launch_bounds__(256) void k_v64l_pass2(V64LArgs a) {
    __shared__ double red[8 * 256];
    const int us = blockIdx.x >> 2, c4 = blockIdx.x & 3, item = us / a.d;
    int4 in4, dim4;
    if (!v64l_begin(a, item, in4, dim4)) return;
    double m2[4], m3[4];
    v64l_port_msg(a, in4, dim4, 2, m2);
    v64l_port_msg(a, in4, dim4, 3, m3);
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
    const int p0 = 4 * c4 + w;
    const double *A = a.sites + (size_t)us * V64L_ELEMS;
    const double *X = a.x + (size_t)us * V64L_ELEMS;
    double *Y = a.y + (size_t)us * V64L_ELEMS;
    v4d acc2 = {0.0, 0.0, 0.0, 0.0}, acc3 = {0.0, 0.0, 0.0, 0.0};
    for (int p1 = 0; p1 < 16; ++p1) {
        const int base = (p0 * 16 + p1) * 256;
        double Er[4], Ec[4], Ar[4], Ac[4];
        v64l_row(X + base, 16, 1, Er);
        v64l_col(X + base, 16, 1, Ec);
        v64l_row(A + base, 16, 1, Ar);
        v64l_col(A + base, 16, 1, Ac);
        v64l_open_last(Er, Ar, m3, acc2);   // out2: open over p2 of E01 x3 m3
        v64l_open_first(Ec, Ac, m2, acc3);  // out3: open over p3 of E01 x2 m2
        const v4d e = v64l_sandwich(Ar, m2, m3);
#pragma unroll
        for (int r = 0; r < 4; ++r) Y[base + (lk + 4 * r) * 16 + li] = e[r];
    }
    double *P = a.part + ((size_t)us * V64L_CHUNKS + c4) * 4 * 256;  // chunks 0..3 of ports 2, 3
    v64l_store_partials(red, acc2, acc3, P + 2 * 256, P + 3 * 256);
}

// pass 3: out0, out1 per column (p2, p3) of E23, one workgroup of 8 waves per (slice, p2),
// wave w takes p3 = 2w, 2w + 1.  One slab of LDS (34 KiB): the E23 slab is staged and its
// tiles' first products made, then the slab takes A for the second products, then the
// reduction.  At 8 waves per SIMD (64 registers, one spilled) four workgroups share a
// compute unit: 569 us per launch at 24x24 D=16, against 634 at 6 waves per SIMD and 701
// with both slabs resident (two workgroups per compute unit).  Half slabs, two workgroups
// per (slice, p2), fetched twice the bytes: each half row drew a whole 128-byte line.
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_v64l_pass3(V64LArgs a) {
    __shared__ double slab[V64L_SLAB];
    const int us = blockIdx.x >> 4, p2 = blockIdx.x & 15, item = us / a.d;
    int4 in4, dim4;
    if (!v64l_begin(a, item, in4, dim4)) return;
    double m0[4], m1[4];
    v64l_port_msg(a, in4, dim4, 0, m0);
    v64l_port_msg(a, in4, dim4, 1, m1);
    const int w = threadIdx.x >> 6;
    const double *A = a.sites + (size_t)us * V64L_ELEMS;
    const double *Y = a.y + (size_t)us * V64L_ELEMS;
    v64l_slab_load(slab, Y, p2);
    __syncthreads();
    v4d g0[2], g1[2];
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        double Er[4], Ec[4];
        v64l_row(slab + 2 * w + k, 16 * V64L_SP, V64L_SP, Er);
        v64l_col(slab + 2 * w + k, 16 * V64L_SP, V64L_SP, Ec);
        g0[k] = v64l_lead(m1, Er);  // (E23 x1 m1)^T
        g1[k] = v64l_lead(m0, Ec);  // m0^T E23
    }
    __syncthreads();
    v64l_slab_load(slab, A, p2);
    __syncthreads();
    v4d acc0 = {0.0, 0.0, 0.0, 0.0}, acc1 = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        double Ar[4], Ac[4];
        v64l_row(slab + 2 * w + k, 16 * V64L_SP, V64L_SP, Ar);
        v64l_col(slab + 2 * w + k, 16 * V64L_SP, V64L_SP, Ac);
        v64l_close(Ar, g0[k], acc0);  // out0: open over p0
        v64l_close(Ac, g1[k], acc1);  // out1: open over p1
    }
    __syncthreads();  // the slab is no longer read: reuse it for the reduction
    double *P = a.part + ((size_t)us * V64L_CHUNKS + p2) * 4 * 256;
    v64l_store_partials(slab, acc0, acc1, P + 0 * 256, P + 1 * 256);
}

// Finish, one workgroup of 256 threads per (item, port): entry e = (a, a') summed over the
// slices and quarters in a fixed order, normalised by the sum when it is nonzero, written
// as D x D, with the fidelity and max-abs residuals against the previous message.
__global__ __launch_bounds__(256) void k_v64l_finish(const double *__restrict__ part,
                                                     const double *__restrict__ mold,
                                                     double *__restrict__ mnew,
                                                     const int4 *__restrict__ vout,
                                                     const int4 *__restrict__ vdim,
                                                     const int *__restrict__ vnet,
                                                     const long long *__restrict__ msg_off,
                                                     NetCtl ctl, int d,
                                                     double *__restrict__ res_fid,
                                                     double *__restrict__ res_abs) {
    __shared__ double red[BP_BLOCK / 64 + 1];
    const int item = blockIdx.x >> 2, j = blockIdx.x & 3;
    const int4 o4 = vout[item], d4 = vdim[item];
    const int q = j == 0 ? o4.x : j == 1 ? o4.y : j == 2 ? o4.z : o4.w;
    if (q < 0) return;  // uniform
    const int D = j == 0 ? d4.x : j == 1 ? d4.y : j == 2 ? d4.z : d4.w;
    const long long mo = msg_off[q];
    const int e = threadIdx.x, a = e >> 4, ap = e & 15;
    const bool real = a < D && ap < D;
    if (ctl.enabled && ctl.done[vnet[item]]) {
        if (real) mnew[mo + a * D + ap] = mold[mo + a * D + ap];
        if (threadIdx.x == 0) {
            res_fid[q] = 0.0;
            res_abs[q] = 0.0;
        }
        return;
    }
    double x = 0.0;
    const int nch = j < 2 ? V64L_CHUNKS : 4;  // pass 3: one per p2; pass 2: one per quarter
    for (int s = 0; s < d; ++s)
        for (int c = 0; c < nch; ++c)
            x += part[(((size_t)(item * d + s) * V64L_CHUNKS + c) * 4 + j) * 256 + e];
    if (!real) x = 0.0;
    const double S = block_sum(x, red);
    const double xn = S != 0.0 ? x / S : x;
    const double xo = real ? mold[mo + a * D + ap] : 0.0;
    if (real) mnew[mo + a * D + ap] = xn;
    const double dn = block_sum(xn * xo, red), nn = block_sum(xn * xn, red),
                 oo = block_sum(xo * xo, red);
    const double mx = block_max(fabs(xn - xo), red);
    if (threadIdx.x == 0) {
        const double c = dn / sqrt(nn * oo);
        res_fid[q] = 1.0 - c * c;
        res_abs[q] = mx;
    }
}

// Vertex scalars (messagecache.jl:139-143): Z_u = <out_2, m_in,2>, the unnormalised port-2
// result of passes 1 and 2 (summed over the slices and the four chunks in a fixed order)
// against the incoming port-2 message; a padded port 2 holds e0 e0^T, so the formula covers
// every degree.  One workgroup of 256 threads per vertex, entry e = (a, a').
__global__ __launch_bounds__(256) void k_v64l_zv(const double *__restrict__ part,
                                                 const double *__restrict__ msgs,
                                                 const int4 *__restrict__ vin,
                                                 const int4 *__restrict__ vdim,
                                                 const long long *__restrict__ msg_off, int d,
                                                 double *__restrict__ zv) {
    __shared__ double red[BP_BLOCK / 64 + 1];
    const int item = blockIdx.x, e = threadIdx.x, a = e >> 4, ap = e & 15;
    const int q = vin[item].z, D = vdim[item].z;
    double m;
    if (q < 0) m = e == 0 ? 1.0 : 0.0;
    else m = (a < D && ap < D) ? msgs[msg_off[q] + a * D + ap] : 0.0;
    double x = 0.0;
    for (int s = 0; s < d; ++s)
        for (int c = 0; c < 4; ++c)
            x += part[(((size_t)(item * d + s) * V64L_CHUNKS + c) * 4 + 2) * 256 + e];
    const double z = block_sum(x * m, red);
    if (threadIdx.x == 0) zv[item] = z;
}

// Every vertex's scalar from the FP64 messages `msgs` into zv[vertex] (passes 1 and 2, then
// k_v64l_zv; E23 is made and not used).
static int vertex64l_scalars(const Vertex64LPlan *p, hipStream_t st, const long long *msg_off,
                             const double *msgs, double *zv) {
    if (p->nv == 0) return 0;
    V64LArgs a{p->d_sites, p->d_x, p->d_y, p->d_part, msgs, p->d_vin, p->d_vdim, p->d_vnet,
               msg_off, NetCtl{nullptr, 0}, p->d};
    const int ns = p->nv * p->d;
    hipLaunchKernelGGL(k_v64l_pass1, dim3(ns * 16), dim3(512), 0, st, a);
    hipLaunchKernelGGL(k_v64l_pass2, dim3(ns * 4), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_v64l_zv, dim3(p->nv), dim3(256), 0, st, p->d_part, msgs, p->d_vin,
                       p->d_vdim, msg_off, p->d, zv);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

// One FP64 sweep of every network: src -> dst (message buffers of the context's FP64 layout).
static int vertex64l_sweep(const Vertex64LPlan *p, hipStream_t st, const long long *msg_off,
                           const double *src, double *dst, const NetCtl &ctl, double *res_fid,
                           double *res_abs) {
    if (p->nv == 0) return 0;
    V64LArgs a{p->d_sites, p->d_x, p->d_y, p->d_part, src, p->d_vin, p->d_vdim, p->d_vnet,
               msg_off, ctl, p->d};
    const int ns = p->nv * p->d;
    hipLaunchKernelGGL(k_v64l_pass1, dim3(ns * 16), dim3(512), 0, st, a);
    hipLaunchKernelGGL(k_v64l_pass2, dim3(ns * 4), dim3(256), 0, st, a);
    hipLaunchKernelGGL(k_v64l_pass3, dim3(ns * 16), dim3(512), 0, st, a);
    hipLaunchKernelGGL(k_v64l_finish, dim3(p->nv * 4), dim3(256), 0, st, p->d_part, src, dst,
                       p->d_vout, p->d_vdim, p->d_vnet, msg_off, ctl, p->d, res_fid, res_abs);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

