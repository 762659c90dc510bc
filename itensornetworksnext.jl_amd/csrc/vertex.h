// vertex.h -- vertex-centric BP sweep for bond dimension 8 on gfx950 (MFMA, FP32 or BF16).
//
// One workgroup per vertex u computes all of u's outgoing messages (the reference
// computes each message separately, beliefpropagation.jl:242-257; the result is the
// same contraction).  One wave per physical index value s: the wave holds the slice
// A_u[s] (4 ports of dimension 8 = 4096 values, 64 per lane) in registers in one of two
// layouts
//     X: register r = (p0, p1), lane l = (p2, p3)      Y: register r = (p2, p3), lane = (p0, p1)
// A ket-side mode product on a register port ("T x_j m_j") runs on the 16-block 4x4
// MFMAs: register i of lane 4k+j accumulates A(lane 4k+i) . B(lane 4k+j) (one product for
// v_mfma_f32_4x4x1_16b_f32, a 4-term dot of BF16 pairs for v_mfma_f32_4x4x4_16b_bf16), so
// with A = m_j[a][4h + lane%4] (the same for every block) and B = the lane's own tensor
// column, each lane gets its output column without any data movement.
// Partial products are shared between the four outgoing messages:
//     E01 = A x0 m0 x1 m1 (layout X)  -> transposed to Y -> out2, out3
//     E23 = A x2 m2 x3 m3 (layout Y)  -> transposed to X -> out0, out1
// (8 mode products + 4 open contractions per slice instead of 16).  The open contraction
// out[a][a'] = sum E~[a..] A[a'..] runs per lane on the VALU (an 8x8x8 product) and is
// reduced across the wave by a halving exchange, so lane l ends with entry l = (a, a').
// Vertices of degree < 4 are padded with dimension-8 ports whose tensor entries vanish
// beyond index 0 and whose message is the identity, and bonds of dimension D < 8 are
// zero-padded to 8 (site entries and incoming messages vanish beyond D, so the padded
// entries of every product are exactly zero and only the D x D block is written back); the
// result is unchanged.
// The two LDS buffers per wave hold A (transposable, row pitch 65 floats: conflict-free
// row and column access) and the transposed partial product.
//
// BF16 mode: every MFMA operand (site entries, messages, partial products) is rounded to
// BF16 (round to nearest even) and products accumulate in FP32; the open contraction
// uses the same BF16-rounded operands with FP32 sums.  Incoming messages are scaled by
// 1 / max|m| before rounding (the scale cancels in the sum normalisation).  Site slices
// are stored as BF16 pairs (r, r+1) per lane so each load is a coalesced 256-byte row.
#pragma once

#include <algorithm>
#include <string>
#include <vector>

#include "common.h"

typedef float v4f __attribute__((ext_vector_type(4)));
typedef short v4s __attribute__((ext_vector_type(4)));

// The plan covers a dispatch bucket: every network whose vertices all have degree <= 4,
// the same physical dimension d (<= 4) and bond dimensions <= 8 on every edge, with the
// network's largest bond dimension D enabled in `mask` (bit D; the autotuner clears the
// bit of a bond dimension that runs faster on the generic kernel).  Work item b (one
// workgroup) is the b-th vertex of the bucket; the remaining networks' message slots go to
// the generic kernel (gen_slots).
struct VertexPlan {
    std::string why;
    int d = 0;                  // physical dimension (waves per workgroup)
    int nv = 0;                 // vertices in the bucket
    int bf16 = 0;               // BF16 operands
    bool all = false;           // the bucket holds every network
    int4 *d_vin = nullptr;      // per item: incoming message slot per padded port (-1: dummy)
    int4 *d_vout = nullptr;     // per item: outgoing message slot per padded port (-1: dummy)
    int4 *d_vdim = nullptr;     // per item: bond dimension per padded port (0: dummy)
    unsigned mask = 0x1FEu;     // bit D: networks of largest bond dimension D may join
    int *d_vnet = nullptr;      // network of each item
    int *d_lead = nullptr;      // per network: its first item in the D <= 8 bucket
    int *d_vlist = nullptr;     // vertex of each item
    int *d_gen_slots = nullptr; // message slots outside the bucket
    std::vector<int> gen_slots; // the same on the host
    int ngen = 0;
    size_t lds_bytes = 0;
    int variant = 0;            // D <= 8 sweep kernel: 0 k_vertex_d8, 1 k_vertex_d8_lo,
                                // 2, 3 k_vertex_d8_mo, 4, 5 k_vertex_d8_hy (FP32)
    // networks whose largest bond dimension is 9..16 in BF16 (k_vertex_d16, vertex16.h) or
    // 9..12 in FP32 (k_vertex_d12, vertex12.h): the large-bond bucket
    int nv16 = 0;
    int4 *d_v16in = nullptr, *d_v16out = nullptr, *d_v16dim = nullptr;
    int *d_v16net = nullptr, *d_v16list = nullptr;
    unsigned short *d_v16sites = nullptr;  // BF16: padded slices, 16^4 entries each
    float *d_v12sites = nullptr;           // FP32: padded slices, 12^4 entries each (2 layouts)
    float *d_v10sites = nullptr;           // FP32: the same at side 10 (items [0, nv16_k10))
    int nv16_k4 = 0;                       // FP32: items [0, nv16_k4) have degree 4, the
                                           // rest degree <= 3 (k_vertex_k3<12>)
    int nv16_k10 = 0;                      // FP32: degree-4 items [0, nv16_k10) have bond
                                           // dimensions <= 10 (k_vertex_d12<10>)
    float *d_v3sites = nullptr;            // FP32, degree <= 3: 12^3 slices (3 layouts)
    // FP32 networks of bond dimension <= 8 with every vertex of degree <= 3 (heavy-hex
    // lattices, trees): the last nv_k3 items of the D <= 8 bucket, swept by k_vertex_k3<8>
    int nv_k3 = 0;
    float *d_vk3sites = nullptr;           // 8^3 slices, 3 layouts
    float *d_vk3part = nullptr;            // per-slice unnormalised messages
    float *d_v16part = nullptr;            // per-slice unnormalised messages
    int v16_waves = 16;                    // waves per slice workgroup: 4, 8 or 16
    float *d_vsites_y = nullptr;           // FP32: the padded slices in layout Y order (the
                                           // bra rows of k_vertex_d8_mo<NS, true>)
    // FP32 networks whose bond dimensions are all <= 4: k_vertex_s<DP> (vertex4.h), one
    // bucket per padded dimension DP = 2, 3, 4 (index DP - 2)
    int ns[3] = {0, 0, 0};
    int *d_vs_ino[3] = {};   // per item and port: offset of the incoming message (-1: padded
                             // port; offsets fit 32 bits, or the bucket is not used)
    int *d_vs_outo[3] = {};  // per item and port: offset of the outgoing message
    int *d_vs_out[3] = {};   // per item and port: outgoing slot (-1: padded port)
    int *d_vs_dim[3] = {};   // per item: the four ports' bond dimensions, one byte each
    int *d_vs_net[3] = {}, *d_vs_list[3] = {};
    float *d_vs_sites[3] = {};
    int ns_total() const { return ns[0] + ns[1] + ns[2]; }
};

static constexpr int VX_D = 8;
static constexpr int VX_PITCH = 65;                 // LDS row pitch (floats)
static constexpr int VX_BUF = 64 * VX_PITCH;        // one transposable 64x64 buffer

static void vertex_plan_release(VertexPlan *p) {
    if (p->d_vin) (void)hipFree(p->d_vin);
    if (p->d_vout) (void)hipFree(p->d_vout);
    if (p->d_vdim) (void)hipFree(p->d_vdim);
    void *v16[12] = {p->d_v16in, p->d_v16out, p->d_v16dim, p->d_v16net, p->d_v16list,
                     p->d_v16sites, p->d_v16part, p->d_v12sites, p->d_v10sites, p->d_v3sites,
                     p->d_vk3sites, p->d_vk3part};
    for (void *x : v16)
        if (x) (void)hipFree(x);
    p->d_v16in = p->d_v16out = p->d_v16dim = nullptr;
    p->d_v16net = p->d_v16list = nullptr;
    p->d_v16sites = nullptr;
    p->d_v12sites = p->d_v10sites = nullptr;
    p->d_v3sites = nullptr;
    p->nv16_k4 = p->nv16_k10 = 0;
    p->d_vk3sites = p->d_vk3part = nullptr;
    p->nv_k3 = 0;
    p->d_v16part = nullptr;
    p->nv16 = 0;
    if (p->d_vsites_y) (void)hipFree(p->d_vsites_y);
    p->d_vsites_y = nullptr;
    for (int k = 0; k < 3; ++k) {
        void *sm[7] = {p->d_vs_ino[k], p->d_vs_outo[k], p->d_vs_out[k], p->d_vs_dim[k],
                       p->d_vs_net[k], p->d_vs_list[k], p->d_vs_sites[k]};
        for (void *x : sm)
            if (x) (void)hipFree(x);
        p->d_vs_ino[k] = p->d_vs_outo[k] = nullptr;
        p->d_vs_out[k] = p->d_vs_dim[k] = nullptr;
        p->d_vs_net[k] = p->d_vs_list[k] = nullptr;
        p->d_vs_sites[k] = nullptr;
        p->ns[k] = 0;
    }
    if (p->d_vnet) (void)hipFree(p->d_vnet);
    if (p->d_lead) (void)hipFree(p->d_lead);
    p->d_lead = nullptr;
    if (p->d_vlist) (void)hipFree(p->d_vlist);
    if (p->d_gen_slots) (void)hipFree(p->d_gen_slots);
    p->d_vin = p->d_vout = p->d_vdim = nullptr;
    p->d_vnet = p->d_vlist = p->d_gen_slots = nullptr;
    p->gen_slots.clear();
    p->nv = p->ngen = 0;
    p->all = false;
}

template <typename T>
static int vx_upload(T **dst, const std::vector<T> &v) {
    if (v.empty()) return 0;
    if (hipMalloc((void **)dst, sizeof(T) * v.size()) != hipSuccess) return 5;
    return hipMemcpy(*dst, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice) == hipSuccess ? 0 : 2;
}

// small != 0: FP32 networks whose bond dimensions are all <= 4 go to the small-bond
// buckets (vertex4.h) instead of the padded D <= 8 kernels.
static int vertex_plan_build(VertexPlan *p, int dtype, int plain, int nv, const int *port_off,
                             const int *port_edge, const int *edge_dim, const int *phys_dim,
                             const long long *site_off, const long long *msg_off,
                             const int *port_rev, const int *vnet, unsigned mask,
                             int small = 0) {
    (void)site_off;
    p->why.clear();
    if (plain) { p->why = "plain (vector-message) networks use the generic path"; return 4; }
    if (dtype != 1 && dtype != 2) { p->why = "the vertex path computes in FP32 or BF16"; return 4; }
    // qualify whole networks (vertices of a network are contiguous); bond dimensions up to 8
    // go to the D <= 8 kernels, larger ones up to 16 (BF16) or 12 (FP32) to the large-bond
    // bucket
    const int dlim = dtype == 2 ? 16 : 12;
    // the small-bond buckets keep message offsets in 32 bits
    const bool off32 = msg_off[port_off[nv]] < (1LL << 31);
    int d = -1;
    std::vector<int> vlist, vlist16, gen_slots, vlist_k3, vlist_s[3];
    p->mask = mask;
    std::string last_reason = "no network has every bond dimension <= 8";
    for (int u0 = 0; u0 < nv;) {
        int u1 = u0;
        while (u1 < nv && vnet[u1] == vnet[u0]) ++u1;
        bool ok = true;
        int dmax = 0, kmax = 0;
        for (int u = u0; u < u1 && ok; ++u) {
            const int k = port_off[u + 1] - port_off[u];
            kmax = std::max(kmax, k);
            if (k > 4) { ok = false; last_reason = "degree > 4"; }
            if (phys_dim[u] < 1 || phys_dim[u] > 4) { ok = false; last_reason = "physical dimension outside 1..4"; }
            if (d >= 0 && phys_dim[u] != d) { ok = false; last_reason = "physical dimension differs between networks"; }
            for (int q = port_off[u]; q < port_off[u + 1] && ok; ++q) {
                const int D = edge_dim[port_edge[q]];
                dmax = std::max(dmax, D);
                if (D > dlim) {
                    ok = false;
                    last_reason = "bond dimension > " + std::to_string(dlim);
                }
            }
        }
        if (ok && dmax > 0 && !((mask >> dmax) & 1u)) {
            ok = false;
            last_reason = "bond dimension " + std::to_string(dmax) +
                          " is routed to the generic kernel (autotuned)";
        }
        if (ok && d < 0) d = phys_dim[u0];
        // FP32 networks of degree <= 3 throughout: their own kernel, at the bucket's end
        const bool k3 = ok && dtype == 1 && dmax <= VX_D && kmax <= 3;
        // FP32 networks of bond dimension <= 4: the small-bond buckets, by padded dimension
        const int sk = (ok && small && dtype == 1 && dmax <= 4 && off32) ? std::max(dmax, 2) - 2 : -1;
        for (int u = u0; u < u1; ++u) {
            if (sk >= 0) vlist_s[sk].push_back(u);
            else if (k3) vlist_k3.push_back(u);
            else if (ok && dmax <= VX_D) vlist.push_back(u);
            else if (ok) vlist16.push_back(u);
            else for (int q = port_off[u]; q < port_off[u + 1]; ++q) gen_slots.push_back(q);
        }
        u0 = u1;
    }
    const int nk3 = (int)vlist_k3.size();
    vlist.insert(vlist.end(), vlist_k3.begin(), vlist_k3.end());
    if (vlist.empty() && vlist16.empty() && vlist_s[0].empty() && vlist_s[1].empty() &&
        vlist_s[2].empty()) {
        p->why = last_reason;
        return 4;
    }
    const int nx = (int)vlist.size();
    std::vector<int4> vin(nx), vout(nx), vdim(nx);
    std::vector<int> vn(nx);
    for (int b = 0; b < nx; ++b) {
        const int u = vlist[b];
        int in[4] = {-1, -1, -1, -1}, out[4] = {-1, -1, -1, -1}, dm[4] = {0, 0, 0, 0};
        for (int j = 0; j < port_off[u + 1] - port_off[u]; ++j) {
            out[j] = port_off[u] + j;
            in[j] = port_rev[port_off[u] + j];
            dm[j] = edge_dim[port_edge[port_off[u] + j]];
        }
        vin[b] = make_int4(in[0], in[1], in[2], in[3]);
        vout[b] = make_int4(out[0], out[1], out[2], out[3]);
        vdim[b] = make_int4(dm[0], dm[1], dm[2], dm[3]);
        vn[b] = vnet[u];
    }
    // the large-bond bucket, laid out the same way; in FP32 the degree-4 items first (those
    // whose bond dimensions are all <= 10 leading, padded to 10), then the degree <= 3 items
    // (their own kernel)
    int n16_k4 = (int)vlist16.size(), n16_k10 = 0;
    if (dtype == 1) {
        auto deg4 = [&](int u) { return port_off[u + 1] - port_off[u] == 4; };
        auto le10 = [&](int u) {
            for (int q = port_off[u]; q < port_off[u + 1]; ++q)
                if (edge_dim[port_edge[q]] > 10) return false;
            return true;
        };
        std::stable_partition(vlist16.begin(), vlist16.end(),
                              [&](int u) { return deg4(u) && le10(u); });
        std::stable_partition(vlist16.begin(), vlist16.end(), deg4);
        n16_k4 = n16_k10 = 0;
        for (int u : vlist16) {
            n16_k4 += deg4(u);
            n16_k10 += deg4(u) && le10(u);
        }
    }
    const int n16 = (int)vlist16.size();
    std::vector<int4> vin16(n16), vout16(n16), vdim16(n16);
    std::vector<int> vn16(n16);
    for (int b = 0; b < n16; ++b) {
        const int u = vlist16[b];
        int in[4] = {-1, -1, -1, -1}, out[4] = {-1, -1, -1, -1}, dm[4] = {0, 0, 0, 0};
        for (int j = 0; j < port_off[u + 1] - port_off[u]; ++j) {
            out[j] = port_off[u] + j;
            in[j] = port_rev[port_off[u] + j];
            dm[j] = edge_dim[port_edge[port_off[u] + j]];
        }
        vin16[b] = make_int4(in[0], in[1], in[2], in[3]);
        vout16[b] = make_int4(out[0], out[1], out[2], out[3]);
        vdim16[b] = make_int4(dm[0], dm[1], dm[2], dm[3]);
        vn16[b] = vnet[u];
    }
    int rc = vx_upload(&p->d_vin, vin);
    if (!rc) rc = vx_upload(&p->d_vout, vout);
    if (!rc) rc = vx_upload(&p->d_vdim, vdim);
    if (!rc) rc = vx_upload(&p->d_v16in, vin16);
    if (!rc) rc = vx_upload(&p->d_v16out, vout16);
    if (!rc) rc = vx_upload(&p->d_v16dim, vdim16);
    if (!rc) rc = vx_upload(&p->d_v16net, vn16);
    if (!rc) rc = vx_upload(&p->d_v16list, vlist16);
    if (!rc && n16 > 0 &&
        hipMalloc((void **)&p->d_v16part, sizeof(float) * (size_t)n16 * d * 4 * 256) != hipSuccess)
        rc = 5;
    if (!rc && nk3 > 0 &&
        hipMalloc((void **)&p->d_vk3part, sizeof(float) * (size_t)nk3 * d * 4 * 256) != hipSuccess)
        rc = 5;
    if (!rc) rc = vx_upload(&p->d_vnet, vn);
    if (!rc && nx > 0) {
        int nn = 0;
        for (int u = 0; u < nv; ++u) nn = std::max(nn, vnet[u] + 1);
        std::vector<int> lead(nn, -1);
        for (int b = 0; b < nx; ++b)
            if (lead[vn[b]] < 0) lead[vn[b]] = b;
        rc = vx_upload(&p->d_lead, lead);
    }
    if (!rc) rc = vx_upload(&p->d_vlist, vlist);
    if (!rc) rc = vx_upload(&p->d_gen_slots, gen_slots);
    for (int k = 0; k < 3 && !rc; ++k) {  // the small-bond buckets
        const int n = (int)vlist_s[k].size();
        std::vector<int> sio(4 * (size_t)n, -1), soo(4 * (size_t)n, 0);
        std::vector<int> so(4 * (size_t)n, -1), sd(n, 0), sn(n);
        for (int b = 0; b < n; ++b) {
            const int u = vlist_s[k][b];
            for (int j = 0; j < port_off[u + 1] - port_off[u]; ++j) {
                const int q = port_off[u] + j;
                so[4 * b + j] = q;
                soo[4 * b + j] = (int)msg_off[q];
                sio[4 * b + j] = (int)msg_off[port_rev[q]];
                sd[b] |= (edge_dim[port_edge[q]] & 255) << (8 * j);
            }
            sn[b] = vnet[u];
        }
        rc = vx_upload(&p->d_vs_ino[k], sio);
        if (!rc) rc = vx_upload(&p->d_vs_outo[k], soo);
        if (!rc) rc = vx_upload(&p->d_vs_out[k], so);
        if (!rc) rc = vx_upload(&p->d_vs_dim[k], sd);
        if (!rc) rc = vx_upload(&p->d_vs_net[k], sn);
        if (!rc) rc = vx_upload(&p->d_vs_list[k], vlist_s[k]);
        p->ns[k] = n;
    }
    if (rc) {
        p->why = "allocation failed";
        vertex_plan_release(p);
        return rc;
    }
    p->d = d;
    p->nv = nx;
    p->nv16 = n16;
    p->nv16_k4 = n16_k4;
    p->nv16_k10 = n16_k10;
    p->nv_k3 = nk3;
    p->gen_slots = gen_slots;
    p->ngen = (int)gen_slots.size();
    p->all = gen_slots.empty();
    p->bf16 = dtype == 2;
    if (!p->all) p->why = last_reason;  // why the rest is not covered
    // per wave: two transposable buffers; per workgroup: 4 messages + d x 4 x 64 partials
    p->lds_bytes = sizeof(float) * ((size_t)d * 2 * VX_BUF + 4 * 64 + (size_t)d * 4 * 64);
    return 0;
}

__device__ __forceinline__ unsigned short bf16_rne(float x) {
    unsigned u = __float_as_uint(x);
    if ((u & 0x7f800000u) == 0x7f800000u) return (unsigned short)((u >> 16) | ((u & 0xffff) ? 0x40 : 0));
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf16_to_f(unsigned short h) {
    return __uint_as_float(((unsigned)h) << 16);
}

// Hardware round-to-nearest-even of two floats into a BF16 pair (v_cvt_pk_bf16_f32).
typedef float vf2 __attribute__((ext_vector_type(2)));
typedef __bf16 vbf2 __attribute__((ext_vector_type(2)));
typedef unsigned vu2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ unsigned bf16x2(float a, float b) {
    const vf2 f = {a, b};
    return __builtin_bit_cast(unsigned, __builtin_convertvector(f, vbf2));
}

__device__ __forceinline__ v4s bf16x4(float a, float b, float c, float d) {
    const vu2 w = {bf16x2(a, b), bf16x2(c, d)};
    return __builtin_bit_cast(v4s, w);
}

// x rounded to BF16 and back, two at a time
__device__ __forceinline__ void bf16_round2(float &a, float &b) {
    const unsigned p = bf16x2(a, b);
    a = __uint_as_float(p << 16);
    b = __uint_as_float(p & 0xffff0000u);
}

// Element e of the padded slice of u (layout X order, e = r * 64 + l): port j's index is
// (e >> (9 - 3 j)) & 7; it is real below the port's bond dimension (padded ports: index 0).
__device__ __forceinline__ float vx_pad_value(const double *sites64, const long long *site_off,
                                              const int *port_off, const int *port_edge,
                                              const int *edge_dim, const int *vlist, int d,
                                              long long us, int e) {
    const int s = (int)(us % d);
    const int u = vlist[us / d];
    const int k = port_off[u + 1] - port_off[u];
    const int idx[4] = {(e >> 9) & 7, (e >> 6) & 7, (e >> 3) & 7, e & 7};
    long long off = 0, n = 1;
    for (int j = 0; j < 4; ++j) {
        if (j < k) {
            const int D = edge_dim[port_edge[port_off[u] + j]];
            if (idx[j] >= D) return 0.f;
            off = off * D + idx[j];
            n *= D;
        } else if (idx[j] != 0) {
            return 0.f;
        }
    }
    return (float)sites64[site_off[u] + s * n + off];
}

// FP32: vsites[u][s][r * 64 + l] (layout X) and, when vsy is given, vsy[u][s][l * 64 + r]
// (layout Y).  BF16: vsites[u][s][(r / 2) * 128 + 2 l + (r % 2)].
__global__ void k_vertex_pad_sites(const double *sites64, const long long *site_off,
                                   const int *port_off, const int *port_edge,
                                   const int *edge_dim, const int *vlist, int d, int nv, int bf16,
                                   void *vsites, float *vsy) {
    const long long total = (long long)nv * d * 4096;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
         i += (long long)gridDim.x * blockDim.x) {
        const int e = (int)(i & 4095);
        const long long us = i >> 12;
        const float v = vx_pad_value(sites64, site_off, port_off, port_edge, edge_dim, vlist, d,
                                     us, e);
        if (bf16) {
            const int r = e >> 6, l = e & 63;
            ((unsigned short *)vsites)[(us << 12) + (r >> 1) * 128 + 2 * l + (r & 1)] = bf16_rne(v);
        } else {
            ((float *)vsites)[i] = v;
            if (vsy) vsy[(us << 12) + (long long)(e & 63) * 64 + (e >> 6)] = v;
        }
    }
}

static int vertex_upload_sites(VertexPlan *p, hipStream_t st, const double *d_sites64,
                               void **d_vsites, const long long *d_site_off,
                               const int *d_port_off, const int *d_port_edge,
                               const int *d_edge_dim) {
    const size_t bytes = (p->bf16 ? 2 : 4) * (size_t)p->nv * p->d * 4096;
    if (*d_vsites) (void)hipFree(*d_vsites);
    *d_vsites = nullptr;
    if (p->nv == 0) return 0;
    if (hipMalloc(d_vsites, bytes) != hipSuccess) return 5;
    if (p->d_vsites_y) (void)hipFree(p->d_vsites_y);
    p->d_vsites_y = nullptr;
    if (!p->bf16 && hipMalloc((void **)&p->d_vsites_y, bytes) != hipSuccess) return 5;
    hipLaunchKernelGGL(k_vertex_pad_sites, dim3(2048), dim3(256), 0, st, d_sites64, d_site_off,
                       d_port_off, d_port_edge, d_edge_dim, p->d_vlist, p->d, p->nv, p->bf16,
                       *d_vsites, p->d_vsites_y);
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

// ---------------------------------------------------------------- device helpers

// Mode product on a register port of an 8x8 register block T[hi*8 + lo]:
//   HI: O[b*8 + lo] = sum_a T[a*8 + lo] m[a][b]     LO: O[hi*8 + b] = sum_a T[hi*8 + a] m[a][b]
// FP32: M[a*2 + h] holds m[a][4h + lane%4].
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

// BF16: Mb[g*2 + h] packs m[4g + t][4h + lane%4], t = 0..3 (g = 0, 1 halves of a).
template <bool HI>
__device__ __forceinline__ void vx_mp_bf16(const float (&T)[64], float (&O)[64], const v4s (&Mb)[4]) {
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        v4s bq[2];
#pragma unroll
        for (int g = 0; g < 2; ++g) {
            float v[4];
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int a = 4 * g + t;
                v[t] = HI ? T[a * 8 + x] : T[x * 8 + a];
            }
            bq[g] = bf16x4(v[0], v[1], v[2], v[3]);
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            v4f acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int g = 0; g < 2; ++g)
                acc = __builtin_amdgcn_mfma_f32_4x4x4bf16_1k(Mb[g * 2 + h], bq[g], acc, 0, 0, 0);
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

__device__ __forceinline__ void vx_load_m_bf16(const float *msg, v4s (&Mb)[4]) {
    const int c = threadIdx.x & 3;
#pragma unroll
    for (int g = 0; g < 2; ++g)
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const float *p = msg + 4 * g * 8 + 4 * h + c;
            Mb[g * 2 + h] = bf16x4(p[0], p[8], p[16], p[24]);
        }
}

// Mode product in the precision of the kernel.
template <bool HI, bool BF>
__device__ __forceinline__ void vx_mode(const float (&T)[64], float (&O)[64], const float *msg) {
    if constexpr (BF) {
        v4s Mb[4];
        vx_load_m_bf16(msg, Mb);
        vx_mp_bf16<HI>(T, O, Mb);
    } else {
        float M[16];
        vx_load_m(msg, M);
        vx_mp<HI>(T, O, M);
    }
}

// One step of a halving exchange inside 16-lane rows: lane l keeps entries i + S * (bit S
// of l) for i < S and adds the same entries from its DPP partner (row_partner<S>).
template <int S, int N>
__device__ __forceinline__ void vx_halve(float (&v)[N], int lane) {
    const bool up = (lane & S) != 0;
#pragma unroll
    for (int i = 0; i < S; ++i) {
        const float lo = v[i], hi = v[i + S];
        const float send = up ? lo : hi;
        const float keep = up ? hi : lo;
        v[i] = keep + row_partner<S>(send);
    }
}

// Sum v[i] over the 64 lanes for every i; lane l returns entry l (halving exchange).
// Step 1 pairs lane l with l^32 through v_permlane32_swap (lanes 32-63 of the first
// operand trade places with lanes 0-31 of the second, so x' + y' holds entry i on the
// low half and entry i+32 on the high half); step 2 does the same for 16-lane rows with
// v_permlane16_swap; the last four steps exchange inside 16-lane rows through DPP.  Values are
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
    vx_halve<8>(v, lane);
    vx_halve<4>(v, lane);
    vx_halve<2>(v, lane);
    vx_halve<1>(v, lane);
    return v[0];
}

// Open contraction with the bra tensor read from the wave's A buffer.
//   HI (open port = register hi index): out[a][a'] = sum_lo G[a*8+lo] A[a'*8+lo]
//   LO (open port = register lo index): out[a][a'] = sum_hi G[hi*8+a] A[hi*8+a']
// XL: the A buffer is read in layout X (else Y).  Returns entry (lane/8, lane%8).  In BF16
// mode the partial product G is rounded to BF16 first (the buffer already holds BF16 values).
template <bool HI, bool XL, int PITCH>
__device__ __forceinline__ float vx_final_acc(const float (&G)[64], const float *bufA) {
    const int lane = threadIdx.x & 63;
    float acc[64];
#pragma unroll
    for (int i = 0; i < 64; ++i) acc[i] = 0.f;
#pragma unroll
    for (int r = 0; r < 64; ++r) {
        const float av = XL ? bufA[r * PITCH + lane] : bufA[lane * PITCH + r];
        const int hi = r >> 3, lo = r & 7;
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            if (HI) acc[a * 8 + hi] += G[a * 8 + lo] * av;   // a' = hi
            else acc[a * 8 + lo] += G[hi * 8 + a] * av;      // a' = lo
        }
    }
    return vx_reduce64(acc);
}

// The open contraction in two halves of 32 accumulators (a < 4, then a >= 4), for half the
// registers of vx_final_acc at twice the reads of the bra tensor.  Each half takes the first two
// halving steps on its own (lane bit 5 against local entry bit 4 through v_permlane32_swap,
// lane bit 4 against bit 3 through v_permlane16_swap), which leaves 8 values per half; the
// two halves' 16 values then share the four DPP steps, lane bit 3 selecting the half.  Lane l
// ends with entry vx_split_entry(l) = 32 b3 + 16 b5 + 8 b4 + l % 8 (b_k = bit k of l).
__device__ __forceinline__ int vx_split_entry(int lane) {
    return ((lane & 8) << 2) | ((lane & 32) >> 1) | ((lane & 16) >> 1) | (lane & 7);
}

template <bool HI, bool XL, int PITCH>
__device__ __forceinline__ float vx_final_half(const float (&G)[64], const float *bufA) {
    const int lane = threadIdx.x & 63;
    float hv[16];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
        float acc[32];
#pragma unroll
        for (int i = 0; i < 32; ++i) acc[i] = 0.f;
#pragma unroll
        for (int r = 0; r < 64; ++r) {
            const float av = XL ? bufA[r * PITCH + lane] : bufA[lane * PITCH + r];
            const int hi = r >> 3, lo = r & 7;
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                const int ag = 4 * h + a;
                if (HI) acc[a * 8 + hi] += G[ag * 8 + lo] * av;
                else acc[a * 8 + lo] += G[hi * 8 + ag] * av;
            }
        }
        // lanes with bit 5 clear keep local entries i, the others i + 16
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(acc[i]),
                                                            __float_as_uint(acc[i + 16]), false,
                                                            false);
            acc[i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
        }
        // lanes with bit 4 clear keep i, the others i + 8
#pragma unroll
        for (int i = 0; i < 8; ++i) {
            const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(acc[i]),
                                                            __float_as_uint(acc[i + 8]), false,
                                                            false);
            hv[h * 8 + i] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
        }
    }
    vx_halve<8>(hv, lane);
    vx_halve<4>(hv, lane);
    vx_halve<2>(hv, lane);
    vx_halve<1>(hv, lane);
    return hv[0];
}

template <bool HI, bool XL, bool BF, int PITCH = VX_PITCH>
__device__ __forceinline__ float vx_final(const float (&Gin)[64], const float *bufA) {
    if constexpr (BF) {
        float G[64];
#pragma unroll
        for (int i = 0; i < 64; i += 2) {
            G[i] = Gin[i];
            G[i + 1] = Gin[i + 1];
            bf16_round2(G[i], G[i + 1]);
        }
        return vx_final_acc<HI, XL, PITCH>(G, bufA);
    } else {
        return vx_final_acc<HI, XL, PITCH>(Gin, bufA);
    }
}

// Whether network `net` is stopped for this sweep.  With fold.prev set, the previous sweep's
// stopping rule (StopWhenConverged on the maximum of its FOLD_SLOTS accumulators) is applied
// first: every wave of every work item of the network computes the same decision (a racing
// read of the lead's done flag sees 0 or 1, and either way ends stopped), and the lead item
// records the iteration, delta, done and at.  StopAfterIteration is left to the kernel that
// ends the sequence: no sequence runs past maxiter.  The lead item also clears fold.clear.
__device__ __forceinline__ bool vx_fold_begin(const NetCtl &ctl, const SweepFold &fold, int net,
                                              int u) {
    const int lane = threadIdx.x & 63;
    const bool lead = fold.lead && fold.lead[net] == u;
    if (lead && fold.clear && threadIdx.x < FOLD_SLOTS)
        fold.clear[(long long)net * FOLD_SLOTS + threadIdx.x] = 0ull;
    const bool was = ctl.enabled && ctl.done[net];
    if (!fold.prev) return was;
    const unsigned long long b =
        lane < FOLD_SLOTS ? fold.prev[(long long)net * FOLD_SLOTS + lane] : 0ull;
    const double m = wave_max_v(__longlong_as_double((long long)b));
    const double tol = fold.rp->tol;
    const bool newly = !was && tol >= 0.0 && m < tol;
    if (lead && threadIdx.x == 0 && !was) {
        const int it = fold.iter[net] + 1;
        fold.iter[net] = it;
        fold.delta[net] = m;
        if (newly) {
            ctl.done[net] = 1;
            fold.at[net] = it;
        }
    }
    return was || newly;
}

// ---------------------------------------------------------------- the sweep kernel

template <int NS, bool BF>
__global__ __launch_bounds__(64 * NS) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_vertex_d8(const void *__restrict__ vsites_v, const float *__restrict__ mold,
                 float *__restrict__ mnew, const int4 *__restrict__ vin,
                 const int4 *__restrict__ vout, const int4 *__restrict__ vdim,
                 const int *__restrict__ vnet, const long long *__restrict__ msg_off, NetCtl ctl,
                 double *__restrict__ res_fid, double *__restrict__ res_abs, SweepFold fold) {
    extern __shared__ __align__(16) float vsm[];
    __shared__ double fold_m[4];
    const int u = blockIdx.x;  // work item: the u-th vertex of the bucket
    const int w = threadIdx.x >> 6;     // wave = physical index s
    const int lane = threadIdx.x & 63;
    float *bufA = vsm + w * 2 * VX_BUF;
    float *bufB = bufA + VX_BUF;
    float *msm = vsm + NS * 2 * VX_BUF;  // 4 x 64 incoming messages
    float *part = msm + 4 * 64;          // [s][port][64] per-slice results
    const int4 vi = vin[u], vo = vout[u], vdm = vdim[u];
    const int ins[4] = {vi.x, vi.y, vi.z, vi.w};
    const int outs[4] = {vo.x, vo.y, vo.z, vo.w};
    const int dims[4] = {vdm.x, vdm.y, vdm.z, vdm.w};
    // message offsets (other networks of a batch may have other bond dimensions)
    long long ino[4], outo[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        ino[j] = ins[j] >= 0 ? msg_off[ins[j]] : 0;
        outo[j] = outs[j] >= 0 ? msg_off[outs[j]] : 0;
    }
    // lane l holds entry (a, a') = (l / 8, l % 8) of an 8 x 8 block
    const int la = lane >> 3, lb = lane & 7;

    const int net = vnet[u];
    if (vx_fold_begin(ctl, fold, net, u)) {  // uniform over the workgroup
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j % NS != w || outs[j] < 0) continue;
            if (lane < dims[j] * dims[j]) mnew[outo[j] + lane] = mold[outo[j] + lane];
            if (lane == 0) { res_fid[outs[j]] = 0.0; res_abs[outs[j]] = 0.0; }
        }
        return;
    }
    // incoming messages zero-padded to 8 x 8 (identity on padded ports); port indices are
    // compile-time so the slot arrays stay in registers.  BF16 mode scales each message by
    // 1 / max|m|.
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j % NS != w) continue;
        const int D = dims[j];
        float x;
        if (ins[j] < 0) x = la == lb ? 1.f : 0.f;
        else x = (la < D && lb < D) ? mold[ino[j] + la * D + lb] : 0.f;
        if constexpr (BF) {
            const float mx = wave_max(fabsf(x));
            x = mx > 0.f ? x / mx : x;
        }
        msm[j * 64 + lane] = x;
    }
    // slice A[s] in layout X
    float T[64], E[64], G[64];
    if constexpr (BF) {
        const unsigned *src = (const unsigned *)vsites_v + ((long long)u * NS + w) * 2048;
#pragma unroll
        for (int r2 = 0; r2 < 32; ++r2) {
            const unsigned pr = src[r2 * 64 + lane];
            T[2 * r2] = bf16_to_f((unsigned short)(pr & 0xffffu));
            T[2 * r2 + 1] = bf16_to_f((unsigned short)(pr >> 16));
        }
    } else {
        const float *src = (const float *)vsites_v + ((long long)u * NS + w) * 4096;
#pragma unroll
        for (int r = 0; r < 64; ++r) T[r] = src[r * 64 + lane];
    }
#pragma unroll
    for (int r = 0; r < 64; ++r) bufA[r * VX_PITCH + lane] = T[r];
    __syncthreads();
    float out[4];

    // ---- phase A: E01 = A x0 m0 x1 m1 (X) -> Y; out2 (open p2), out3 (open p3)
    vx_mode<true, BF>(T, E, msm + 0 * 64);
    vx_mode<false, BF>(E, G, msm + 1 * 64);
#pragma unroll
    for (int r = 0; r < 64; ++r) bufB[r * VX_PITCH + lane] = G[r];
    wave_lds_sync();  // bufB belongs to this wave
#pragma unroll
    for (int r = 0; r < 64; ++r) E[r] = bufB[lane * VX_PITCH + r];  // E01 in Y
    vx_mode<false, BF>(E, G, msm + 3 * 64);
    out[2] = vx_final<true, false, BF>(G, bufA);
    vx_mode<true, BF>(E, G, msm + 2 * 64);
    out[3] = vx_final<false, false, BF>(G, bufA);

    // ---- phase B: E23 = A x2 m2 x3 m3 (Y) -> X; out0 (open p0), out1 (open p1)
#pragma unroll
    for (int r = 0; r < 64; ++r) T[r] = bufA[lane * VX_PITCH + r];  // A in Y
    vx_mode<true, BF>(T, E, msm + 2 * 64);
    vx_mode<false, BF>(E, G, msm + 3 * 64);
    wave_lds_sync();  // every lane of the wave has finished reading bufB (phase A)
#pragma unroll
    for (int r = 0; r < 64; ++r) bufB[lane * VX_PITCH + r] = G[r];  // write from Y
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 64; ++r) E[r] = bufB[r * VX_PITCH + lane];  // E23 in X
    vx_mode<false, BF>(E, G, msm + 1 * 64);
    out[0] = vx_final<true, true, BF>(G, bufA);
    vx_mode<true, BF>(E, G, msm + 0 * 64);
    out[1] = vx_final<false, true, BF>(G, bufA);

    // ---- sum over s, normalise, residual, write
    double fm = 0.0;  // this wave's largest residual under the run's criterion (lane 0)
    const bool fcrit = fold.acc && fold.rp->crit != 0;
#pragma unroll
    for (int j = 0; j < 4; ++j) part[(w * 4 + j) * 64 + lane] = out[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int q = outs[j];
        if (j % NS != w || q < 0) continue;
        float x = 0.f;
#pragma unroll
        for (int s = 0; s < NS; ++s) x += part[(s * 4 + j) * 64 + lane];
        const double S = wave_sum_v((double)x);
        const double xn = S != 0.0 ? (double)x / S : (double)x;
        const float xf = (float)xn;  // exactly 0 on padded entries
        const int D = dims[j];
        const bool real = la < D && lb < D;
        const long long o = outo[j] + la * D + lb;
        const double xo = real ? (double)mold[o] : 0.0;
        if (real) mnew[o] = xf;
        const double dn = wave_sum_v((double)xf * xo),
                     nn = wave_sum_v((double)xf * (double)xf), oo = wave_sum_v(xo * xo);
        const double mx = wave_max_v(fabs((double)xf - xo));
        if (lane == 0) {
            const double c = dn / sqrt(nn * oo);
            res_fid[q] = 1.0 - c * c;
            res_abs[q] = mx;
            fm = fold_max(fm, fcrit ? mx : 1.0 - c * c);
        }
    }
    if (fold.acc) {
        if (lane == 0) fold_m[w] = fm;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = 0.0;
            for (int s = 0; s < NS; ++s) m = fold_max(m, fold_m[s]);
            atomicMax(fold.acc + (long long)net * FOLD_SLOTS + (u % FOLD_SLOTS),
                      (unsigned long long)__double_as_longlong(m));
        }
    }
}

// ---------------------------------------------------------------- MFMA open contractions
//
// Variant 2 of the D <= 8 FP32 sweep (k_vertex_d8_mo): the open contractions run on the
// MFMA pipe too.  out[a][a'] = sum_K G[a][K] A[a'][K] over the K = 512 values of the three
// other ports is, per v_mfma_f32_4x4x1_16b_f32, 16 blocks of 4x4 outer products: block k
// (lanes 4k..4k+3) takes one K value per instruction, lane 4k+q supplying G[4t+q][K] as the
// A operand and A[4t'+q][K] as the B operand, so register i of lane 4k+j accumulates
// out[4t+i][4t'+j] over the K values of block k (K = 16c + k, chunk c = 0..31); the four
// 4x4 tiles (t, t') and a sum over the 16 blocks finish it.  The operands need K spread over
// lane blocks, but the registers hold K along one register port and the lanes, so G is first
// written to LDS as gbuf[K][a] (K = o * 64 + lane, o the register index that is not the open
// port; a stored in the order 0,4,1,5,2,6,3,7 so that rows q and q + 4 are one 8-byte read)
// and the bra values are gathered from the wave's copy of A in LDS.  Per slice that is 512
// MFMAs and 96 reduction instructions per open instead of 2048 VALU FMAs and about 1500
// reduction instructions.

// G (registers; open index a = register hi (HI) or lo (LO)) to gbuf[(o*64 + lane)*8 + perm(a)].
template <bool HI>
__device__ __forceinline__ void vmo_stage(const float (&G)[64], float *gbuf) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 0; o < 8; ++o) {
        float v[8];
#pragma unroll
        for (int a = 0; a < 8; ++a) v[2 * (a & 3) + (a >> 2)] = HI ? G[a * 8 + o] : G[o * 8 + a];
        float4 *dst = reinterpret_cast<float4 *>(gbuf + (o * 64 + lane) * 8);
        dst[0] = make_float4(v[0], v[1], v[2], v[3]);
        dst[1] = make_float4(v[4], v[5], v[6], v[7]);
    }
}

template <int N>
__device__ __forceinline__ float vmo_ror(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x120 + N, 0xF, 0xF, false));
}

// Open contraction from the staged G (gbuf) and the bra A (bufA, the slice stored as
// bufA[(p0*8 + p1) * VX_PITCH + p2*8 + p3]).  OPEN names the open port and the layout of G:
// 2 (layout Y, port 2 = register hi), 3 (Y, port 3 = register lo), 0 (X, port 0 = hi),
// 1 (X, port 1 = lo).  Entry (a, a') of the slice's result goes to out[a*8 + a'] (LDS),
// written by lanes 0..3.
//
// GB (bra from global): `bufA` is instead the slice's padded copy in HBM in the layout of G
// (layout Y for OPEN 2, 3; layout X for OPEN 0, 1; row r of lane l at [r * 64 + l]), so the
// bra element (a', K) sits at (a' * 8 + K / 64) * 64 + K % 64 for an open on the register hi
// port and at ((K / 64) * 8 + a') * 64 + K % 64 on the lo port: four 64-byte runs per load.
template <int OPEN, bool GB = false>
__device__ __forceinline__ void vmo_open(const float *gbuf, const float *bufA, float *out) {
    static_assert(VX_PITCH == 65, "bra address offsets assume a row pitch of 65");
    const int lane = threadIdx.x & 63;
    const int k = lane >> 2, q = lane & 3;
    constexpr bool HIO = OPEN == 0 || OPEN == 2;
    // bra element (a' = q, K = 16c + k): K % 64 = 16 (c % 4) + k, K / 64 = c / 4
    const int base = GB ? (HIO ? q * 512 + k : q * 64 + k)
                   : OPEN == 2 ? k * 65 + q * 8 : OPEN == 3 ? k * 65 + q
                   : OPEN == 0 ? q * 520 + k : q * 65 + k;
    constexpr int D4 = GB ? (HIO ? 2048 : 256)
                     : OPEN == 2 ? 32 : OPEN == 3 ? 4 : OPEN == 0 ? 2080 : 260;  // a' + 4
    const float *bp = bufA + base;
    const float *gp = gbuf + k * 8 + 2 * q;
    v4f acc00 = {0.f, 0.f, 0.f, 0.f}, acc01 = acc00, acc10 = acc00, acc11 = acc00;
#pragma unroll
    for (int c = 0; c < 32; ++c) {
        const int coff = GB ? (HIO ? (c >> 2) * 64 + (c & 3) * 16 : (c >> 2) * 512 + (c & 3) * 16)
                       : OPEN == 2 ? (c & 3) * 1040 + (c >> 2)
                       : OPEN == 3 ? (c & 3) * 1040 + (c >> 2) * 8
                       : OPEN == 0 ? (c >> 2) * 65 + (c & 3) * 16
                                   : (c >> 2) * 520 + (c & 3) * 16;
        const float2 g = *reinterpret_cast<const float2 *>(gp + c * 128);
        const float b0 = bp[coff], b1 = bp[coff + D4];
        acc00 = __builtin_amdgcn_mfma_f32_4x4x1f32(g.x, b0, acc00, 0, 0, 0);
        acc01 = __builtin_amdgcn_mfma_f32_4x4x1f32(g.x, b1, acc01, 0, 0, 0);
        acc10 = __builtin_amdgcn_mfma_f32_4x4x1f32(g.y, b0, acc10, 0, 0, 0);
        acc11 = __builtin_amdgcn_mfma_f32_4x4x1f32(g.y, b1, acc11, 0, 0, 0);
    }
    float v[16];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        v[i] = acc00[i];
        v[4 + i] = acc01[i];
        v[8 + i] = acc10[i];
        v[12 + i] = acc11[i];
    }
    // sum over the 16 blocks: lane bits 2 and 3 by rotations inside 16-lane rows, bits 4
    // and 5 by the row and half swaps (each leaves the total on every lane)
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        float x = v[i] + vmo_ror<4>(v[i]);
        x += vmo_ror<8>(x);
        const auto r16 = __builtin_amdgcn_permlane16_swap(__float_as_uint(x), __float_as_uint(x),
                                                          false, false);
        x = __uint_as_float(r16[0]) + __uint_as_float(r16[1]);
        const auto r32 = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(x),
                                                          false, false);
        v[i] = __uint_as_float(r32[0]) + __uint_as_float(r32[1]);
    }
    if (lane < 4) {  // lane j holds out[4t + i][4t' + j] in v[(2t + t') * 4 + i]
#pragma unroll
        for (int tt = 0; tt < 4; ++tt)
#pragma unroll
            for (int i = 0; i < 4; ++i)
                out[(4 * (tt >> 1) + i) * 8 + 4 * (tt & 1) + lane] = v[tt * 4 + i];
    }
}

// GB = false (variant 2): the bra A lives in a second LDS buffer per wave (one wave per
// SIMD).  GB = true (variant 3): the bra rows are gathered from the padded X and Y copies of
// the sites in HBM (L2 hits after the first), so one LDS buffer per wave suffices and two
// waves share a SIMD.
template <int NS, bool GB>
__global__ __launch_bounds__(64 * NS) __attribute__((amdgpu_waves_per_eu(GB ? 2 : 1, GB ? 2 : 1)))
void k_vertex_d8_mo(const void *__restrict__ vsites_v, const float *__restrict__ vsites_y,
                    const float *__restrict__ mold,
                    float *__restrict__ mnew, const int4 *__restrict__ vin,
                    const int4 *__restrict__ vout, const int4 *__restrict__ vdim,
                    const int *__restrict__ vnet, const long long *__restrict__ msg_off,
                    NetCtl ctl, double *__restrict__ res_fid, double *__restrict__ res_abs,
                    SweepFold fold) {
    extern __shared__ __align__(16) float vsm[];
    __shared__ double fold_m[4];
    const int u = blockIdx.x;
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    float *bufA = vsm + w * (GB ? 1 : 2) * VX_BUF;  // A (written once; unused with GB)
    float *bufT = GB ? bufA : bufA + VX_BUF;        // transposes, then each open's staged G
    float *msm = vsm + NS * (GB ? 1 : 2) * VX_BUF;
    float *part = msm + 4 * 64;
    const int4 vi = vin[u], vo = vout[u], vdm = vdim[u];
    const int ins[4] = {vi.x, vi.y, vi.z, vi.w};
    const int outs[4] = {vo.x, vo.y, vo.z, vo.w};
    const int dims[4] = {vdm.x, vdm.y, vdm.z, vdm.w};
    long long ino[4], outo[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        ino[j] = ins[j] >= 0 ? msg_off[ins[j]] : 0;
        outo[j] = outs[j] >= 0 ? msg_off[outs[j]] : 0;
    }
    const int la = lane >> 3, lb = lane & 7;
    const int net = vnet[u];
    if (vx_fold_begin(ctl, fold, net, u)) {  // uniform over the workgroup
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j % NS != w || outs[j] < 0) continue;
            if (lane < dims[j] * dims[j]) mnew[outo[j] + lane] = mold[outo[j] + lane];
            if (lane == 0) { res_fid[outs[j]] = 0.0; res_abs[outs[j]] = 0.0; }
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j % NS != w) continue;
        const int D = dims[j];
        float x;
        if (ins[j] < 0) x = la == lb ? 1.f : 0.f;
        else x = (la < D && lb < D) ? mold[ino[j] + la * D + lb] : 0.f;
        msm[j * 64 + lane] = x;
    }
    float T[64], E[64], G[64];
    const float *sx = (const float *)vsites_v + ((long long)u * NS + w) * 4096;
    const float *sy = GB ? vsites_y + ((long long)u * NS + w) * 4096 : nullptr;
#pragma unroll
    for (int r = 0; r < 64; ++r) T[r] = sx[r * 64 + lane];
    if constexpr (!GB) {
#pragma unroll
        for (int r = 0; r < 64; ++r) bufA[r * VX_PITCH + lane] = T[r];
    }
    __syncthreads();  // messages of every wave, and this wave's A
    float *po = part + w * 4 * 64;
    const float *braY = GB ? sy : bufA;  // bra of the phase-A opens (layout Y)
    const float *braX = GB ? sx : bufA;  // bra of the phase-B opens (layout X)

    // ---- phase A: E01 = A x0 m0 x1 m1 (X) -> Y; out2 (open p2), out3 (open p3)
    vx_mode<true, false>(T, E, msm + 0 * 64);
    vx_mode<false, false>(E, G, msm + 1 * 64);
#pragma unroll
    for (int r = 0; r < 64; ++r) bufT[r * VX_PITCH + lane] = G[r];
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 64; ++r) E[r] = bufT[lane * VX_PITCH + r];  // E01 in Y
    vx_mode<false, false>(E, G, msm + 3 * 64);
    wave_lds_sync();  // every lane has read E01 from bufT
    vmo_stage<true>(G, bufT);
    wave_lds_sync();
    vmo_open<2, GB>(bufT, braY, po + 2 * 64);
    vx_mode<true, false>(E, G, msm + 2 * 64);
    wave_lds_sync();  // the staged G of out2 is consumed
    vmo_stage<false>(G, bufT);
    wave_lds_sync();
    vmo_open<3, GB>(bufT, braY, po + 3 * 64);

    // ---- phase B: E23 = A x2 m2 x3 m3 (Y) -> X; out0 (open p0), out1 (open p1)
    if constexpr (GB) {
#pragma unroll
        for (int r = 0; r < 64; ++r) T[r] = sy[r * 64 + lane];  // A in Y
    } else {
#pragma unroll
        for (int r = 0; r < 64; ++r) T[r] = bufA[lane * VX_PITCH + r];  // A in Y
    }
    vx_mode<true, false>(T, E, msm + 2 * 64);
    vx_mode<false, false>(E, G, msm + 3 * 64);
    wave_lds_sync();  // the staged G of out3 is consumed
#pragma unroll
    for (int r = 0; r < 64; ++r) bufT[lane * VX_PITCH + r] = G[r];
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 64; ++r) E[r] = bufT[r * VX_PITCH + lane];  // E23 in X
    vx_mode<false, false>(E, G, msm + 1 * 64);
    wave_lds_sync();
    vmo_stage<true>(G, bufT);
    wave_lds_sync();
    vmo_open<0, GB>(bufT, braX, po + 0 * 64);
    vx_mode<true, false>(E, G, msm + 0 * 64);
    wave_lds_sync();
    vmo_stage<false>(G, bufT);
    wave_lds_sync();
    vmo_open<1, GB>(bufT, braX, po + 1 * 64);

    // ---- sum over s, normalise, residual, write (as k_vertex_d8)
    double fm = 0.0;
    const bool fcrit = fold.acc && fold.rp->crit != 0;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int q = outs[j];
        if (j % NS != w || q < 0) continue;
        float x = 0.f;
#pragma unroll
        for (int s = 0; s < NS; ++s) x += part[(s * 4 + j) * 64 + lane];
        const double S = wave_sum_v((double)x);
        const double xn = S != 0.0 ? (double)x / S : (double)x;
        const float xf = (float)xn;
        const int D = dims[j];
        const bool real = la < D && lb < D;
        const long long o = outo[j] + la * D + lb;
        const double xo = real ? (double)mold[o] : 0.0;
        if (real) mnew[o] = xf;
        const double dn = wave_sum_v((double)xf * xo),
                     nn = wave_sum_v((double)xf * (double)xf), oo = wave_sum_v(xo * xo);
        const double mx = wave_max_v(fabs((double)xf - xo));
        if (lane == 0) {
            const double c = dn / sqrt(nn * oo);
            res_fid[q] = 1.0 - c * c;
            res_abs[q] = mx;
            fm = fold_max(fm, fcrit ? mx : 1.0 - c * c);
        }
    }
    if (fold.acc) {
        if (lane == 0) fold_m[w] = fm;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = 0.0;
            for (int s = 0; s < NS; ++s) m = fold_max(m, fold_m[s]);
            atomicMax(fold.acc + (long long)net * FOLD_SLOTS + (u % FOLD_SLOTS),
                      (unsigned long long)__double_as_longlong(m));
        }
    }
}

// BF16 rounding (nearest even) of a partial product in place, before an open contraction.
__device__ __forceinline__ void vx_round_bf16(float (&G)[64]) {
#pragma unroll
    for (int i = 0; i < 64; i += 2) bf16_round2(G[i], G[i + 1]);
}

// ---------------------------------------------------------------- low-LDS variant (FP32 and BF16)
//
// The same sweep with one LDS buffer per wave instead of two (36 KB per two-wave workgroup,
// so four workgroups fit a CU) and a 256-register budget, i.e. two waves per SIMD: the
// buffer first carries the E01 transpose, then a copy of A for phase A's bra (layout Y) and
// phase B's ket (layout Y), then the E23 transpose, then A again (kept in registers across
// phase B's mode products) for phase B's bra (layout X).  The open contractions accumulate
// in two halves of 32 registers (vx_final_half) and each result goes to LDS as soon as it
// is made, which keeps the kernel within the budget (a few values spill, once per wave).
template <int NS, bool BF>
__global__ __launch_bounds__(64 * NS) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_vertex_d8_lo(const void *__restrict__ vsites_v, const float *__restrict__ mold,
                    float *__restrict__ mnew, const int4 *__restrict__ vin,
                    const int4 *__restrict__ vout, const int4 *__restrict__ vdim,
                    const int *__restrict__ vnet, const long long *__restrict__ msg_off,
                    NetCtl ctl, double *__restrict__ res_fid, double *__restrict__ res_abs,
                    SweepFold fold) {
    extern __shared__ __align__(16) float vsm[];
    __shared__ double fold_m[4];
    const int u = blockIdx.x;
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    float *buf = vsm + w * VX_BUF;
    float *msm = vsm + NS * VX_BUF;      // 4 x 64 incoming messages
    float *part = msm + 4 * 64;          // [s][port][64] per-slice results
    const int4 vi = vin[u], vo = vout[u], vdm = vdim[u];
    const int ins[4] = {vi.x, vi.y, vi.z, vi.w};
    const int outs[4] = {vo.x, vo.y, vo.z, vo.w};
    const int dims[4] = {vdm.x, vdm.y, vdm.z, vdm.w};
    long long ino[4], outo[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        ino[j] = ins[j] >= 0 ? msg_off[ins[j]] : 0;
        outo[j] = outs[j] >= 0 ? msg_off[outs[j]] : 0;
    }
    const int la = lane >> 3, lb = lane & 7;
    const int net = vnet[u];
    if (vx_fold_begin(ctl, fold, net, u)) {  // uniform over the workgroup
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j % NS != w || outs[j] < 0) continue;
            if (lane < dims[j] * dims[j]) mnew[outo[j] + lane] = mold[outo[j] + lane];
            if (lane == 0) { res_fid[outs[j]] = 0.0; res_abs[outs[j]] = 0.0; }
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j % NS != w) continue;
        const int D = dims[j];
        float x;
        if (ins[j] < 0) x = la == lb ? 1.f : 0.f;
        else x = (la < D && lb < D) ? mold[ino[j] + la * D + lb] : 0.f;
        if constexpr (BF) {
            const float mx = wave_max(fabsf(x));
            x = mx > 0.f ? x / mx : x;
        }
        msm[j * 64 + lane] = x;
    }
    float T[64], E[64], G[64];
    if constexpr (BF) {
        const unsigned *src = (const unsigned *)vsites_v + ((long long)u * NS + w) * 2048;
#pragma unroll
        for (int r2 = 0; r2 < 32; ++r2) {
            const unsigned pr = src[r2 * 64 + lane];
            T[2 * r2] = bf16_to_f((unsigned short)(pr & 0xffffu));
            T[2 * r2 + 1] = bf16_to_f((unsigned short)(pr >> 16));
        }
    } else {
        const float *src = (const float *)vsites_v + ((long long)u * NS + w) * 4096;
#pragma unroll
        for (int r = 0; r < 64; ++r) T[r] = src[r * 64 + lane];
    }
    __syncthreads();
    // results go to LDS as they are made (fewer live registers; with one wave per
    // workgroup the compiler allocates best when they stay in registers to the end); lane l
    // holds entry vx_split_entry(l) of each result
    float *mypart = part + w * 4 * 64 + vx_split_entry(lane);
    float out[4];
    auto put = [&](int j, float v) {
        if constexpr (NS == 1) out[j] = v; else mypart[j * 64] = v;
    };


    // ---- phase A: E01 (X) -> Y through the buffer; then A into the buffer for the bra.
    // The buffer belongs to this wave: wave-level LDS ordering suffices between its uses.
    vx_mode<true, BF>(T, E, msm + 0 * 64);
    vx_mode<false, BF>(E, G, msm + 1 * 64);
#pragma unroll
    for (int r = 0; r < 64; ++r) buf[r * VX_PITCH + lane] = G[r];
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 64; ++r) E[r] = buf[lane * VX_PITCH + r];  // E01 in Y
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 64; ++r) buf[r * VX_PITCH + lane] = T[r];  // A, readable as X or Y
    wave_lds_sync();
    vx_mode<false, BF>(E, G, msm + 3 * 64);
    if constexpr (BF) vx_round_bf16(G);  // the partial product dies here
    put(2, vx_final_half<true, false, VX_PITCH>(G, buf));
    vx_mode<true, BF>(E, G, msm + 2 * 64);
    if constexpr (BF) vx_round_bf16(G);  // the partial product dies here
    put(3, vx_final_half<false, false, VX_PITCH>(G, buf));

    // ---- phase B: E23 (Y) -> X through the buffer, then A written back for the bra (X)
#pragma unroll
    for (int r = 0; r < 64; ++r) T[r] = buf[lane * VX_PITCH + r];  // A in Y
    vx_mode<true, BF>(T, E, msm + 2 * 64);
    vx_mode<false, BF>(E, G, msm + 3 * 64);
    wave_lds_sync();  // every lane has finished reading A from the buffer
#pragma unroll
    for (int r = 0; r < 64; ++r) buf[lane * VX_PITCH + r] = G[r];
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 64; ++r) E[r] = buf[r * VX_PITCH + lane];  // E23 in X
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 64; ++r) buf[lane * VX_PITCH + r] = T[r];  // A again
    wave_lds_sync();
    vx_mode<false, BF>(E, G, msm + 1 * 64);
    if constexpr (BF) vx_round_bf16(G);  // the partial product dies here
    put(0, vx_final_half<true, true, VX_PITCH>(G, buf));
    vx_mode<true, BF>(E, G, msm + 0 * 64);
    if constexpr (BF) vx_round_bf16(G);  // the partial product dies here
    put(1, vx_final_half<false, true, VX_PITCH>(G, buf));

    // ---- sum over s, normalise, residual, write (as k_vertex_d8); the slot data are read
    // again here rather than kept live through the contractions
    double fm = 0.0;
    const bool fcrit = fold.acc && fold.rp->crit != 0;
    if constexpr (NS == 1) {
#pragma unroll
        for (int j = 0; j < 4; ++j) mypart[j * 64] = out[j];
    }
    __syncthreads();
    const int4 vo2 = vout[u], vd2 = vdim[u];
    const int outs2[4] = {vo2.x, vo2.y, vo2.z, vo2.w};
    const int dims2[4] = {vd2.x, vd2.y, vd2.z, vd2.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int q = outs2[j];
        if (j % NS != w || q < 0) continue;
        float x = 0.f;
#pragma unroll
        for (int s = 0; s < NS; ++s) x += part[(s * 4 + j) * 64 + lane];
        const double S = wave_sum_v((double)x);
        const double xn = S != 0.0 ? (double)x / S : (double)x;
        const float xf = (float)xn;
        const int D = dims2[j];
        const bool real = la < D && lb < D;
        const long long o = msg_off[q] + la * D + lb;
        const double xo = real ? (double)mold[o] : 0.0;
        if (real) mnew[o] = xf;
        const double dn = wave_sum_v((double)xf * xo),
                     nn = wave_sum_v((double)xf * (double)xf), oo = wave_sum_v(xo * xo);
        const double mx = wave_max_v(fabs((double)xf - xo));
        if (lane == 0) {
            const double c = dn / sqrt(nn * oo);
            res_fid[q] = 1.0 - c * c;
            res_abs[q] = mx;
            fm = fold_max(fm, fcrit ? mx : 1.0 - c * c);
        }
    }
    if (fold.acc) {
        if (lane == 0) fold_m[w] = fm;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = 0.0;
            for (int s = 0; s < NS; ++s) m = fold_max(m, fold_m[s]);
            atomicMax(fold.acc + (long long)net * FOLD_SLOTS + (u % FOLD_SLOTS),
                      (unsigned long long)__double_as_longlong(m));
        }
    }
}

// ---------------------------------------------------------------- lane-port MFMA opens (FP32)
//
// Variant 4 of the D <= 8 FP32 sweep (k_vertex_d8_hy): the low-LDS kernel with two of the four
// open contractions per slice on the MFMA pipe.  An open on a port that sits in the low three
// lane bits needs no staging: v_mfma_f32_4x4x1_16b_f32 pairs lane 4b + i (first operand, row
// i) with lane 4b + j (second operand, column j), so with the ket in registers (K = register
// index and lane bits 3..5) and the bra read from LDS, block b accumulates
// out[4c + i][4c' + j] over the registers, c = lane bit 2 and c' = c (the bra at the lane
// itself) or 1 - c (the bra at lane ^ 4).  The other partner message is absorbed into the
// bra, not the ket:
//   out3 = open_p3(E01, A x2 m2^T)   (E01 in layout X: port 3 = lane bits 0..2)
//   out1 = open_p1(E23, A x0 m0^T)   (E23 in layout Y: port 1 = lane bits 0..2)
// which keeps eight mode products per slice.  ALL = false: out2 and out0 stay on the VALU as
// in the low-LDS kernel (variant 4).  ALL = true (variant 5): they run the same way from the
// lane-swapped layouts X' (registers p0 p1, lanes p3 (bits 3..5) p2 (bits 0..2)) and Y'
// (registers p2 p3, lanes p1 p0), made through the buffer like the X <-> Y transposes:
//   out2 = open_p2(E01 in X', A x3 m3^T in X')   out0 = open_p0(E23 in Y', A x1 m1^T in Y')

// Mode product with the transposed message: O[b] = sum_a T[a] m[b][a] on a register port, in
// the precision of the kernel (BF16: Mb[g * 2 + h] packs m[4h + lane % 4][4g + t], t = 0..3).
template <bool HI, bool BF = false>
__device__ __forceinline__ void vx_mode_t(const float (&T)[64], float (&O)[64], const float *msg) {
    const int c = threadIdx.x & 3;
    if constexpr (BF) {
        v4s Mb[4];
#pragma unroll
        for (int g = 0; g < 2; ++g)
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const float *p = msg + (4 * h + c) * 8 + 4 * g;
                Mb[g * 2 + h] = bf16x4(p[0], p[1], p[2], p[3]);
            }
        vx_mp_bf16<HI>(T, O, Mb);
    } else {
        float M[16];
#pragma unroll
        for (int a = 0; a < 8; ++a)
#pragma unroll
            for (int h = 0; h < 2; ++h) M[a * 2 + h] = msg[(4 * h + c) * 8 + a];
        vx_mp<HI>(T, O, M);
    }
}

// Entry (a * 8 + a') of the open contraction that lane l holds after vx_open_lane:
// a = 4 b2 + 2 b4 + b3, a' = 4 (b2 ^ b5) + l % 4 (b_k = bit k of l).
__device__ __forceinline__ int vx_lane_entry(int lane) {
    const int b2 = (lane >> 2) & 1, b3 = (lane >> 3) & 1, b4 = (lane >> 4) & 1,
              b5 = (lane >> 5) & 1;
    return (4 * b2 + 2 * b4 + b3) * 8 + 4 * (b2 ^ b5) + (lane & 3);
}

// The bras of these opens are stored lane-major, bra element (r, lane) at
// bra[lane * VH_PITCH + r]: each lane reads its own row (and that of lane ^ 4) as 16-byte
// words, and a pitch of 68 floats puts the 16 lanes of one read cycle on distinct banks.  The
// buffer is VH_BUF floats; the X <-> Y transposes inside it keep the pitch VX_PITCH.
constexpr int VH_PITCH = 68;
constexpr int VH_BUF = 64 * VH_PITCH;

// Accumulates, per MFMA block, the products of vx_open_lane for one ket and bra: acc[0], acc[1]
// the bra at the lane itself, acc[2], acc[3] the bra at lane ^ 4.
__device__ __forceinline__ void vx_open_acc(const float (&K)[64], const float *bra, v4f (&acc)[4]) {
    const int lane = threadIdx.x & 63;
    const float4 *bd = reinterpret_cast<const float4 *>(bra + lane * VH_PITCH);
    const float4 *bo = reinterpret_cast<const float4 *>(bra + (lane ^ 4) * VH_PITCH);
#pragma unroll
    for (int q = 0; q < 16; ++q) {
        const float4 x = bd[q], y = bo[q];
        acc[0] = __builtin_amdgcn_mfma_f32_4x4x1f32(K[4 * q], x.x, acc[0], 0, 0, 0);
        acc[2] = __builtin_amdgcn_mfma_f32_4x4x1f32(K[4 * q], y.x, acc[2], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_4x4x1f32(K[4 * q + 1], x.y, acc[1], 0, 0, 0);
        acc[3] = __builtin_amdgcn_mfma_f32_4x4x1f32(K[4 * q + 1], y.y, acc[3], 0, 0, 0);
        acc[0] = __builtin_amdgcn_mfma_f32_4x4x1f32(K[4 * q + 2], x.z, acc[0], 0, 0, 0);
        acc[2] = __builtin_amdgcn_mfma_f32_4x4x1f32(K[4 * q + 2], y.z, acc[2], 0, 0, 0);
        acc[1] = __builtin_amdgcn_mfma_f32_4x4x1f32(K[4 * q + 3], x.w, acc[1], 0, 0, 0);
        acc[3] = __builtin_amdgcn_mfma_f32_4x4x1f32(K[4 * q + 3], y.w, acc[3], 0, 0, 0);
    }
}

// The sum over the MFMA blocks of the same open port (lane bits 5, 4, 3): lanes with bit 5
// set keep the other-half tile, bit 4 and bit 3 select the row i.
__device__ __forceinline__ float vx_open_fin(const v4f (&acc)[4]) {
    const int lane = threadIdx.x & 63;
    float v[8];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        v[i] = acc[0][i] + acc[1][i];
        v[4 + i] = acc[2][i] + acc[3][i];
    }
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v[k]),
                                                        __float_as_uint(v[k + 4]), false, false);
        v[k] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v[k]),
                                                        __float_as_uint(v[k + 2]), false, false);
        v[k] = __uint_as_float(r[0]) + __uint_as_float(r[1]);
    }
    const bool up = (lane & 8) != 0;
    const float send = up ? v[0] : v[1], keep = up ? v[1] : v[0];
    return keep + vmo_ror<8>(send);  // row_ror:8 pairs lane l with l ^ 8 exactly
}

// out[a][a'] = sum over the registers r and lane bits 3..5 of K[r](a) bra[r](a'), the open
// port in lane bits 0..2.
__device__ __forceinline__ float vx_open_lane(const float (&K)[64], const float *bra) {
    v4f acc[4] = {{0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f}, {0.f, 0.f, 0.f, 0.f},
                  {0.f, 0.f, 0.f, 0.f}};
    vx_open_acc(K, bra, acc);
    return vx_open_fin(acc);
}

// Lane index with its two 3-bit halves swapped ((p, q) -> (q, p)).
__device__ __forceinline__ int vx_lane_swap(int l) { return ((l & 7) << 3) | (l >> 3); }

template <int NS, bool ALL, bool BF = false>
__global__ __launch_bounds__(64 * NS) __attribute__((amdgpu_waves_per_eu(2, 2)))
void k_vertex_d8_hy(const void *__restrict__ vsites_v, const float *__restrict__ mold,
                    float *__restrict__ mnew, const int4 *__restrict__ vin,
                    const int4 *__restrict__ vout, const int4 *__restrict__ vdim,
                    const int *__restrict__ vnet, const long long *__restrict__ msg_off,
                    NetCtl ctl, double *__restrict__ res_fid, double *__restrict__ res_abs,
                    SweepFold fold) {
    extern __shared__ __align__(16) float vsm[];
    __shared__ double fold_m[4];
    const int u = blockIdx.x;
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    float *buf = vsm + w * VH_BUF;
    float *msm = vsm + NS * VH_BUF;      // 4 x 64 incoming messages
    float *part = msm + 4 * 64;          // [s][port][64] per-slice results
    const int4 vi = vin[u], vo = vout[u], vdm = vdim[u];
    const int ins[4] = {vi.x, vi.y, vi.z, vi.w};
    const int outs[4] = {vo.x, vo.y, vo.z, vo.w};
    const int dims[4] = {vdm.x, vdm.y, vdm.z, vdm.w};
    long long ino[4], outo[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        ino[j] = ins[j] >= 0 ? msg_off[ins[j]] : 0;
        outo[j] = outs[j] >= 0 ? msg_off[outs[j]] : 0;
    }
    const int la = lane >> 3, lb = lane & 7;
    const int net = vnet[u];
    if (vx_fold_begin(ctl, fold, net, u)) {  // uniform over the workgroup
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j % NS != w || outs[j] < 0) continue;
            if (lane < dims[j] * dims[j]) mnew[outo[j] + lane] = mold[outo[j] + lane];
            if (lane == 0) { res_fid[outs[j]] = 0.0; res_abs[outs[j]] = 0.0; }
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j % NS != w) continue;
        const int D = dims[j];
        float x;
        if (ins[j] < 0) x = la == lb ? 1.f : 0.f;
        else x = (la < D && lb < D) ? mold[ino[j] + la * D + lb] : 0.f;
        if constexpr (BF) {
            const float mx = wave_max(fabsf(x));
            x = mx > 0.f ? x / mx : x;
        }
        msm[j * 64 + lane] = x;
    }
    float T[64], E[64], G[64];
    if constexpr (BF) {
        const unsigned *src = (const unsigned *)vsites_v + ((long long)u * NS + w) * 2048;
#pragma unroll
        for (int r2 = 0; r2 < 32; ++r2) {
            const unsigned pr = src[r2 * 64 + lane];
            T[2 * r2] = bf16_to_f((unsigned short)(pr & 0xffffu));
            T[2 * r2 + 1] = bf16_to_f((unsigned short)(pr >> 16));
        }
    } else {
        const float *src = (const float *)vsites_v + ((long long)u * NS + w) * 4096;
#pragma unroll
        for (int r = 0; r < 64; ++r) T[r] = src[r * 64 + lane];
    }
    __syncthreads();
    float *mypart = part + w * 4 * 64;
    const int e_split = vx_split_entry(lane), e_lane = vx_lane_entry(lane);

    // ---- phase A (layout X: registers p0 p1, lanes p2 (bits 3..5) p3 (bits 0..2))
    vx_mode<true, BF>(T, E, msm + 0 * 64);
    vx_mode<false, BF>(E, G, msm + 1 * 64);  // G = E01 (X), kept for out3
    if constexpr (ALL) {
#pragma unroll
        for (int r = 0; r < 64; ++r) buf[r * VX_PITCH + lane] = T[r];  // A (X)
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < 64; ++r) E[r] = buf[lane * VX_PITCH + r];  // A in Y (kept to phase B)
    } else {
#pragma unroll
        for (int r = 0; r < 64; ++r) buf[r * VX_PITCH + lane] = G[r];
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < 64; ++r) E[r] = buf[lane * VX_PITCH + r];  // E01 in Y
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < 64; ++r) buf[r * VX_PITCH + lane] = T[r];  // A, the bra of out2
        wave_lds_sync();
        vx_mode<false, BF>(E, T, msm + 3 * 64);  // E01 x3 m3 (Y)
        mypart[2 * 64 + e_split] = vx_final_half<true, false, VX_PITCH>(T, buf);
#pragma unroll
        for (int r = 0; r < 64; ++r) E[r] = buf[lane * VX_PITCH + r];  // A in Y (kept to phase B)
    }
    vx_mode_t<true, BF>(E, T, msm + 2 * 64);        // B2 = A x2 m2^T (Y)
    wave_lds_sync();  // every lane has finished reading A from the buffer
#pragma unroll
    for (int r = 0; r < 64; ++r) buf[r * VH_PITCH + lane] = T[r];  // B2, lane-major in X
    wave_lds_sync();
    mypart[3 * 64 + e_lane] = vx_open_lane(G, buf);
    if constexpr (ALL) {
        vx_mode_t<false, BF>(E, T, msm + 3 * 64);   // B3 = A x3 m3^T (Y)
        wave_lds_sync();  // every lane has finished reading B2
#pragma unroll
        for (int r = 0; r < 64; ++r) buf[r * VX_PITCH + lane] = G[r];
        wave_lds_sync();
        const int ls = vx_lane_swap(lane);
#pragma unroll
        for (int r = 0; r < 64; ++r) G[r] = buf[r * VX_PITCH + ls];  // E01 in X'
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < 64; ++r) buf[vx_lane_swap(r) * VH_PITCH + lane] = T[r];  // B3 in X'
        wave_lds_sync();
        mypart[2 * 64 + e_lane] = vx_open_lane(G, buf);
    }

    // ---- phase B (layout Y: registers p2 p3, lanes p0 (bits 3..5) p1 (bits 0..2))
    vx_mode<true, BF>(E, T, msm + 2 * 64);
    vx_mode<false, BF>(T, G, msm + 3 * 64);  // G = E23 (Y), kept for out1
    wave_lds_sync();  // every lane has finished reading the last bra
#pragma unroll
    for (int r = 0; r < 64; ++r) buf[lane * VX_PITCH + r] = E[r];  // A, readable as X
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 64; ++r) T[r] = buf[r * VX_PITCH + lane];  // A in X
    vx_mode_t<true, BF>(T, E, msm + 0 * 64);        // B0 = A x0 m0^T (X)
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 64; ++r) buf[r * VH_PITCH + lane] = E[r];  // B0, lane-major in Y
    wave_lds_sync();
    mypart[1 * 64 + e_lane] = vx_open_lane(G, buf);
    if constexpr (ALL) {
        vx_mode_t<false, BF>(T, E, msm + 1 * 64);   // B1 = A x1 m1^T (X)
        wave_lds_sync();  // every lane has finished reading B0
#pragma unroll
        for (int r = 0; r < 64; ++r) buf[r * VX_PITCH + lane] = G[r];
        wave_lds_sync();
        const int ls = vx_lane_swap(lane);
#pragma unroll
        for (int r = 0; r < 64; ++r) G[r] = buf[r * VX_PITCH + ls];  // E23 in Y'
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < 64; ++r) buf[vx_lane_swap(r) * VH_PITCH + lane] = E[r];  // B1 in Y'
        wave_lds_sync();
        mypart[0 * 64 + e_lane] = vx_open_lane(G, buf);
    } else {
        wave_lds_sync();  // every lane has finished reading B0
#pragma unroll
        for (int r = 0; r < 64; ++r) buf[r * VX_PITCH + lane] = G[r];
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < 64; ++r) E[r] = buf[lane * VX_PITCH + r];  // E23 in X
        wave_lds_sync();
#pragma unroll
        for (int r = 0; r < 64; ++r) buf[r * VX_PITCH + lane] = T[r];  // A, the bra of out0
        wave_lds_sync();
        vx_mode<false, BF>(E, G, msm + 1 * 64);  // E23 x1 m1 (X)
        mypart[0 * 64 + e_split] = vx_final_half<true, true, VX_PITCH>(G, buf);
    }

    // ---- sum over s, normalise, residual, write (as k_vertex_d8_lo)
    double fm = 0.0;
    const bool fcrit = fold.acc && fold.rp->crit != 0;
    __syncthreads();
    const int4 vo2 = vout[u], vd2 = vdim[u];
    const int outs2[4] = {vo2.x, vo2.y, vo2.z, vo2.w};
    const int dims2[4] = {vd2.x, vd2.y, vd2.z, vd2.w};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int q = outs2[j];
        if (j % NS != w || q < 0) continue;
        float x = 0.f;
#pragma unroll
        for (int s = 0; s < NS; ++s) x += part[(s * 4 + j) * 64 + lane];
        const double S = wave_sum_v((double)x);
        const double xn = S != 0.0 ? (double)x / S : (double)x;
        const float xf = (float)xn;
        const int D = dims2[j];
        const bool real = la < D && lb < D;
        const long long o = msg_off[q] + la * D + lb;
        const double xo = real ? (double)mold[o] : 0.0;
        if (real) mnew[o] = xf;
        const double dn = wave_sum_v((double)xf * xo),
                     nn = wave_sum_v((double)xf * (double)xf), oo = wave_sum_v(xo * xo);
        const double mx = wave_max_v(fabs((double)xf - xo));
        if (lane == 0) {
            const double c = dn / sqrt(nn * oo);
            res_fid[q] = 1.0 - c * c;
            res_abs[q] = mx;
            fm = fold_max(fm, fcrit ? mx : 1.0 - c * c);
        }
    }
    if (fold.acc) {
        if (lane == 0) fold_m[w] = fm;
        __syncthreads();
        if (threadIdx.x == 0) {
            double m = 0.0;
            for (int s = 0; s < NS; ++s) m = fold_max(m, fold_m[s]);
            atomicMax(fold.acc + (long long)net * FOLD_SLOTS + (u % FOLD_SLOTS),
                      (unsigned long long)__double_as_longlong(m));
        }
    }
}

static size_t vertex_lo_lds(int ns) {
    return sizeof(float) * ((size_t)ns * VX_BUF + 4 * 64 + (size_t)ns * 4 * 64);
}

static size_t vertex_hy_lds(int ns) {
    return sizeof(float) * ((size_t)ns * VH_BUF + 4 * 64 + (size_t)ns * 4 * 64);
}

static int vertex_sweep(VertexPlan *p, hipStream_t st, DevTopo tp, const void *vsites,
                        const void *src, void *dst, const NetCtl &ctl, double *res_fid,
                        double *res_abs, const SweepFold &fold) {
    (void)tp;
    const int n4 = p->nv - p->nv_k3;  // items on the four-port kernels
    if (n4 == 0) return 0;
    if ((p->variant == 2 || p->variant == 3) && !p->bf16) {  // MFMA open contractions (FP32)
        const bool gb = p->variant == 3;
        const size_t lds = gb ? vertex_lo_lds(p->d) : p->lds_bytes;
        switch (p->d * 2 + (gb ? 1 : 0)) {
#define VX_MO(NS, GB)                                                                          \
    case NS * 2 + GB:                                                                          \
        hipLaunchKernelGGL((k_vertex_d8_mo<NS, GB>), dim3(n4), dim3(64 * NS), lds, st, vsites, \
                           (const float *)p->d_vsites_y, (const float *)src, (float *)dst,     \
                           p->d_vin, p->d_vout, p->d_vdim, p->d_vnet, tp.msg_off, ctl,         \
                           res_fid, res_abs, fold);                                            \
        break;
            VX_MO(1, 0)
            VX_MO(2, 0)
            VX_MO(3, 0)
            VX_MO(4, 0)
            VX_MO(1, 1)
            VX_MO(2, 1)
            VX_MO(3, 1)
            VX_MO(4, 1)
#undef VX_MO
            default:
                return 4;
        }
        return hipGetLastError() == hipSuccess ? 0 : 2;
    }
    // lane-port MFMA opens: variant 4 (FP32), variant 5 (FP32 or BF16)
    if ((p->variant == 4 && !p->bf16) || p->variant == 5) {
        const size_t lds = vertex_hy_lds(p->d);
        switch (p->d * 4 + (p->variant == 5 ? 2 : 0) + (p->bf16 ? 1 : 0)) {
#define VX_HY(NS, ALL, BF)                                                                     \
    case NS * 4 + ALL * 2 + BF:                                                                \
        hipLaunchKernelGGL((k_vertex_d8_hy<NS, ALL, BF>), dim3(n4), dim3(64 * NS), lds, st,    \
                           vsites, (const float *)src, (float *)dst, p->d_vin, p->d_vout,      \
                           p->d_vdim, p->d_vnet, tp.msg_off, ctl, res_fid, res_abs, fold);     \
        break;
            VX_HY(1, 0, 0)
            VX_HY(2, 0, 0)
            VX_HY(3, 0, 0)
            VX_HY(4, 0, 0)
            VX_HY(1, 1, 0)
            VX_HY(2, 1, 0)
            VX_HY(3, 1, 0)
            VX_HY(4, 1, 0)
            VX_HY(1, 1, 1)
            VX_HY(2, 1, 1)
            VX_HY(3, 1, 1)
            VX_HY(4, 1, 1)
#undef VX_HY
            default:
                return 4;
        }
        return hipGetLastError() == hipSuccess ? 0 : 2;
    }
    if (p->variant == 1) {
        const size_t lds = vertex_lo_lds(p->d);
        switch (p->d * 2 + (p->bf16 ? 1 : 0)) {
#define VX_LO(NS, BF)                                                                          \
    case NS * 2 + BF:                                                                          \
        hipLaunchKernelGGL((k_vertex_d8_lo<NS, BF>), dim3(n4), dim3(64 * NS), lds, st, vsites, \
                           (const float *)src, (float *)dst, p->d_vin, p->d_vout, p->d_vdim,   \
                           p->d_vnet, tp.msg_off, ctl, res_fid, res_abs, fold);                \
        break;
            VX_LO(1, 0)
            VX_LO(2, 0)
            VX_LO(3, 0)
            VX_LO(4, 0)
            VX_LO(1, 1)
            VX_LO(2, 1)
            VX_LO(3, 1)
            VX_LO(4, 1)
#undef VX_LO
            default:
                return 4;
        }
        return hipGetLastError() == hipSuccess ? 0 : 2;
    }
    const int key = p->d * 2 + (p->bf16 ? 1 : 0);
    switch (key) {
#define VX_CASE(NS, BF)                                                                      \
    case NS * 2 + BF:                                                                        \
        hipLaunchKernelGGL((k_vertex_d8<NS, BF>), dim3(n4), dim3(64 * NS), p->lds_bytes,     \
                           st, vsites, (const float *)src, (float *)dst, p->d_vin, p->d_vout, \
                           p->d_vdim, p->d_vnet, tp.msg_off, ctl, res_fid, res_abs, fold);   \
        break;
        VX_CASE(1, 0)
        VX_CASE(2, 0)
        VX_CASE(3, 0)
        VX_CASE(4, 0)
        VX_CASE(1, 1)
        VX_CASE(2, 1)
        VX_CASE(3, 1)
        VX_CASE(4, 1)
#undef VX_CASE
        default:
            return 4;
    }
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

static int vertex_set_lds_limits(const VertexPlan *p) {
    const void *fns[8] = {(const void *)k_vertex_d8<1, false>, (const void *)k_vertex_d8<2, false>,
                          (const void *)k_vertex_d8<3, false>, (const void *)k_vertex_d8<4, false>,
                          (const void *)k_vertex_d8<1, true>, (const void *)k_vertex_d8<2, true>,
                          (const void *)k_vertex_d8<3, true>, (const void *)k_vertex_d8<4, true>};
    const void *mo[4] = {(const void *)k_vertex_d8_mo<1, false>, (const void *)k_vertex_d8_mo<2, false>,
                         (const void *)k_vertex_d8_mo<3, false>, (const void *)k_vertex_d8_mo<4, false>};
    const void *mg[4] = {(const void *)k_vertex_d8_mo<1, true>, (const void *)k_vertex_d8_mo<2, true>,
                         (const void *)k_vertex_d8_mo<3, true>, (const void *)k_vertex_d8_mo<4, true>};
    const void *hy[12] = {(const void *)k_vertex_d8_hy<1, false>, (const void *)k_vertex_d8_hy<2, false>,
                          (const void *)k_vertex_d8_hy<3, false>, (const void *)k_vertex_d8_hy<4, false>,
                          (const void *)k_vertex_d8_hy<1, true>, (const void *)k_vertex_d8_hy<2, true>,
                          (const void *)k_vertex_d8_hy<3, true>, (const void *)k_vertex_d8_hy<4, true>,
                          (const void *)k_vertex_d8_hy<1, true, true>, (const void *)k_vertex_d8_hy<2, true, true>,
                          (const void *)k_vertex_d8_hy<3, true, true>, (const void *)k_vertex_d8_hy<4, true, true>};
    if (p->d < 1 || p->d > 4) return 4;
    if (!p->bf16 && (hipFuncSetAttribute(mo[p->d - 1], hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)p->lds_bytes) != hipSuccess ||
                     hipFuncSetAttribute(mg[p->d - 1], hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)vertex_lo_lds(p->d)) != hipSuccess ||
                     hipFuncSetAttribute(hy[p->d - 1], hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)vertex_hy_lds(p->d)) != hipSuccess ||
                     hipFuncSetAttribute(hy[4 + p->d - 1], hipFuncAttributeMaxDynamicSharedMemorySize,
                                         (int)vertex_hy_lds(p->d)) != hipSuccess))
        return 2;
    if (p->bf16 && hipFuncSetAttribute(hy[8 + p->d - 1], hipFuncAttributeMaxDynamicSharedMemorySize,
                                       (int)vertex_hy_lds(p->d)) != hipSuccess)
        return 2;
    return hipFuncSetAttribute(fns[(p->bf16 ? 4 : 0) + p->d - 1],
                               hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)p->lds_bytes) == hipSuccess ? 0 : 2;
}
