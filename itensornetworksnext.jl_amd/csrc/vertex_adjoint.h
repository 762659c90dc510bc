// vertex_adjoint.h -- fixed-point adjoint and BP energy on the bond-dimension-8 vertex path
// (FP32 MFMA), for buckets that hold every network.
//
// Every quantity the adjoint needs is one of two per-vertex contractions of the slice
// layout of vertex.h, with the ket tensor op.A (op = I, X or Z on the physical index: a
// signed permutation of the slices, so wave s reads ket slice pi(s) with sign sigma(s)):
//
//   message form, for every port j of u:
//     R_j = cv * open(op.A, A, m, j) + sum_{t != j} open(op.A, A, {t <- W_t}, j)
//   site form, for every slice s of u:
//     R[s] = sigma(s) * ( cv * (A[pi(s)] x_all M) + sum_t A[pi(s)] x_all {M, t <- W_t} )
//     with M the incoming messages (or their transposes, for the ket-side derivative).
//
// The sums over t are directional derivatives of multilinear maps and run as dual-number
// absorptions: value V_{i+1} = V_i x_i m_i and tangent T_{i+1} = T_i x_i m_i + V_i x_i W_i,
// sharing the partial products E01 / E23 exactly as the forward sweep kernel does.  The
// value term folds into the tangent before the open contraction (open is linear), so a
// message-form launch costs 18 mode products per slice against the forward's 8.
//
// Message form with cv = 1 and no W gives the unnormalised forward messages G (and with
// op = X or Z the operator-inserted ones), their sums S and the vertex scalars
// Z_u = <G_{u, port 0}, m_{in, port 0}>.  With W = the normalisation cotangent of lambda it
// is one adjoint sweep lambda' = dE/dm + J_m^T lambda (every in-slot has exactly one
// consumer vertex, so results are written, not accumulated).  The site form gives the
// bra-side (M) and ket-side (M^T) derivatives of <op.A, A>_M.
#pragma once

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "adjoint.h"
#include "vertex.h"

// Mode products that finish both output halves of a row/column before writing, so the
// output may be the same array as the first input.
template <bool HI>
__device__ __forceinline__ void vd_mp1(const float (&T)[64], const float (&M)[16], float (&O)[64]) {
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        v4f acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const float b = HI ? T[a * 8 + x] : T[x * 8 + a];
            acc0 = __builtin_amdgcn_mfma_f32_4x4x1f32(M[a * 2 + 0], b, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_4x4x1f32(M[a * 2 + 1], b, acc1, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (HI) { O[i * 8 + x] = acc0[i]; O[(4 + i) * 8 + x] = acc1[i]; }
            else { O[x * 8 + i] = acc0[i]; O[x * 8 + 4 + i] = acc1[i]; }
        }
    }
}

// O = T1 x M1 + T2 x M2 on the same register port (O may be T1).
template <bool HI>
__device__ __forceinline__ void vd_mp2(const float (&T1)[64], const float (&M1)[16],
                                       const float (&T2)[64], const float (&M2)[16],
                                       float (&O)[64]) {
#pragma unroll
    for (int x = 0; x < 8; ++x) {
        v4f acc0 = {0.f, 0.f, 0.f, 0.f}, acc1 = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const float b = HI ? T1[a * 8 + x] : T1[x * 8 + a];
            acc0 = __builtin_amdgcn_mfma_f32_4x4x1f32(M1[a * 2 + 0], b, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_4x4x1f32(M1[a * 2 + 1], b, acc1, 0, 0, 0);
        }
#pragma unroll
        for (int a = 0; a < 8; ++a) {
            const float b = HI ? T2[a * 8 + x] : T2[x * 8 + a];
            acc0 = __builtin_amdgcn_mfma_f32_4x4x1f32(M2[a * 2 + 0], b, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_4x4x1f32(M2[a * 2 + 1], b, acc1, 0, 0, 0);
        }
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            if (HI) { O[i * 8 + x] = acc0[i]; O[(4 + i) * 8 + x] = acc1[i]; }
            else { O[x * 8 + i] = acc0[i]; O[x * 8 + 4 + i] = acc1[i]; }
        }
    }
}

// Register transposes through the wave's own LDS buffer (row pitch 65: conflict-free);
// only the wave's lanes touch it, so wave-level ordering suffices.
__device__ __forceinline__ void vd_x_to_y(float (&R)[64], float *buf, int lane) {
#pragma unroll
    for (int r = 0; r < 64; ++r) buf[r * VX_PITCH + lane] = R[r];
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 64; ++r) R[r] = buf[lane * VX_PITCH + r];
    wave_lds_sync();
}

__device__ __forceinline__ void vd_y_to_x(float (&R)[64], float *buf, int lane) {
#pragma unroll
    for (int r = 0; r < 64; ++r) buf[lane * VX_PITCH + r] = R[r];
    wave_lds_sync();
#pragma unroll
    for (int r = 0; r < 64; ++r) R[r] = buf[r * VX_PITCH + lane];
    wave_lds_sync();
}

// Incoming messages (identity on padded ports) and overrides W at the out-slots (zero on
// padded ports), zero-padded to 8 x 8, into LDS, one port per wave in turn; `trans` stores
// every matrix transposed.  Port indices are compile-time so the slot arrays stay in
// registers.
struct VdSlots {
    int ins[4], outs[4], dims[4];
    long long ino[4], outo[4];
};

__device__ __forceinline__ VdSlots vd_slots(const int4 *vin, const int4 *vout, const int4 *vdim,
                                            const long long *msg_off, int u) {
    VdSlots s;
    const int4 vi = vin[u], vo = vout[u], vd = vdim[u];
    s.ins[0] = vi.x; s.ins[1] = vi.y; s.ins[2] = vi.z; s.ins[3] = vi.w;
    s.outs[0] = vo.x; s.outs[1] = vo.y; s.outs[2] = vo.z; s.outs[3] = vo.w;
    s.dims[0] = vd.x; s.dims[1] = vd.y; s.dims[2] = vd.z; s.dims[3] = vd.w;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        s.ino[j] = s.ins[j] >= 0 ? msg_off[s.ins[j]] : 0;
        s.outo[j] = s.outs[j] >= 0 ? msg_off[s.outs[j]] : 0;
    }
    return s;
}

// With Gn != nullptr, W holds lambda and the override is its cotangent through the sum
// normalisation, (lambda - <lambda, G>/S) / S (G the unnormalised forward message, S its
// sum; lambda unchanged where S = 0).
// With W2 != nullptr (two slices, d = 2), wave w gets its own overrides W + sigma_w W2 at
// wsm + 64 * 4 * w, sigma = (+1, -1): the Z-inserted launch with overrides W2 folded into
// the plain launch with overrides W (both results are linear in the overrides, and Z only
// flips the sign of slice 1).  Otherwise the overrides sit at wsm, shared by the waves.
template <int NS>
__device__ __forceinline__ void vd_load_mats(const float *msgs, const float *W, const VdSlots &sl,
                                             int trans, float *msm, float *wsm,
                                             const float *Gn = nullptr,
                                             const double *Sn = nullptr,
                                             const float *W2 = nullptr) {
    const int w = threadIdx.x >> 6, e = threadIdx.x & 63;
    const int a = trans ? (e & 7) : (e >> 3), b = trans ? (e >> 3) : (e & 7);
    if (W && W2) {
        const float sgn = w == 1 ? -1.f : 1.f;
        float *ws = wsm + w * 256;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const int D = sl.dims[j];
            const bool have = sl.outs[j] >= 0 && a < D && b < D;
            const long long off = sl.outo[j] + a * D + b;
            ws[j * 64 + e] = have ? W[off] + sgn * W2[off] : 0.f;
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j % NS != w) continue;
        const int D = sl.dims[j];
        const bool real = a < D && b < D;
        msm[j * 64 + e] = sl.ins[j] >= 0 ? (real ? msgs[sl.ino[j] + a * D + b] : 0.f)
                                         : (a == b ? 1.f : 0.f);
        if (W && !W2) {
            const bool have = sl.outs[j] >= 0 && real;
            float x = have ? W[sl.outo[j] + a * D + b] : 0.f;
            if (Gn) {
                const float g = have ? Gn[sl.outo[j] + a * D + b] : 0.f;
                const double dot = wave_sum_v((double)x * (double)g);
                const double S = sl.outs[j] >= 0 ? Sn[sl.outs[j]] : 0.0;
                if (S != 0.0) x = (float)(((double)x - dot / S) / S);
            }
            wsm[j * 64 + e] = x;
        }
    }
}

// Convergence bookkeeping of the adjoint sweeps, kept on the device so that sweeps are
// queued in chunks without a host round trip per sweep.  Sweep j accumulates its maxima in
// slot j % 3 of the accumulator (3 x 128 words); the check of sweep j runs at the start of
// sweep j + 1 (every workgroup forms the same r from the same words) or in k_vd_adj_check.
struct VdLoopState {
    int stop;  // set once the relative change fell below tol
    int it;    // sweeps checked (= sweeps that ran, once the loop has ended)
    double r;  // relative change of the last checked sweep
};

// r = max|lambda' - lambda| / max|lambda'| of sweep j from the workgroup maxima in slot j % 3
// (NaN in the change propagates; NaN in the magnitude is ignored).  One wave, uniform result.
__device__ __forceinline__ double vd_adj_r(const unsigned long long *acc, int j, int lane) {
    const unsigned long long *a = acc + (j % 3) * 128;
    const double x = __longlong_as_double((long long)a[lane]);
    double y = __longlong_as_double((long long)a[64 + lane]);
    if (y != y) y = 0.0;
    const double ra = wave_max_v(x), rr = wave_max_v(y);
    return rr > 0.0 ? ra / rr : ra;
}

// Check of sweep j (thread 0 of the grid records it); returns whether sweeps go on.
__device__ __forceinline__ bool vd_adj_check(const unsigned long long *acc, int j, double tol,
                                             VdLoopState *st, int lane, bool record) {
    const double r = vd_adj_r(acc, j, lane);
    const bool go = !(r < tol);
    if (record && lane == 0) {
        st->it = j + 1;
        st->r = r;
        if (!go) st->stop = 1;
    }
    return go;
}

__global__ __launch_bounds__(64) void k_vd_adj_check(const unsigned long long *__restrict__ acc,
                                                     int j, double tol, VdLoopState *st) {
    if (st->stop) return;  // uniform
    (void)vd_adj_check(acc, j, tol, st, threadIdx.x, true);
}

// ---- message form.  to_in == 0: R_j goes to the out-slot of port j (raw, with its sum in
// S_out and Z_u in z_out when given).  to_in == 1: add[in-slot] + R_j goes to the in-slot of
// port j.  `add` may be `out` (each in-slot is read and written by the same lane).
// Gn / Sn: W is lambda, overrides are its normalisation cotangent (see vd_load_mats).
// resacc (adjoint sweeps): max|result - ref| and max|result| of the workgroup go, as the
// bit patterns of non-negative doubles, into resacc[rslot][0|1][blockIdx % 64] by atomic
// max, rslot = sweep % 3; workgroup 0 clears the slot of the next sweep (written two sweeps
// ago and checked one sweep ago).  lst (adjoint sweeps): the launch does nothing once
// lst->stop is set; when chk >= 0 it first checks sweep chk (the previous one) against tol
// and does nothing if that sweep met it.
template <int NS, bool HASW>
__global__ __launch_bounds__(64 * NS) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_vd_msg(const float *__restrict__ vsites, const float *__restrict__ msgs,
              const float *__restrict__ W, int op, double cv_const,
              const double *__restrict__ cv_vertex, const int4 *__restrict__ vin,
              const int4 *__restrict__ vout, const int4 *__restrict__ vdim,
              const int *__restrict__ vlist, const long long *__restrict__ msg_off, int to_in,
              const float *add, float *out,
              const float *__restrict__ ref, double *__restrict__ S_out,
              double *__restrict__ z_out, const float *__restrict__ Gn,
              const double *__restrict__ Sn, unsigned long long *__restrict__ resacc,
              int rslot, VdLoopState *lst, double tol, int chk,
              const float *__restrict__ W2, float *__restrict__ out_z) {
    extern __shared__ __align__(16) float vsm[];
    const int u = blockIdx.x;
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    if (lst) {  // adjoint loop control, uniform over the grid
        __shared__ int s_go;
        if (w == 0) {
            bool go = lst->stop == 0;
            if (go && chk >= 0) go = vd_adj_check(resacc, chk, tol, lst, lane, u == 0);
            if (lane == 0) s_go = go ? 1 : 0;
        }
        __syncthreads();
        if (!s_go) return;
    }
    float *bufA = vsm + w * 2 * VX_BUF;
    float *bufB = bufA + VX_BUF;
    float *msm = vsm + NS * 2 * VX_BUF;
    float *wsm0 = msm + 4 * 64;
    float *part = wsm0 + NS * 4 * 64;
    const VdSlots sl = vd_slots(vin, vout, vdim, msg_off, u);
    const int vid = vlist[u];
    const float cv = (float)(cv_vertex ? cv_vertex[vid] : cv_const);
    vd_load_mats<NS>(msgs, HASW ? W : nullptr, sl, 0, msm, wsm0, Gn, Sn, HASW ? W2 : nullptr);
    const float *wsm = wsm0 + (W2 ? w * 256 : 0);
    if (resacc && blockIdx.x == 0 && threadIdx.x < 128)
        resacc[((rslot + 1) % 3) * 128 + threadIdx.x] = 0ull;
    const int ks = op == OP_X ? (w ^ 1) : w;
    const float sg = (op == OP_Z && w == 1) ? -1.f : 1.f;
    float R1[64], R2[64], R3[64];
    {
        const float *own = vsites + ((long long)u * NS + w) * 4096;
        const float *kst = vsites + ((long long)u * NS + ks) * 4096;
#pragma unroll
        for (int r = 0; r < 64; ++r) bufA[r * VX_PITCH + lane] = own[r * 64 + lane];
#pragma unroll
        for (int r = 0; r < 64; ++r) R1[r] = kst[r * 64 + lane];
    }
    __syncthreads();
    float M[16], Wm[16];
    float o[4];

    // ---- phase A: E01 (and its tangent) -> Y; out2 (open p2), out3 (open p3)
    vx_load_m(msm + 0 * 64, M);
    vd_mp1<true>(R1, M, R2);
    if constexpr (HASW) { vx_load_m(wsm + 0 * 64, Wm); vd_mp1<true>(R1, Wm, R3); }
    vx_load_m(msm + 1 * 64, M);
    vd_mp1<false>(R2, M, R1);
    if constexpr (HASW) { vx_load_m(wsm + 1 * 64, Wm); vd_mp2<false>(R3, M, R2, Wm, R3); }
    vd_x_to_y(R1, bufB, lane);
    if constexpr (HASW) {
        vd_x_to_y(R3, bufB, lane);
        if (cv != 0.f) {
#pragma unroll
            for (int r = 0; r < 64; ++r) R3[r] += cv * R1[r];
        }
    }
    // both operands of the open contractions first, so E01 and its tangent die before
    // the (register-hungry) contractions: G2 -> R2, G3 -> R3 (or R1 without overrides)
    vx_load_m(msm + 3 * 64, M);
    if constexpr (HASW) { vx_load_m(wsm + 3 * 64, Wm); vd_mp2<false>(R3, M, R1, Wm, R2); }
    else vd_mp1<false>(R1, M, R2);
    vx_load_m(msm + 2 * 64, M);
    if constexpr (HASW) {
        vx_load_m(wsm + 2 * 64, Wm);
        vd_mp2<true>(R3, M, R1, Wm, R3);
        o[2] = vx_final_half<true, false, VX_PITCH>(R2, bufA);
        o[3] = vx_final_half<false, false, VX_PITCH>(R3, bufA);
    } else {
        vd_mp1<true>(R1, M, R1);
        o[2] = vx_final_half<true, false, VX_PITCH>(R2, bufA);
        o[3] = vx_final_half<false, false, VX_PITCH>(R1, bufA);
    }

    // ---- phase B: E23 (and its tangent) from the ket slice in Y -> X; out0, out1
    {
        const float *kbuf = vsm + ks * 2 * VX_BUF;
#pragma unroll
        for (int r = 0; r < 64; ++r) R1[r] = kbuf[lane * VX_PITCH + r];
    }
    vx_load_m(msm + 2 * 64, M);
    vd_mp1<true>(R1, M, R2);
    if constexpr (HASW) { vx_load_m(wsm + 2 * 64, Wm); vd_mp1<true>(R1, Wm, R3); }
    vx_load_m(msm + 3 * 64, M);
    vd_mp1<false>(R2, M, R1);
    if constexpr (HASW) { vx_load_m(wsm + 3 * 64, Wm); vd_mp2<false>(R3, M, R2, Wm, R3); }
    vd_y_to_x(R1, bufB, lane);
    if constexpr (HASW) {
        vd_y_to_x(R3, bufB, lane);
        if (cv != 0.f) {
#pragma unroll
            for (int r = 0; r < 64; ++r) R3[r] += cv * R1[r];
        }
    }
    vx_load_m(msm + 1 * 64, M);
    if constexpr (HASW) { vx_load_m(wsm + 1 * 64, Wm); vd_mp2<false>(R3, M, R1, Wm, R2); }
    else vd_mp1<false>(R1, M, R2);
    vx_load_m(msm + 0 * 64, M);
    if constexpr (HASW) {
        vx_load_m(wsm + 0 * 64, Wm);
        vd_mp2<true>(R3, M, R1, Wm, R3);
        o[0] = vx_final_half<true, true, VX_PITCH>(R2, bufA);
        o[1] = vx_final_half<false, true, VX_PITCH>(R3, bufA);
    } else {
        vd_mp1<true>(R1, M, R1);
        o[0] = vx_final_half<true, true, VX_PITCH>(R2, bufA);
        o[1] = vx_final_half<false, true, VX_PITCH>(R1, bufA);
    }

    // ---- sum over slices, write (lane l holds entry vx_split_entry(l) of each result)
    const float scale = HASW ? sg : sg * cv;
    const int ent_l = vx_split_entry(lane);
#pragma unroll
    for (int j = 0; j < 4; ++j) part[(w * 4 + j) * 64 + ent_l] = o[j] * scale;
    __syncthreads();
    double rd = 0.0, rv = 0.0;  // this wave's residual maxima (resacc)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        if (j % NS != w) continue;
        float x = 0.f, xz = 0.f;  // xz: the Z-inserted result (slice 1 negated, d = 2)
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            const float v = part[(s * 4 + j) * 64 + lane];
            x += v;
            xz += s == 1 ? -v : v;
        }
        if (j == 0 && z_out) {
            const double z = wave_sum_v((double)x * (double)msm[lane]);
            if (lane == 0) z_out[vid] = z;
        }
        // lane = entry (a, a') of the padded 8 x 8 block; real below the bond dimension
        const int D = sl.dims[j];
        const bool real = (lane >> 3) < D && (lane & 7) < D;
        const int ent = (lane >> 3) * D + (lane & 7);
        if (!to_in) {
            if (sl.outs[j] < 0) continue;
            if (out && real) out[sl.outo[j] + ent] = x;
            if (out_z && real) out_z[sl.outo[j] + ent] = xz;
            if (S_out) {
                const double S = wave_sum_v((double)x);
                if (lane == 0) S_out[sl.outs[j]] = S;
            }
        } else {
            if (sl.ins[j] < 0) continue;
            const long long off = sl.ino[j] + ent;
            const float v = real ? x + (add ? add[off] : 0.f) : 0.f;
            if (resacc) {
                const double dm = real ? fabs((double)v - (double)ref[off]) : 0.0;
                rd = (dm > rd || dm != dm) ? dm : rd;
                rv = fmax(rv, fabs((double)v));
            }
            if (real) out[off] = v;
        }
    }
    if (resacc) {  // uniform over the workgroup
        rd = wave_max_v(rd);
        rv = wave_max_v(rv);
        __syncthreads();  // part[] is free again
        double *red = reinterpret_cast<double *>(part);
        if (lane == 0) { red[2 * w] = rd; red[2 * w + 1] = rv; }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int i = 1; i < NS; ++i) {
                rd = (red[2 * i] > rd || red[2 * i] != red[2 * i]) ? red[2 * i] : rd;
                rv = fmax(rv, red[2 * i + 1]);
            }
            unsigned long long *slot = resacc + rslot * 128 + (blockIdx.x & 63);
            atomicMax(slot, (unsigned long long)__double_as_longlong(rd));
            atomicMax(slot + 64, (unsigned long long)__double_as_longlong(rv));
        }
    }
}

// ---- site form: da[u][s] += R[s] at the real (unpadded) entries of u's site tensor.
template <int NS, bool HASW>
__global__ __launch_bounds__(64 * NS) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_vd_site(const float *__restrict__ vsites, const float *__restrict__ msgs,
               const float *__restrict__ W, int trans, int op, double cv_const,
               const double *__restrict__ cv_vertex, const int4 *__restrict__ vin,
               const int4 *__restrict__ vout, const int4 *__restrict__ vdim,
               const int *__restrict__ vlist, const long long *__restrict__ msg_off,
               const long long *__restrict__ site_off,
               const int *__restrict__ port_off, float *__restrict__ da,
               const float *__restrict__ W2, const double *__restrict__ cv_swap) {
    extern __shared__ __align__(16) float vsm[];
    const int u = blockIdx.x;
    const int w = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    float *bufB = vsm + w * VX_BUF;
    float *msm = vsm + NS * VX_BUF;
    float *wsm0 = msm + 4 * 64;
    const VdSlots sl = vd_slots(vin, vout, vdim, msg_off, u);
    const int vid = vlist[u];
    const float cv = (float)(cv_vertex ? cv_vertex[vid] : cv_const);
    vd_load_mats<NS>(msgs, HASW ? W : nullptr, sl, trans, msm, wsm0, nullptr, nullptr,
                     HASW ? W2 : nullptr);
    const float *wsm = wsm0 + (W2 ? w * 256 : 0);
    const int ks = op == OP_X ? (w ^ 1) : w;
    const float sg = (op == OP_Z && w == 1) ? -1.f : 1.f;
    float R1[64], R2[64], R3[64];
    // cv_swap given (d = 2): the X-inserted value term cv_swap A[w ^ 1] x_all M of this
    // slice runs first, into a wave-private LDS block (entry r * 64 + lane), and joins the
    // result at the end; the tangent chain needs the value chain of A[w] alone
    const bool fold = cv_swap != nullptr;
    float *xbuf = wsm0 + NS * 4 * 64 + w * 4096;
    {
        const float *kst = vsites + ((long long)u * NS + ks) * 4096;
        if (fold) {
            const float cx = (float)cv_swap[vid];
            const float *xst = vsites + ((long long)u * NS + (w ^ 1)) * 4096;
#pragma unroll
            for (int r = 0; r < 64; ++r) R2[r] = cx * xst[r * 64 + lane];
        }
#pragma unroll
        for (int r = 0; r < 64; ++r) R1[r] = kst[r * 64 + lane];
    }
    __syncthreads();
    float M[16], Wm[16];
    if (fold) {
        vx_load_m(msm + 0 * 64, M);
        vd_mp1<true>(R2, M, R2);
        vx_load_m(msm + 1 * 64, M);
        vd_mp1<false>(R2, M, R2);
        vd_x_to_y(R2, bufB, lane);
        vx_load_m(msm + 2 * 64, M);
        vd_mp1<true>(R2, M, R2);
        vx_load_m(msm + 3 * 64, M);
        vd_mp1<false>(R2, M, R2);
#pragma unroll
        for (int r = 0; r < 64; ++r) xbuf[r * 64 + lane] = R2[r];
    }
    vx_load_m(msm + 0 * 64, M);
    vd_mp1<true>(R1, M, R2);                                                   // V1
    if constexpr (HASW) { vx_load_m(wsm + 0 * 64, Wm); vd_mp1<true>(R1, Wm, R3); }
    vx_load_m(msm + 1 * 64, M);
    vd_mp1<false>(R2, M, R1);                                                  // V2
    if constexpr (HASW) { vx_load_m(wsm + 1 * 64, Wm); vd_mp2<false>(R3, M, R2, Wm, R3); }
    vd_x_to_y(R1, bufB, lane);
    if constexpr (HASW) vd_x_to_y(R3, bufB, lane);
    vx_load_m(msm + 2 * 64, M);
    vd_mp1<true>(R1, M, R2);                                                   // V3
    if constexpr (HASW) { vx_load_m(wsm + 2 * 64, Wm); vd_mp2<true>(R3, M, R1, Wm, R3); }
    vx_load_m(msm + 3 * 64, M);
    vd_mp1<false>(R2, M, R1);                                                  // V4
    if constexpr (HASW) { vx_load_m(wsm + 3 * 64, Wm); vd_mp2<false>(R3, M, R2, Wm, R3); }
    const float c = sg * cv;
#pragma unroll
    for (int r = 0; r < 64; ++r) R1[r] = HASW ? c * R1[r] + sg * R3[r] : c * R1[r];
    if (fold) {
#pragma unroll
        for (int r = 0; r < 64; ++r) R1[r] += xbuf[r * 64 + lane];
    }
    vd_y_to_x(R1, bufB, lane);
    // entry e = r * 64 + lane of the padded slice: ports (r >> 3, r & 7, lane >> 3, lane & 7)
    const int k = port_off[vid + 1] - port_off[vid];
    long long n = 1;
#pragma unroll
    for (int j = 0; j < 4; ++j)
        if (j < k) n *= sl.dims[j];
    // every load is issued before the first store (the stores may alias later loads as far
    // as the compiler can tell, which would otherwise serialise the read-modify-writes)
    float *dst = da + site_off[vid] + (long long)w * n;
    // offset of padded entry (r, lane) in the real slice, or -1 for a padding entry
    auto real_off = [&](int r) {
        const int idx[4] = {r >> 3, r & 7, lane >> 3, lane & 7};
        int off = 0;
        bool real = true;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j < k) {
                off = off * sl.dims[j] + idx[j];
                real = real && idx[j] < sl.dims[j];
            } else {
                real = real && idx[j] == 0;
            }
        }
        return real ? off : -1;
    };
#pragma unroll
    for (int r = 0; r < 64; ++r) {
        const int off = real_off(r);
        R2[r] = off >= 0 ? dst[off] : 0.f;
    }
#pragma unroll
    for (int r = 0; r < 64; ++r) {
        const int off = real_off(r);
        if (off >= 0) dst[off] = R2[r] + R1[r];
    }
}

// ---- small kernels of the orchestration
// dm[rev q] = cv1[owner q] * GX[q] + cv0[owner q] * G[q]   (vertex terms of the energy)
__global__ void k_vd_vertex_terms(DevTopo tp, const float *G, const float *GX, const double *cv0,
                                  const double *cv1, float *dm) {
    const int q = blockIdx.x;
    if (q >= tp.nports) return;
    const int u = tp.owner[q];
    const long long mo = tp.msg_off[q], ro = tp.msg_off[tp.port_rev[q]];
    const int len = (int)(tp.msg_off[q + 1] - mo);
    const float a = (float)cv1[u], b = (float)cv0[u];
    for (int i = threadIdx.x; i < len; i += blockDim.x) dm[ro + i] = a * GX[mo + i] + b * G[mo + i];
}

// dst[q] = coef[q] * src[rev q]  (+ dst[q] when acc)
__global__ void k_vd_scale_rev(DevTopo tp, const double *coef, const float *src, float *dst,
                               int acc) {
    const int q = blockIdx.x;
    if (q >= tp.nports) return;
    const long long mo = tp.msg_off[q], ro = tp.msg_off[tp.port_rev[q]];
    const int len = (int)(tp.msg_off[q + 1] - mo);
    const float c = (float)coef[q];
    for (int i = threadIdx.x; i < len; i += blockDim.x)
        dst[mo + i] = c * src[ro + i] + (acc ? dst[mo + i] : 0.f);
}

__global__ void k_vd_axpy(const float *a, float *b, long long n) {
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n;
         i += (long long)gridDim.x * blockDim.x)
        b[i] += a[i];
}

__global__ void k_vd_inv(const double *z, double *out, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = 1.0 / z[i];
}

// ---------------------------------------------------------------- host side

// LDS: slice buffers, messages (4 x 64), overrides (one set per wave), partial results
static size_t vd_msg_lds(int ns) {
    return sizeof(float) * ((size_t)ns * 2 * VX_BUF + 256 + (size_t)ns * 256 + (size_t)ns * 256);
}
static size_t vd_site_lds(int ns, bool fold = true) {
    return sizeof(float) * ((size_t)ns * VX_BUF + 256 + (size_t)ns * 256 +
                            (fold ? (size_t)ns * 4096 : 0));
}

struct VdCtx {
    const VertexPlan *vp;
    hipStream_t st;
    DevTopo tp;
    const float *vsites;
};

static int vd_set_limits(int ns) {
    static bool done[5] = {false, false, false, false, false};
    if (ns < 1 || ns > 4) return 4;
    if (done[ns]) return 0;
    const void *fm[4][2] = {
        {(const void *)k_vd_msg<1, false>, (const void *)k_vd_msg<1, true>},
        {(const void *)k_vd_msg<2, false>, (const void *)k_vd_msg<2, true>},
        {(const void *)k_vd_msg<3, false>, (const void *)k_vd_msg<3, true>},
        {(const void *)k_vd_msg<4, false>, (const void *)k_vd_msg<4, true>}};
    const void *fs[4][2] = {
        {(const void *)k_vd_site<1, false>, (const void *)k_vd_site<1, true>},
        {(const void *)k_vd_site<2, false>, (const void *)k_vd_site<2, true>},
        {(const void *)k_vd_site<3, false>, (const void *)k_vd_site<3, true>},
        {(const void *)k_vd_site<4, false>, (const void *)k_vd_site<4, true>}};
    for (int h = 0; h < 2; ++h) {
        if (hipFuncSetAttribute(fm[ns - 1][h], hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)vd_msg_lds(ns)) != hipSuccess)
            return 2;
        if (hipFuncSetAttribute(fs[ns - 1][h], hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)vd_site_lds(ns)) != hipSuccess)
            return 2;
    }
    done[ns] = true;
    return 0;
}

static int vd_msg(const VdCtx &c, const float *msgs, const float *W, int op, double cv_const,
                  const double *cv_vertex, int to_in, const float *add, float *out,
                  const float *ref, double *S_out, double *z_out, const float *Gn = nullptr,
                  const double *Sn = nullptr, unsigned long long *resacc = nullptr,
                  int rslot = 0, VdLoopState *lst = nullptr, double tol = 0.0,
                  int chk = -1, const float *W2 = nullptr, float *out_z = nullptr) {
    const VertexPlan &p = *c.vp;
    const dim3 grid(p.nv), block(64 * p.d);
    const size_t lds = vd_msg_lds(p.d);
    const bool hw = W != nullptr;
    if (op != OP_I && p.d != 2) return 4;  // X and Z pair the two slices of a qubit
    if (W2 && (p.d != 2 || op != OP_I || !W || Gn)) return 4;
    if (out_z && (p.d != 2 || op != OP_I || W || to_in)) return 4;
#define VD_MSG(NS, HW)                                                                         \
    hipLaunchKernelGGL((k_vd_msg<NS, HW>), grid, block, lds, c.st, c.vsites, msgs, W, op,      \
                       cv_const, cv_vertex, p.d_vin, p.d_vout, p.d_vdim, p.d_vlist,            \
                       c.tp.msg_off, to_in,                                                    \
                       add, out, ref, S_out, z_out, Gn, Sn, resacc, rslot, lst, tol, chk, W2,    \
                       out_z)
    switch (p.d * 2 + (hw ? 1 : 0)) {
        case 2: VD_MSG(1, false); break;
        case 3: VD_MSG(1, true); break;
        case 4: VD_MSG(2, false); break;
        case 5: VD_MSG(2, true); break;
        case 6: VD_MSG(3, false); break;
        case 7: VD_MSG(3, true); break;
        case 8: VD_MSG(4, false); break;
        case 9: VD_MSG(4, true); break;
        default: return 4;
    }
#undef VD_MSG
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

// W2 (Z-inserted overrides) and cv_swap (X-inserted value coefficients) fold the Z and X
// launches of a d = 2 vertex into this op = I launch.
static int vd_site(const VdCtx &c, const float *msgs, const float *W, int trans, int op,
                   double cv_const, const double *cv_vertex, float *da,
                   const float *W2 = nullptr, const double *cv_swap = nullptr) {
    const VertexPlan &p = *c.vp;
    const dim3 grid(p.nv), block(64 * p.d);
    const size_t lds = vd_site_lds(p.d, cv_swap != nullptr);
    const bool hw = W != nullptr;
    if (op != OP_I && p.d != 2) return 4;
    if ((W2 || cv_swap) && (p.d != 2 || op != OP_I)) return 4;
    if (W2 && !W) return 4;
#define VD_SITE(NS, HW)                                                                         \
    hipLaunchKernelGGL((k_vd_site<NS, HW>), grid, block, lds, c.st, c.vsites, msgs, W, trans,   \
                       op, cv_const, cv_vertex, p.d_vin, p.d_vout, p.d_vdim, p.d_vlist,         \
                       c.tp.msg_off,                                                            \
                       c.tp.site_off, c.tp.port_off, da, W2, cv_swap)
    switch (p.d * 2 + (hw ? 1 : 0)) {
        case 2: VD_SITE(1, false); break;
        case 3: VD_SITE(1, true); break;
        case 4: VD_SITE(2, false); break;
        case 5: VD_SITE(2, true); break;
        case 6: VD_SITE(3, false); break;
        case 7: VD_SITE(3, true); break;
        case 8: VD_SITE(4, false); break;
        case 9: VD_SITE(4, true); break;
        default: return 4;
    }
#undef VD_SITE
    return hipGetLastError() == hipSuccess ? 0 : 2;
}

#define VD_RET(x)                  \
    do {                           \
        const int rc_ = (x);       \
        if (rc_) {                 \
            s->free_temps();       \
            return rc_;            \
        }                          \
    } while (0)

// Gradient of log Z_BP: dlogZ/dA_u = (bra + ket derivative of Z_u) / Z_u at the fixed point.
static int vadj_grad_lognorm(AdjointScratch *s, const VdCtx &c, const float *msgs, long long nsite,
                             const GradOut &go) {
    const int nv = c.tp.nv;
    VD_RET(vd_set_limits(c.vp->d));
    const bool direct = go.dev && !go.dev_f64;  // accumulate straight into the caller's buffer
    double *z = (double *)s->alloc(sizeof(double) * nv);
    double *iz = (double *)s->alloc(sizeof(double) * nv);
    float *da = direct ? (float *)go.dev : (float *)s->alloc(sizeof(float) * nsite);
    double *g64 = go.host ? (double *)s->alloc(sizeof(double) * nsite) : nullptr;
    if (!z || !iz || !da || (go.host && !g64)) VD_RET(5);
    VD_RET(vd_msg(c, msgs, nullptr, OP_I, 1.0, nullptr, 0, nullptr, nullptr, nullptr, nullptr, z,
                  nullptr, nullptr));
    hipLaunchKernelGGL(k_vd_inv, dim3((nv + 255) / 256), dim3(256), 0, c.st, z, iz, nv);
    if (hipMemsetAsync(da, 0, sizeof(float) * nsite, c.st) != hipSuccess) VD_RET(2);
    VD_RET(vd_site(c, msgs, nullptr, 0, OP_I, 0.0, iz, da));
    VD_RET(vd_site(c, msgs, nullptr, 1, OP_I, 0.0, iz, da));
    if (!direct) VD_RET(adj_emit_grad(c.st, (const float *)da, g64, nsite, go));
    const hipError_t e = hipStreamSynchronize(c.st);
    s->free_temps();
    return e == hipSuccess ? 0 : 2;
}

// TFIM energy E = sum_<uv> -J <Z_u Z_v>_BP + sum_v -h <X_v>_BP and its gradient.
static int vadj_tfim(AdjointScratch *s, const VdCtx &c, int ne, const float *msgs,
                     long long nsite, long long nmsg, double J, double h, double *energy,
                     const GradOut &go, double adj_tol, int adj_maxiter, int *out_iters,
                     double *out_res) {
    const DevTopo &tp = c.tp;
    const int nv = tp.nv, np = tp.nports;
    VD_RET(vd_set_limits(c.vp->d));
    double *z = (double *)s->alloc(sizeof(double) * nv);
    double *zx = (double *)s->alloc(sizeof(double) * nv);
    float *G = (float *)s->alloc(sizeof(float) * nmsg);
    float *GX = (float *)s->alloc(sizeof(float) * nmsg);
    float *GZ = (float *)s->alloc(sizeof(float) * nmsg);
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
    if (!z || !zx || !G || !GX || !GZ || !S || !nz || !dd || !cv1 || !cv0 || !cq1 || !cq0 ||
        !ev || !eq || !esum)
        VD_RET(5);
    // forward opens with I, X, Z on the ket: G (+ sums, Z_u) and GZ from one launch (Z only
    // negates slice 1), GX (+ Z_u^X)
    VD_RET(vd_msg(c, msgs, nullptr, OP_I, 1.0, nullptr, 0, nullptr, G, nullptr, S, z, nullptr,
                  nullptr, nullptr, 0, nullptr, 0.0, -1, nullptr, GZ));
    VD_RET(vd_msg(c, msgs, nullptr, OP_X, 1.0, nullptr, 0, nullptr, GX, nullptr, nullptr, zx,
                  nullptr, nullptr));
    if (np > 0)
        hipLaunchKernelGGL(k_adj_edge_dots<float>, dim3((np + 255) / 256), dim3(256), 0, c.st, tp,
                           (const float *)G, (const float *)GZ, nz, dd);
    const int nmax = std::max(nv, np);
    hipLaunchKernelGGL(k_adj_tfim_coefs, dim3((nmax + 255) / 256), dim3(256), 0, c.st, tp, J, h,
                       z, zx, nz, dd, cv1, cv0, cq1, cq0, ev, eq);
    hipLaunchKernelGGL(k_adj_sum2, dim3(1), dim3(BP_BLOCK), 0, c.st, ev, nv, eq, np, esum);
    if (!go.any()) {
        if (hipMemcpyAsync(energy, esum, sizeof(double), hipMemcpyDeviceToHost, c.st) !=
                hipSuccess ||
            hipStreamSynchronize(c.st) != hipSuccess)
            VD_RET(2);
        s->free_temps();
        return 0;
    }
    // with a gradient, the energy is read back with the loop state (esum stays untouched)
    const bool direct = go.dev && !go.dev_f64;
    float *dm = (float *)s->alloc(sizeof(float) * nmsg);
    float *WI = (float *)s->alloc(sizeof(float) * nmsg);
    float *WZ = (float *)s->alloc(sizeof(float) * nmsg);
    float *lbuf[2] = {(float *)s->alloc(sizeof(float) * nmsg),
                      (float *)s->alloc(sizeof(float) * nmsg)};
    float *gg = (float *)s->alloc(sizeof(float) * nmsg);
    float *da = direct ? (float *)go.dev : (float *)s->alloc(sizeof(float) * nsite);
    // residual accumulator (3 x 128 words) followed by the loop state, cleared together
    const size_t acc_bytes = sizeof(unsigned long long) * 384;
    unsigned long long *acc =
        (unsigned long long *)s->alloc(acc_bytes + sizeof(VdLoopState));
    VdLoopState *lst = acc ? (VdLoopState *)(acc + 384) : nullptr;
    double *g64 = go.host ? (double *)s->alloc(sizeof(double) * nsite) : nullptr;
    if (!dm || !WI || !WZ || !lbuf[0] || !lbuf[1] || !gg || !da || !acc || !lst ||
        (go.host && !g64))
        VD_RET(5);
    if (hipMemsetAsync(da, 0, sizeof(float) * nsite, c.st) != hipSuccess) VD_RET(2);
    // explicit dE/dm: vertex terms, then the edge terms (port t carries the other side's G / GZ)
    if (np > 0) {
        hipLaunchKernelGGL(k_vd_vertex_terms, dim3(np), dim3(64), 0, c.st, tp, (const float *)G,
                           (const float *)GX, (const double *)cv0, (const double *)cv1, dm);
        hipLaunchKernelGGL(k_vd_scale_rev, dim3(np), dim3(64), 0, c.st, tp, (const double *)cq0,
                           (const float *)G, WI, 0);
        hipLaunchKernelGGL(k_vd_scale_rev, dim3(np), dim3(64), 0, c.st, tp, (const double *)cq1,
                           (const float *)GZ, WZ, 0);
        // the I- and Z-inserted edge terms in one launch (overrides WI + sigma WZ per slice)
        VD_RET(vd_msg(c, msgs, WI, OP_I, 0.0, nullptr, 1, dm, dm, nullptr, nullptr, nullptr,
                      nullptr, nullptr, nullptr, 0, nullptr, 0.0, -1, WZ));
    }
    // adjoint sweeps lambda <- dE/dm + J_m^T lambda
    const int cblocks = (int)std::max<long long>(1, std::min<long long>((nmsg + 255) / 256, 4096));
    hipLaunchKernelGGL(k_adj_axpy_copy<float>, dim3(cblocks), dim3(256), 0, c.st,
                       (const float *)dm, lbuf[0], nmsg);
    // sweep i reads lbuf[i & 1] and writes lbuf[(i + 1) & 1]; sweeps queued after the
    // stop flag was set return at once, so lambda sits in lbuf[it & 1] for the `it` sweeps
    // that ran
    VdLoopState hst{0, 0, INFINITY};
    double hE = 0.0;
    if (hipMemsetAsync(acc, 0, acc_bytes + sizeof(VdLoopState), c.st) != hipSuccess) VD_RET(2);
    // sweeps queued between host reads of the loop state: 16, then doubling up to 64 (a
    // sweep queued after convergence costs one empty launch, a host read one round trip)
    int chunk = 16;
    int queued = 0;
    AdjFloor floor_;  // FP32 sweeps: stop at the rounding floor
    while (np > 0 && queued < adj_maxiter) {
        const int n = std::min(chunk, adj_maxiter - queued);
        chunk = std::min(2 * chunk, 64);
        for (int i = queued; i < queued + n; ++i) {
            float *lam = lbuf[i & 1], *nxt = lbuf[(i + 1) & 1];
            // lambda' = dE/dm + J_m^T lambda, the normalisation cotangent formed in the kernel
            VD_RET(vd_msg(c, msgs, lam, OP_I, 0.0, nullptr, 1, dm, nxt, lam, nullptr, nullptr, G,
                          S, acc, i % 3, lst, adj_tol, i - 1));
        }
        queued += n;
        hipLaunchKernelGGL(k_vd_adj_check, dim3(1), dim3(64), 0, c.st,
                           (const unsigned long long *)acc, queued - 1, adj_tol, lst);
        if (hipMemcpyAsync(&hst, lst, sizeof(VdLoopState), hipMemcpyDeviceToHost, c.st) !=
                hipSuccess ||
            hipStreamSynchronize(c.st) != hipSuccess)
            VD_RET(2);
        if (hst.stop || floor_.stalled(hst.r, hst.it)) break;
    }
    const int it = hst.it;
    const double r = np > 0 ? hst.r : INFINITY;
    const float *lam = lbuf[it & 1];
    // grad = dE/dA + J_A^T lambda: op I carries W = cq0 G_rev + gg(lambda), op Z carries cq1 GZ_rev,
    // op X and the value part of op I carry the vertex coefficients
    if (np > 0) {
        hipLaunchKernelGGL(k_adj_gg<float>, dim3(np), dim3(64), 0, c.st, tp, (const float *)lam,
                           (const float *)G, (const double *)S, gg);
        hipLaunchKernelGGL(k_vd_axpy, dim3(cblocks), dim3(256), 0, c.st, (const float *)gg, WI,
                           nmsg);
    }
    const float *wi = np > 0 ? WI : nullptr, *wz = np > 0 ? WZ : nullptr;
    // per side: the I (overrides wi, value cv0), Z (overrides wz) and X (value cv1) terms
    // in one launch
    for (int tr = 0; tr < 2; ++tr) VD_RET(vd_site(c, msgs, wi, tr, OP_I, 0.0, cv0, da, wz, cv1));
    if (!direct) VD_RET(adj_emit_grad(c.st, (const float *)da, g64, nsite, go));
    if (hipMemcpyAsync(&hE, esum, sizeof(double), hipMemcpyDeviceToHost, c.st) != hipSuccess)
        VD_RET(2);
    const hipError_t e = hipStreamSynchronize(c.st);
    *energy = hE;
    if (out_iters) *out_iters = it;
    if (out_res) *out_res = r;
    s->free_temps();
    return e == hipSuccess ? 0 : 2;
}

#undef VD_RET
