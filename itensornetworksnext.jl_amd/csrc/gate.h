// gate.h -- BP-gauged gate application: BPApplyGate / apply_gate_bp_nsite!
// (src/apply/apply_operators.jl:181-283), one workgroup per gate, FP64.
//
// Two-site gate on the bond between sites 1 and 2 (Val(2), :229-283):
//  1. every incoming message M on a port other than the shared one is factored
//     M = V diag(lam) V^T (parallel Jacobi over all the gate's messages); the gauge is
//     X = V diag(sqrt(lam)) (X X^T = M) and its pseudo-inverse Xi = diag(1/sqrt(lam)) V^T
//     (lam <= 1e-13 max(lam) treated as zero) -- the roles of gram_eigh_full_with_pinv;
//  2. each site tensor absorbs X on those ports and is laid out as the matrix
//     [other ports] x [(physical, shared bond)];
//  3. that matrix is factored Q R with Q orthonormal (qr_compact's role) by a one-sided
//     Jacobi SVD (X = U S W^T, Q = U, R = S W^T; the wide case transposed), which is
//     accurate for rank-deficient and ill-conditioned matrices;
//  4. Theta'[(i1,t1),(i2,t2)] = sum G[t1,t2,s1,s2] R1[i1,(s1,b)] R2[i2,(s2,b)] is split by the
//     same one-sided Jacobi SVD, singular values sorted, truncated (maxdim, relative cutoff)
//     and optionally normalised (svd_trunc, normalize);
//  5. the new site tensors are Q1 U sqrt(S) and Q2 V sqrt(S) with Xi absorbed back on the
//     other ports; the new bond takes the shared port's place, and the messages on the gate
//     edge become diag(S) both ways.
// One-site gate (Val(1), :214-227): the operator acts on the physical index; with
// normalize, the tensor is divided by the norm of the tensor with every port's gauge X.
//
// Every phase is a __host__ __device__ function of (thread id, thread count) with the
// block barrier supplied by the caller, so the same code also runs as a single-thread host
// loop.
#pragma once

#include <cmath>

#include "common.h"

#define GT_KMAX 8   // ports per site
#define GT_NT 256   // threads per workgroup
#define GT_FN __host__ __device__ inline

// gate descriptor (ints per gate)
enum {
    GD_KIND = 0,  // 1: one-site, 2: two-site
    GD_D1, GD_D2, GD_K1, GD_K2, GD_P1, GD_P2, GD_MAXDIM, GD_NORM,
    GD_DIMS1,
    GD_DIMS2 = GD_DIMS1 + GT_KMAX,
    GD_N = GD_DIMS2 + GT_KMAX
};

// per-gate offsets (long long): packed inputs, packed outputs, workspace
enum { GO_IN = 0, GO_OUT, GO_WS, GO_N };

#define GT_QMAX 128  // largest min(m1, m2) (kept ranks) a gate may have

// sizes and workspace layout of one gate, all in doubles
struct GtLayout {
    int kind, d1, d2, k1, k2, p1, p2, maxdim, norm;
    int dims1[GT_KMAX], dims2[GT_KMAX];
    int Db;                      // shared bond dimension (two-site)
    long long P1, P2;            // products of the other ports' dimensions
    long long size1, size2;      // site tensor sizes
    int n1, n2;                  // (physical, shared bond) columns
    int r1, r2;                  // QR ranks min(P, n)
    int m1, m2, q, chimax;       // Theta' shape, min(m1, m2), largest kept rank
    long long out1, out2;        // new site tensor capacities (chimax on the shared port)
    // workspace offsets
    long long msg[2][GT_KMAX];   // eigen blocks per other port: W, V, X, Xi (D^2 each), lam
    long long rot;               // rotation table of the message eigensolves (2 per pair)
    long long bufA[2], bufB[2], jac[2], jv[2], Q[2], R[2], ssig[2];
    long long th, thp, tv, sig, perm, raw, ws_total;
    int nmsg1, nmsg2;            // messages per site (other ports; one-site: all ports)
    long long in_total, out_total;
};

GT_FN long long gt_max(long long a, long long b) { return a > b ? a : b; }
GT_FN long long gt_min(long long a, long long b) { return a < b ? a : b; }

// Fills the layout from a descriptor; returns 0, or 1 for an invalid descriptor.
GT_FN int gt_layout(const int *dsc, GtLayout &L) {
    L.kind = dsc[GD_KIND];
    L.d1 = dsc[GD_D1];
    L.d2 = dsc[GD_D2];
    L.k1 = dsc[GD_K1];
    L.k2 = dsc[GD_K2];
    L.p1 = dsc[GD_P1];
    L.p2 = dsc[GD_P2];
    L.maxdim = dsc[GD_MAXDIM];
    L.norm = dsc[GD_NORM];
    if (L.kind != 1 && L.kind != 2) return 1;
    if (L.d1 < 1 || L.k1 < 0 || L.k1 > GT_KMAX) return 1;
    for (int j = 0; j < L.k1; ++j) {
        L.dims1[j] = dsc[GD_DIMS1 + j];
        if (L.dims1[j] < 1) return 1;
    }
    long long ws = 0;
    auto take = [&](long long n) {
        const long long at = ws;
        ws += (n + 3) / 4 * 4;
        return at;
    };
    if (L.kind == 1) {
        L.size1 = L.d1;
        for (int j = 0; j < L.k1; ++j) L.size1 *= L.dims1[j];
        L.nmsg1 = L.norm ? L.k1 : 0;
        L.nmsg2 = 0;
        long long npair = 0;
        for (int j = 0; j < L.nmsg1; ++j) {
            const long long D = L.dims1[j];
            L.msg[0][j] = take(4 * D * D + D);
            npair += (D + 1) / 2;
        }
        L.rot = take(2 * npair + 2);
        L.bufA[0] = take(L.size1);
        L.bufB[0] = take(L.size1);
        L.out1 = L.size1;
        L.out2 = 0;
        L.chimax = 0;
        L.ws_total = ws;
        L.in_total = L.size1 + (long long)L.d1 * L.d1;
        for (int j = 0; j < L.nmsg1; ++j) L.in_total += (long long)L.dims1[j] * L.dims1[j];
        L.out_total = L.out1;
        return 0;
    }
    if (L.d2 < 1 || L.k2 < 1 || L.k2 > GT_KMAX || L.k1 < 1) return 1;
    if (L.p1 < 0 || L.p1 >= L.k1 || L.p2 < 0 || L.p2 >= L.k2) return 1;
    for (int j = 0; j < L.k2; ++j) {
        L.dims2[j] = dsc[GD_DIMS2 + j];
        if (L.dims2[j] < 1) return 1;
    }
    if (L.dims1[L.p1] != L.dims2[L.p2]) return 1;
    L.Db = L.dims1[L.p1];
    L.P1 = L.P2 = 1;
    for (int j = 0; j < L.k1; ++j)
        if (j != L.p1) L.P1 *= L.dims1[j];
    for (int j = 0; j < L.k2; ++j)
        if (j != L.p2) L.P2 *= L.dims2[j];
    L.n1 = L.d1 * L.Db;
    L.n2 = L.d2 * L.Db;
    L.size1 = L.P1 * L.n1;
    L.size2 = L.P2 * L.n2;
    // 32-bit site indices: a grid-stride loop idx += GT_NT must not pass INT_MAX
    if (L.size1 > 0x7fffffffLL - GT_NT || L.size2 > 0x7fffffffLL - GT_NT) return 1;
    L.r1 = (int)gt_min(L.P1, L.n1);
    L.r2 = (int)gt_min(L.P2, L.n2);
    L.m1 = L.r1 * L.d1;
    L.m2 = L.r2 * L.d2;
    L.q = L.m1 < L.m2 ? L.m1 : L.m2;
    L.chimax = L.maxdim > 0 && L.maxdim < L.q ? L.maxdim : L.q;
    L.out1 = L.P1 * L.d1 * L.chimax;
    L.out2 = L.P2 * L.d2 * L.chimax;
    L.nmsg1 = L.k1 - 1;
    L.nmsg2 = L.k2 - 1;
    long long npair = 0;
    for (int s = 0; s < 2; ++s) {
        const int k = s ? L.k2 : L.k1, p = s ? L.p2 : L.p1;
        const int *dims = s ? L.dims2 : L.dims1;
        for (int j = 0; j < k; ++j) {
            const long long D = dims[j];
            L.msg[s][j] = j == p ? -1 : take(4 * D * D + D);
            if (j != p) npair += (D + 1) / 2;
        }
    }
    L.rot = take(2 * npair + 2);
    for (int s = 0; s < 2; ++s) {
        const long long P = s ? L.P2 : L.P1;
        const long long n = s ? L.n2 : L.n1;
        const long long d = s ? L.d2 : L.d1;
        const long long sz = gt_max(P * n, P * d * L.chimax);
        const long long r = gt_min(P, n);
        L.bufA[s] = take(sz);
        L.bufB[s] = take(sz);
        L.jac[s] = take(P * n);
        L.jv[s] = take(r * r);
        L.Q[s] = take(P * r);
        L.R[s] = take(r * n);
        L.ssig[s] = take(n);
    }
    if (L.q > GT_QMAX) return 1;
    L.th = take((long long)L.m1 * L.m2);
    L.thp = take((long long)L.m1 * L.m2);
    L.tv = take((long long)L.q * L.q);
    L.sig = take(L.q);
    L.perm = take(L.q);
    L.raw = take(L.q);
    L.ws_total = ws;
    L.in_total = L.size1 + L.size2 + (long long)L.d1 * L.d2 * L.d1 * L.d2;
    for (int j = 0; j < L.k1; ++j)
        if (j != L.p1) L.in_total += (long long)L.dims1[j] * L.dims1[j];
    for (int j = 0; j < L.k2; ++j)
        if (j != L.p2) L.in_total += (long long)L.dims2[j] * L.dims2[j];
    L.out_total = L.out1 + L.out2 + L.chimax;
    return 0;
}

// doubles of LDS that stage every phase of the gate: the message eigensolves, the Jacobi
// sweeps, the rebuild's factors
GT_FN long long gt_lds_need(const GtLayout &L) {
    long long g = 0;  // the message eigensolves: W and V of every message
    for (int j = 0; j < L.k1; ++j)
        if (L.kind == 1 ? L.norm : j != L.p1) g += 2LL * L.dims1[j] * L.dims1[j];
    if (L.kind != 2) return g;
    for (int j = 0; j < L.k2; ++j)
        if (j != L.p2) g += 2LL * L.dims2[j] * L.dims2[j];
    const long long s1 = L.P1 * L.n1 + (long long)L.r1 * L.r1;
    const long long s2 = L.P2 * L.n2 + (long long)L.r2 * L.r2;
    const long long t = (long long)L.m1 * L.m2 + (long long)L.q * L.q;
    const long long b1 = L.P1 * L.r1 + (long long)L.r1 * L.d1 * L.chimax;
    const long long b2 = L.P2 * L.r2 + (long long)L.r2 * L.d2 * L.chimax;
    return gt_max(gt_max(gt_max(s1, s2), gt_max(t, g)), gt_max(b1, b2));
}

// ------------------------------------------------------------------ block helpers

struct GtDev {
    int tid, nt;
    double *red;   // [nt] reduction scratch
    int *flag;     // one int
    double *lds;   // dynamic LDS for staged Jacobi sweeps
    long long lds_cap;
    __device__ void sync() const { __syncthreads(); }
    __device__ void mark(int) const {}  // phase boundary (a hook for instrumented contexts)
    // sum over an aligned group of gs lanes (a power of two <= 64, uniform over the wave)
    // on the VALU: DPP partners inside 16-lane rows, then the row and half swaps; every
    // partner lies in the same group, so lanes of other (possibly inactive) groups are not
    // read
    __device__ double group_sum(double v, int gs) const {
        if (gs >= 2) v += row_partner<1>(v);
        if (gs >= 4) v += row_partner<2>(v);
        if (gs >= 8) v += row_partner<4>(v);
        if (gs >= 16) v += row_partner<8>(v);
        if (gs >= 32) {
            const DPair r = rows16_pair(v);
            v = r.first + r.second;
        }
        if (gs >= 64) {
            const DPair h = halves32_pair(v);
            v = h.first + h.second;
        }
        return v;
    }
};

// sum over the block of one value per thread (all threads get the total)
template <class Ctx>
GT_FN double gt_block_sum(const Ctx &c, double v) {
    c.red[c.tid] = v;
    c.sync();
    double s = 0.0;
    for (int i = 0; i < c.nt; ++i) s += c.red[i];  // every thread sums in the same order
    c.sync();
    return s;
}

// Gauges of a gate's messages, all at once: for message i (dimension D[i], input M[i],
// block E[i] = W, V, X, Xi (D^2 each), lam (D)) the symmetric part W = (M + M^T) / 2 is
// diagonalised W = V diag(lam) V^T by parallel two-sided Jacobi: every round applies the
// disjoint plane rotations of a round-robin pairing of every message together (angles,
// then the column pass W J and V J, then the row pass J^T W), until no off-diagonal
// element is above 1e-15 of its diagonal pair.  Then X = V sqrt(lam) (X X^T = W) and
// Xi = pinv(X), lam <= 1e-13 max(lam) dropped.  `rot` holds (c, s) per (message, pair).
template <class Ctx>
GT_FN void gt_gauges(const Ctx &c, int nm, const double *const *M, double *const *E,
                     const int *Ds, double *rot) {
    int dmax = 0;
    for (int i = 0; i < nm; ++i) dmax = Ds[i] > dmax ? Ds[i] : dmax;
    long long tot = 0;
    for (int i = 0; i < nm; ++i) tot += (long long)Ds[i] * Ds[i];
    // W and V of every message: in LDS when they all fit, else in the E blocks
    double *Wp[2 * GT_KMAX], *Vp[2 * GT_KMAX];
    {
        const bool staged = 2 * tot <= c.lds_cap;
        long long off = 0;
        for (int i = 0; i < nm; ++i) {
            const long long DD = (long long)Ds[i] * Ds[i];
            Wp[i] = staged ? c.lds + off : E[i];
            Vp[i] = staged ? c.lds + off + DD : E[i] + DD;
            off += 2 * DD;
        }
    }
    // W = sym(M), V = I
    for (long long t = c.tid; t < tot; t += c.nt) {
        long long rem = t;
        int i = 0;
        while (rem >= (long long)Ds[i] * Ds[i]) rem -= (long long)Ds[i] * Ds[i], ++i;
        const int D = Ds[i], a = (int)(rem / D), b = (int)(rem % D);
        Wp[i][a * D + b] = 0.5 * (M[i][a * D + b] + M[i][b * D + a]);
        Vp[i][a * D + b] = a == b ? 1.0 : 0.0;
    }
    c.sync();
    const int ne = dmax + (dmax & 1);
    int npair = 0;
    for (int i = 0; i < nm; ++i) npair += (Ds[i] + 1) / 2;
    for (int sweep = 0; sweep < 60 && ne > 1; ++sweep) {
        if (c.tid == 0) *c.flag = 0;
        c.sync();
        for (int r = 0; r < ne - 1; ++r) {
            // angles, one task per (message, pair)
            for (int t = c.tid; t < npair; t += c.nt) {
                int i = 0, k = t;
                while (k >= (Ds[i] + 1) / 2) k -= (Ds[i] + 1) / 2, ++i;
                const int D = Ds[i], nei = D + (D & 1);
                double cs = 1.0, sn = 0.0;
                if (r < nei - 1) {
                    int p, q;
                    if (k == 0) {
                        p = r;
                        q = nei - 1;
                    } else {
                        p = (r + k) % (nei - 1);
                        q = (r - k + nei - 1) % (nei - 1);
                    }
                    if (p > q) {
                        const int x = p;
                        p = q;
                        q = x;
                    }
                    const double *W = Wp[i];
                    if (q < D) {
                        const double apq = W[p * D + q], app = W[p * D + p], aqq = W[q * D + q];
                        if (apq != 0.0 && fabs(apq) > 1e-15 * sqrt(fabs(app * aqq))) {
                            const double th = (aqq - app) / (2.0 * apq);
                            const double tt =
                                (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                            cs = 1.0 / sqrt(tt * tt + 1.0);
                            sn = tt * cs;
                            if (fabs(apq) > 1e-13 * sqrt(fabs(app * aqq))) *c.flag = 1;
                        }
                    }
                }
                rot[2 * t] = cs;
                rot[2 * t + 1] = sn;
            }
            c.sync();
            // column pass: W <- W J, V <- V J (one task per message, pair, row)
            for (int pass = 0; pass < 2; ++pass) {
                long long ntask = 0;
                for (int i = 0; i < nm; ++i) ntask += (long long)((Ds[i] + 1) / 2) * Ds[i];
                for (long long t = c.tid; t < ntask; t += c.nt) {
                    long long rem = t;
                    int i = 0, base = 0;
                    while (rem >= (long long)((Ds[i] + 1) / 2) * Ds[i]) {
                        rem -= (long long)((Ds[i] + 1) / 2) * Ds[i];
                        base += (Ds[i] + 1) / 2;
                        ++i;
                    }
                    const int D = Ds[i], nei = D + (D & 1);
                    const int k = (int)(rem / D), x = (int)(rem % D);
                    if (r >= nei - 1) continue;
                    int p, q;
                    if (k == 0) {
                        p = r;
                        q = nei - 1;
                    } else {
                        p = (r + k) % (nei - 1);
                        q = (r - k + nei - 1) % (nei - 1);
                    }
                    if (p > q) {
                        const int y = p;
                        p = q;
                        q = y;
                    }
                    if (q >= D) continue;
                    const double cs = rot[2 * (base + k)], sn = rot[2 * (base + k) + 1];
                    if (sn == 0.0) continue;
                    double *W = Wp[i], *V = Vp[i];
                    if (pass == 0) {  // columns p, q of W and V, row x
                        const double wp = W[x * D + p], wq = W[x * D + q];
                        W[x * D + p] = cs * wp - sn * wq;
                        W[x * D + q] = sn * wp + cs * wq;
                        const double vp = V[x * D + p], vq = V[x * D + q];
                        V[x * D + p] = cs * vp - sn * vq;
                        V[x * D + q] = sn * vp + cs * vq;
                    } else {  // rows p, q of W, column x
                        const double wp = W[p * D + x], wq = W[q * D + x];
                        W[p * D + x] = cs * wp - sn * wq;
                        W[q * D + x] = sn * wp + cs * wq;
                    }
                }
                c.sync();
            }
        }
        const int more = *c.flag;
        c.sync();
        if (!more) break;
    }
    // lam, X, Xi
    for (long long t = c.tid; t < tot; t += c.nt) {
        long long rem = t;
        int i = 0;
        while (rem >= (long long)Ds[i] * Ds[i]) rem -= (long long)Ds[i] * Ds[i], ++i;
        const int D = Ds[i], a = (int)(rem / D), x = (int)(rem % D);
        const double *W = Wp[i], *V = Vp[i];
        double lmax = 0.0;
        for (int y = 0; y < D; ++y) lmax = W[y * D + y] > lmax ? W[y * D + y] : lmax;
        const double l = W[x * D + x] > 1e-13 * lmax ? W[x * D + x] : 0.0;
        const double sq = sqrt(l), inv = l > 0.0 ? 1.0 / sq : 0.0;
        E[i][2 * D * D + a * D + x] = V[a * D + x] * sq;
        E[i][3 * D * D + x * D + a] = V[a * D + x] * inv;
        if (a == 0) E[i][4 * D * D + x] = W[x * D + x];
    }
    c.sync();
}

template <typename I, class Ctx>
GT_FN void gt_mode_idx(const Ctx &c, const double *in, double *out, I size, I inner, int D,
                       const double *Mx) {
    for (I idx = c.tid; idx < size; idx += c.nt) {
        const I q = idx / inner;
        const I i = idx - q * inner, o = q / D;
        const int b = (int)(q - o * D);
        const double *src = in + (long long)o * inner * D + i;
        double s = 0.0;
#pragma unroll 4
        for (int a = 0; a < D; ++a) s += src[(long long)a * inner] * Mx[a * D + b];
        out[idx] = s;
    }
}

// out = in x_j Mx (mode product on port j of a [d0][dims...] tensor):
// out[o, b, i] = sum_a in[o, a, i] Mx[a * D + b]; `dims` are the port dimensions (the
// port j dimension D is unchanged)
template <class Ctx>
GT_FN void gt_mode(const Ctx &c, const double *in, double *out, long long size, const int *dims,
                   int k, int j, const double *Mx) {
    long long inner = 1;
    for (int l = j + 1; l < k; ++l) inner *= dims[l];
    const int D = dims[j];
    if (size <= 0x7fffffffLL - GT_NT)  // 32-bit index arithmetic (64-bit division is a long
                                       // routine); idx += nt stays below INT_MAX
        gt_mode_idx<int>(c, in, out, (int)size, (int)inner, D, Mx);
    else
        gt_mode_idx<long long>(c, in, out, size, inner, D, Mx);
}

// One-sided Jacobi SVD of A (rows x cols, column-major: column c at A + c * rows): the
// column pairs of a round-robin schedule (odd cols get a phantom partner) are rotated by
// groups of gs threads (a power of two <= 64, one wave at most) with group reductions; on
// return the columns of A are U S and V (cols x cols, column-major) holds the right vectors.
template <class Ctx>
GT_FN void gt_jacobi(const Ctx &c, double *A, int rows, int cols, double *V) {
    for (long long i = c.tid; i < (long long)cols * cols; i += c.nt)
        V[i] = (i / cols) == (i % cols) ? 1.0 : 0.0;
    const int ne = cols + (cols & 1);  // even number of players
    const int np = ne / 2;
    int gs = 1;
    while (gs * 2 <= 64 && gs * 2 * np <= c.nt) gs *= 2;
    const int g = c.tid / gs, l = c.tid % gs;
    c.sync();
    for (int sweep = 0; sweep < 40; ++sweep) {
        if (c.tid == 0) *c.flag = 0;
        c.sync();
        for (int r = 0; r < ne - 1; ++r) {
            // the threads of groups 0..np-1 each rotate one pair of this round
            for (int pg = g; pg < np; pg += c.nt / gs > 0 ? c.nt / gs : 1) {
                int i, jj;
                if (pg == 0) {
                    i = r;
                    jj = ne - 1;
                } else {
                    i = (r + pg) % (ne - 1);
                    jj = (r - pg + ne - 1) % (ne - 1);
                }
                if (i > jj) {
                    const int t = i;
                    i = jj;
                    jj = t;
                }
                const bool real = jj < cols;  // a phantom partner leaves i alone
                double al = 0.0, be = 0.0, ga = 0.0;
                double *ai = A + (long long)i * rows, *aj = A + (long long)jj * rows;
                if (real)
#pragma unroll 4
                    for (int x = l; x < rows; x += gs) {
                        const double u = ai[x], v = aj[x];
                        al += u * u;
                        be += v * v;
                        ga += u * v;
                    }
                al = c.group_sum(al, gs);
                be = c.group_sum(be, gs);
                ga = c.group_sum(ga, gs);
                if (real && fabs(ga) > 1e-15 * sqrt(al * be) && ga != 0.0) {
                    const double ze = (be - al) / (2.0 * ga);
                    const double t = (ze >= 0.0 ? 1.0 : -1.0) / (fabs(ze) + sqrt(1.0 + ze * ze));
                    const double cs = 1.0 / sqrt(1.0 + t * t), sn = cs * t;
#pragma unroll 4
                    for (int x = l; x < rows; x += gs) {
                        const double u = ai[x], v = aj[x];
                        ai[x] = cs * u - sn * v;
                        aj[x] = sn * u + cs * v;
                    }
                    double *vi = V + (long long)i * cols, *vj = V + (long long)jj * cols;
                    for (int x = l; x < cols; x += gs) {
                        const double u = vi[x], v = vj[x];
                        vi[x] = cs * u - sn * v;
                        vj[x] = sn * u + cs * v;
                    }
                    if (l == 0 && fabs(ga) > 1e-13 * sqrt(al * be)) *c.flag = 1;
                }
            }
            c.sync();
        }
        const int more = *c.flag;
        c.sync();
        if (!more) break;
    }
}

// gt_jacobi in LDS when the matrix and V fit in the context's `lds` (lds_cap doubles):
// copied in, rotated there (LDS latency instead of a global round trip per step), copied
// back; in global memory otherwise.
template <class Ctx>
GT_FN void gt_jacobi_staged(const Ctx &c, double *A, int rows, int cols, double *V) {
    const long long na = (long long)rows * cols, nv = (long long)cols * cols;
    if (na + nv > c.lds_cap) {
        gt_jacobi(c, A, rows, cols, V);
        return;
    }
    double *la = c.lds, *lv = c.lds + na;
    for (long long i = c.tid; i < na; i += c.nt) la[i] = A[i];
    c.sync();
    gt_jacobi(c, la, rows, cols, lv);
    for (long long i = c.tid; i < na; i += c.nt) A[i] = la[i];
    for (long long i = c.tid; i < nv; i += c.nt) V[i] = lv[i];
    c.sync();
}

// Householder QR of a tall matrix A (rows >= cols, column-major, in place), LAPACK's geqr2
// followed by org2r: R (cols x cols, row-major, upper triangular) is written to `R`, and A
// is overwritten by Q (rows x cols, orthonormal columns, Q R = the input).  Column c belongs
// to thread group c % ng (groups of gs threads, gs a power of two <= 64, gs * cols <= nt
// when possible) and row i to thread i % gs of a group, in every step: a thread reads
// elements of its group's columns only in rows it owns, and values from other threads
// arrive through group_sum (dot products, the diagonal element) or after a barrier.  A
// factorisation step needs one barrier (after the owner of column j has formed its
// reflector), a Q step two.  tau: cols doubles of scratch.
template <class Ctx>
GT_FN void gt_householder(const Ctx &c, double *A, int rows, int cols, double *tau,
                          double *R) {
    int gs = 1;
    while (gs * 2 <= 64 && gs * 2 * cols <= c.nt) gs *= 2;
    const int ng = c.nt / gs > 0 ? c.nt / gs : 1;
    const int g = c.tid / gs, l = c.tid % gs;
    const bool in_group = g < ng;
    // first row >= i0 that this thread owns
    auto first = [&](int i0) { return i0 + ((l - i0 % gs) % gs + gs) % gs; };
    // H_j (v_j = 1, v below the diagonal of column j, factor tj) applied to this group's
    // columns > j, rows >= j
    auto reflect = [&](int j, double tj) {
        const double *aj = A + (long long)j * rows;
        const int i1 = first(j + 1);
        for (int col = j + 1 + ((g - (j + 1) % ng) % ng + ng) % ng; col < cols; col += ng) {
            double *ac = A + (long long)col * rows;
            double w = l == j % gs ? ac[j] : 0.0;  // v_j = 1; row j is this thread's
#pragma unroll 8
            for (int i = i1; i < rows; i += gs) w += aj[i] * ac[i];
            w = c.group_sum(w, gs) * tj;
            if (l == j % gs) ac[j] -= w;
#pragma unroll 8
            for (int i = i1; i < rows; i += gs) ac[i] -= w * aj[i];
        }
    };
    c.sync();
    // factorisation: reflector j maps A[j:, j] to beta e_j
    for (int j = 0; j < cols; ++j) {
        double *aj = A + (long long)j * rows;
        if (in_group && j % ng == g) {
            const int i1 = first(j + 1);
            // alpha = A[j, j] from the thread that owns row j
            const double alpha = c.group_sum(l == j % gs ? aj[j] : 0.0, gs);
            double s2 = 0.0;
#pragma unroll 8
            for (int i = i1; i < rows; i += gs) s2 += aj[i] * aj[i];
            s2 = c.group_sum(s2, gs);
            double t = 0.0;
            if (s2 > 0.0) {
                const double nrm = sqrt(alpha * alpha + s2);
                const double beta = alpha >= 0.0 ? -nrm : nrm;
                t = (beta - alpha) / beta;
                const double sc = 1.0 / (alpha - beta);
#pragma unroll 8
                for (int i = i1; i < rows; i += gs) aj[i] *= sc;
                if (l == j % gs) aj[j] = beta;
            }
            if (l == 0) tau[j] = t;
        }
        c.sync();
        const double tj = tau[j];
        if (in_group && tj != 0.0) reflect(j, tj);
    }
    c.sync();
    for (long long idx = c.tid; idx < (long long)cols * cols; idx += c.nt) {
        const int i = (int)(idx / cols), col = (int)(idx % cols);
        R[idx] = col >= i ? A[(long long)col * rows + i] : 0.0;
    }
    c.sync();
    // Q = H_0 ... H_{cols-1} [I; 0], accumulated backwards in place
    for (int j = cols - 1; j >= 0; --j) {
        const double tj = tau[j];
        if (in_group && tj != 0.0) reflect(j, tj);
        c.sync();
        if (in_group && j % ng == g) {
            double *aj = A + (long long)j * rows;
            for (int i = first(0); i < rows; i += gs) {
                if (i < j) aj[i] = 0.0;
                else if (i == j) aj[i] = 1.0 - tj;
                else aj[i] *= -tj;
            }
        }
        c.sync();
    }
}

// ------------------------------------------------------------------ the pipeline

// Site s's gauged matrix, factored: Q (P x r, row-major) and R (r x n, row-major).
template <class Ctx>
GT_FN void gt_site_qr(const Ctx &c, const GtLayout &L, int s, const double *A, double *ws) {
    const int k = s ? L.k2 : L.k1, p = s ? L.p2 : L.p1, d = s ? L.d2 : L.d1;
    const int *dims = s ? L.dims2 : L.dims1;
    const long long P = s ? L.P2 : L.P1, size = s ? L.size2 : L.size1;
    const int n = s ? L.n2 : L.n1, Db = L.Db;
    const int r = (int)gt_min(P, n);
    double *cur = ws + L.bufA[s], *nxt = ws + L.bufB[s];
    // gauge every other port
    const double *src = A;
    for (int j = 0; j < k; ++j) {
        if (j == p) continue;
        gt_mode(c, src, nxt, size, dims, k, j, ws + L.msg[s][j] + 2 * dims[j] * dims[j]);
        c.sync();
        src = nxt;
        double *t = cur;
        cur = nxt;
        nxt = t;
    }
    if (s == 0) c.mark(9);
    // to the Jacobi layout: tall (P >= n) columns (t, b), wide columns x
    double *J = ws + L.jac[s];
    const bool tall = P >= n;
    // (32-bit index arithmetic: the layout rejects site tensors of INT_MAX - GT_NT entries
    // or more, so idx += nt cannot overflow)
    for (int idx = c.tid; idx < (int)size; idx += c.nt) {
        int rem = idx, x = 0, xs = 1;
        int bb = 0;
        for (int j = k - 1; j >= 0; --j) {
            const int qt = rem / dims[j];
            const int a = rem - qt * dims[j];
            rem = qt;
            if (j == p) {
                bb = a;
            } else {
                x += a * xs;
                xs *= dims[j];
            }
        }
        const int t = rem;
        const long long col = (long long)t * Db + bb;
        if (tall)
            J[col * P + x] = src[idx];
        else
            J[(long long)x * n + col] = src[idx];
    }
    c.sync();
    double *JV = ws + L.jv[s];
    double *Q = ws + L.Q[s], *R = ws + L.R[s];
    if (tall) {
        // Householder QR (in LDS when the matrix fits), Q written row-major
        const bool staged = P * n + n <= c.lds_cap;
        double *A = staged ? c.lds : J;
        double *tau = staged ? c.lds + P * n : ws + L.ssig[s];
        if (staged)
            for (long long i = c.tid; i < P * n; i += c.nt) A[i] = J[i];
        if (s == 0) c.mark(10);
        gt_householder(c, A, (int)P, n, tau, R);
        if (s == 0) c.mark(11);
        for (long long idx = c.tid; idx < P * n; idx += c.nt) {
            const long long x = idx / n;
            const int i = (int)(idx % n);
            Q[idx] = A[(long long)i * P + x];
        }
        c.sync();
        (void)d;
        return;
    }
    gt_jacobi_staged(c, J, n, (int)P, JV);
    // Q and R
    if (tall) {
        // columns of J: U S (P x n); V = JV (n x n)
        for (int col = c.tid; col < n; col += c.nt) {
            double s2 = 0.0;
#pragma unroll 8
            for (long long x = 0; x < P; ++x) s2 += J[col * P + x] * J[col * P + x];
            ws[L.ssig[s] + col] = sqrt(s2);
        }
        c.sync();
        for (long long idx = c.tid; idx < P * r; idx += c.nt) {
            const long long x = idx / r;
            const int i = (int)(idx % r);
            const double sg = ws[L.ssig[s] + i];
            Q[idx] = sg > 0.0 ? J[(long long)i * P + x] / sg : 0.0;
        }
        for (long long idx = c.tid; idx < (long long)r * n; idx += c.nt) {
            const int i = (int)(idx / n), col = (int)(idx % n);
            R[idx] = ws[L.ssig[s] + i] * JV[(long long)i * n + col];
        }
    } else {
        // X^T (n x P) = U' S V'^T: Q = V' (P x P), R = S U'^T = columns of J transposed
        for (long long idx = c.tid; idx < P * r; idx += c.nt) {
            const long long x = idx / r;
            const int i = (int)(idx % r);
            Q[idx] = JV[(long long)i * P + x];
        }
        for (long long idx = c.tid; idx < (long long)r * n; idx += c.nt) {
            const int i = (int)(idx / n), col = (int)(idx % n);
            R[idx] = J[(long long)i * n + col];
        }
    }
    c.sync();
    (void)d;
}

// new site tensor of site s from its Q and the kept singular vectors:
// psi[t][ports, chi at p] = sum_i Q[x, i] F[(i, t), c] with F[(i, t), c] = coef(i, t, c);
// then the inverse gauges on the other ports, into `out`
template <class Ctx>
GT_FN void gt_site_rebuild(const Ctx &c, const GtLayout &L, int s, int chi, double *ws,
                           double *out) {
    const int k = s ? L.k2 : L.k1, p = s ? L.p2 : L.p1, d = s ? L.d2 : L.d1;
    int dims[GT_KMAX];
    for (int j = 0; j < k; ++j) dims[j] = (s ? L.dims2 : L.dims1)[j];
    dims[p] = chi;
    const long long P = s ? L.P2 : L.P1;
    const int r = (int)gt_min(P, s ? L.n2 : L.n1);
    const long long size = P * d * chi;
    const double *Q = ws + L.Q[s];
    const int m1 = L.m1, m2 = L.m2;
    const bool tall = m1 >= m2;
    const double *J = ws + L.thp, *TV = ws + L.tv, *sig = ws + L.sig, *perm = ws + L.perm;
    double *cur = ws + L.bufA[s], *nxt = ws + L.bufB[s];
    // F[(i, t), c] = the kept singular vector entry times sqrt(S_c): site 1's left vectors
    // U[:, c] (tall: Jacobi column / sigma, wide: V column), site 2's right vectors;
    // F and Q in LDS when they fit
    const long long nf = (long long)r * d * chi;
    const bool staged = nf + P * r <= c.lds_cap;
    double *F = staged ? c.lds : nxt;
    for (long long idx = c.tid; idx < nf; idx += c.nt) {
        const int cc = (int)(idx % chi);
        const long long row = idx / chi;  // i * d + t
        const int pc = (int)perm[cc];
        const double rs = ws[L.raw + pc];
        const double inv = rs > 0.0 ? 1.0 / rs : 0.0;  // raw sigma of the Jacobi column
        double f;
        if (s == 0)
            f = tall ? J[(long long)pc * m1 + row] * inv : TV[(long long)pc * m1 + row];
        else
            f = tall ? TV[(long long)pc * m2 + row] : J[(long long)pc * m2 + row] * inv;
        F[idx] = f * sqrt(sig[cc]);  // sig holds the kept, sorted, scaled values
    }
    const double *Qs = Q;
    if (staged) {
        double *lq = c.lds + nf;
        for (long long idx = c.tid; idx < P * r; idx += c.nt) lq[idx] = Q[idx];
        Qs = lq;
    }
    c.sync();
    for (long long idx = c.tid; idx < size; idx += c.nt) {
        long long rem = idx, x = 0, xs = 1;
        int cc = 0;
        for (int j = k - 1; j >= 0; --j) {
            const int a = (int)(rem % dims[j]);
            rem /= dims[j];
            if (j == p) {
                cc = a;
            } else {
                x += a * xs;
                xs *= dims[j];
            }
        }
        const int t = (int)rem;
        double acc = 0.0;
#pragma unroll 4
        for (int i = 0; i < r; ++i) acc += Qs[x * r + i] * F[((long long)i * d + t) * chi + cc];
        cur[idx] = acc;
    }
    c.sync();
    // inverse gauges on the other ports
    const double *src = cur;
    int left = 0;
    for (int j = 0; j < k; ++j)
        if (j != p) ++left;
    for (int j = 0; j < k; ++j) {
        if (j == p) continue;
        --left;
        double *dst = left == 0 ? out : nxt;
        gt_mode(c, src, dst, size, dims, k, j, ws + L.msg[s][j] + 3 * dims[j] * dims[j]);
        c.sync();
        src = dst;
        double *t = cur;
        cur = nxt;
        nxt = t;
    }
    if (left == 0 && src != out) {  // no other ports: copy
        for (long long idx = c.tid; idx < size; idx += c.nt) out[idx] = src[idx];
        c.sync();
    }
}

// The whole gate.  `in`: the gate's packed inputs; `out`: its packed outputs (new site
// tensors at capacity, singular values); ws: its workspace; chi_out: the kept rank.
template <class Ctx>
GT_FN void gt_run(const Ctx &c, const GtLayout &L, const double *in, double *out, double *ws,
                  int *chi_out, double rtol) {
    if (L.kind == 1) {
        // packed: the tensor, the messages (normalize only), then the operator
        const double *A = in, *M = in + L.size1;
        long long mtot = 0;
        for (int j = 0; j < L.nmsg1; ++j) mtot += (long long)L.dims1[j] * L.dims1[j];
        const double *O = M + mtot;
        double *res = out;
        const long long rest = L.size1 / L.d1;
        for (long long idx = c.tid; idx < L.size1; idx += c.nt) {
            const long long t = idx / rest, x = idx % rest;
            double acc = 0.0;
            for (int s = 0; s < L.d1; ++s) acc += O[t * L.d1 + s] * A[s * rest + x];
            res[idx] = acc;
        }
        c.sync();
        if (L.norm) {
            // port j's message follows those of ports 0..j-1
            const double *Ms[GT_KMAX];
            double *Es[GT_KMAX];
            int Ds[GT_KMAX];
            long long mo = 0;
            for (int j = 0; j < L.nmsg1; ++j) {
                Ms[j] = M + mo;
                Es[j] = ws + L.msg[0][j];
                Ds[j] = L.dims1[j];
                mo += (long long)L.dims1[j] * L.dims1[j];
            }
            gt_gauges(c, L.nmsg1, Ms, Es, Ds, ws + L.rot);
        }
        if (L.norm) {
            // gauge every port of res, then the norm
            double *cur = ws + L.bufA[0], *nxt = ws + L.bufB[0];
            const double *src = res;
            for (int j = 0; j < L.k1; ++j) {
                gt_mode(c, src, nxt, L.size1, L.dims1, L.k1, j,
                        ws + L.msg[0][j] + 2 * L.dims1[j] * L.dims1[j]);
                c.sync();
                src = nxt;
                double *t = cur;
                cur = nxt;
                nxt = t;
            }
            double s2 = 0.0;
            for (long long idx = c.tid; idx < L.size1; idx += c.nt) s2 += src[idx] * src[idx];
            const double nrm = sqrt(gt_block_sum(c, s2));
            const double inv = nrm > 0.0 ? 1.0 / nrm : 0.0;
            for (long long idx = c.tid; idx < L.size1; idx += c.nt) res[idx] *= inv;
            c.sync();
        }
        if (c.tid == 0) *chi_out = 0;
        return;
    }
    c.mark(0);
    // inputs
    const double *A1 = in, *A2 = in + L.size1;
    const double *mp = A2 + L.size2;
    const double *mptr[2][GT_KMAX];
    for (int s = 0; s < 2; ++s) {
        const int k = s ? L.k2 : L.k1, p = s ? L.p2 : L.p1;
        const int *dims = s ? L.dims2 : L.dims1;
        for (int j = 0; j < k; ++j) {
            mptr[s][j] = nullptr;
            if (j == p) continue;
            mptr[s][j] = mp;
            mp += (long long)dims[j] * dims[j];
        }
    }
    const double *G = mp;
    // 1. gauges of every message of the gate
    {
        const double *Ms[2 * GT_KMAX];
        double *Es[2 * GT_KMAX];
        int Ds[2 * GT_KMAX], nm = 0;
        for (int s = 0; s < 2; ++s) {
            const int k = s ? L.k2 : L.k1, p = s ? L.p2 : L.p1;
            const int *dims = s ? L.dims2 : L.dims1;
            for (int j = 0; j < k; ++j) {
                if (j == p) continue;
                Ms[nm] = mptr[s][j];
                Es[nm] = ws + L.msg[s][j];
                Ds[nm] = dims[j];
                ++nm;
            }
        }
        gt_gauges(c, nm, Ms, Es, Ds, ws + L.rot);
    }
    c.mark(1);
    // 2-3. gauged matrices and their factorisations
    gt_site_qr(c, L, 0, A1, ws);
    c.mark(2);
    gt_site_qr(c, L, 1, A2, ws);
    c.mark(3);
    // 4. Theta' in the Jacobi layout
    const int d1 = L.d1, d2 = L.d2, Db = L.Db, r1 = L.r1, r2 = L.r2, m1 = L.m1, m2 = L.m2;
    const double *R1 = ws + L.R[0], *R2 = ws + L.R[1];
    double *TH = ws + L.th, *J = ws + L.thp;
    for (long long idx = c.tid; idx < (long long)m1 * m2; idx += c.nt) {
        const int row = (int)(idx / m2), col = (int)(idx % m2);
        const int i1 = row / d1, s1 = row % d1, i2 = col / d2, s2 = col % d2;
        double acc = 0.0;
        for (int b = 0; b < Db; ++b)
            acc += R1[(long long)i1 * L.n1 + s1 * Db + b] * R2[(long long)i2 * L.n2 + s2 * Db + b];
        TH[idx] = acc;
    }
    c.sync();
    const bool tall = m1 >= m2;
    for (long long idx = c.tid; idx < (long long)m1 * m2; idx += c.nt) {
        const int row = (int)(idx / m2), col = (int)(idx % m2);
        const int i1 = row / d1, t1 = row % d1, i2 = col / d2, t2 = col % d2;
        double acc = 0.0;
        for (int s1 = 0; s1 < d1; ++s1)
            for (int s2 = 0; s2 < d2; ++s2)
                acc += G[((long long)(t1 * d2 + t2) * d1 + s1) * d2 + s2] *
                       TH[(long long)(i1 * d1 + s1) * m2 + i2 * d2 + s2];
        if (tall)
            J[(long long)col * m1 + row] = acc;
        else
            J[(long long)row * m2 + col] = acc;
    }
    c.sync();
    c.mark(4);
    // 5. SVD, sort, truncate, normalise
    if (tall)
        gt_jacobi_staged(c, J, m1, m2, ws + L.tv);
    else
        gt_jacobi_staged(c, J, m2, m1, ws + L.tv);
    c.mark(5);
    const int q = L.q, rows = tall ? m1 : m2;
    double *sig = ws + L.sig, *perm = ws + L.perm, *raw = ws + L.raw;
    for (int i = c.tid; i < q; i += c.nt) {
        double s2 = 0.0;
        for (int x = 0; x < rows; ++x) s2 += J[(long long)i * rows + x] * J[(long long)i * rows + x];
        raw[i] = sqrt(s2);  // sigma of Jacobi column i
    }
    c.sync();
    // rank of every Jacobi column by sigma, descending (ties by index): perm[rank] = column
    for (int i = c.tid; i < q; i += c.nt) {
        int rank = 0;
        for (int j = 0; j < q; ++j) rank += raw[j] > raw[i] || (raw[j] == raw[i] && j < i);
        perm[rank] = (double)i;
    }
    c.sync();
    if (c.tid == 0) {
        const double top = raw[(int)perm[0]];
        int chi = L.chimax;
        if (rtol >= 0.0)
            while (chi > 1 && raw[(int)perm[chi - 1]] <= rtol * top) --chi;
        double nrm = 1.0;
        if (L.norm) {
            double s2 = 0.0;
            for (int i = 0; i < chi; ++i) s2 += raw[(int)perm[i]] * raw[(int)perm[i]];
            nrm = s2 > 0.0 ? sqrt(s2) : 1.0;
        }
        *chi_out = chi;
        c.red[0] = (double)chi;
        c.red[1] = nrm;
    }
    c.sync();
    {
        const int chi = (int)c.red[0];
        const double nrm = c.red[1];
        for (int i = c.tid; i < q; i += c.nt)  // kept, sorted, scaled
            sig[i] = i < chi ? raw[(int)perm[i]] / nrm : 0.0;
    }
    c.sync();
    const int chi = (int)c.red[0];
    c.sync();
    c.mark(6);
    // 6. new site tensors and the singular values
    gt_site_rebuild(c, L, 0, chi, ws, out);
    c.mark(7);
    gt_site_rebuild(c, L, 1, chi, ws, out + L.out1);
    c.mark(8);
    for (int i = c.tid; i < chi; i += c.nt) out[L.out1 + L.out2 + i] = sig[i];
    c.sync();
    (void)r1;
    (void)r2;
}

// one workgroup per gate: desc (GD_N ints per gate), offs (GO_N per gate), chi (per gate)
// LDS per workgroup: lds_cap doubles of dynamic LDS for the Jacobi sweeps
__global__ __launch_bounds__(GT_NT) void k_gate(const int *__restrict__ desc,
                                                const long long *__restrict__ offs,
                                                const double *__restrict__ in,
                                                double *__restrict__ out,
                                                double *__restrict__ ws, int *__restrict__ chi,
                                                double rtol, long long lds_cap) {
    __shared__ GtLayout L;
    __shared__ double red[GT_NT];
    __shared__ int flag;
    extern __shared__ double gt_lds[];
    const int g = blockIdx.x;
    if (threadIdx.x == 0) gt_layout(desc + (long long)g * GD_N, L);
    __syncthreads();
    GtDev c{(int)threadIdx.x, GT_NT, red, &flag, gt_lds, lds_cap};
    const long long *o = offs + (long long)g * GO_N;
    gt_run(c, L, in + o[GO_IN], out + o[GO_OUT], ws + o[GO_WS], chi + g, rtol);
}

// ---------------------------------------------------------- gates on a resident network

// Packs gate inputs from the engine's resident arrays: segment e = (source, source offset,
// destination offset, length), source 0 the FP64 sites, 1 the messages, 2 the gate matrices.
template <typename TM>
__global__ void k_gate_gather(const long long *__restrict__ seg, int nseg,
                              const double *__restrict__ sites, const TM *__restrict__ msgs,
                              const double *__restrict__ gates, double *__restrict__ dst) {
    for (int e = blockIdx.x; e < nseg; e += gridDim.x) {
        const long long *t = seg + 4LL * e;
        const int src = (int)t[0];
        const long long so = t[1], dof = t[2], len = t[3];
        for (long long i = threadIdx.x; i < len; i += blockDim.x)
            dst[dof + i] = src == 0 ? sites[so + i] : src == 1 ? (double)msgs[so + i] : gates[so + i];
    }
}

// per-gate scatter record (long long)
enum {
    GS_KIND = 0, GS_OUT, GS_CAP1, GS_CAP2, // packed output offset, the sites' output capacities
    GS_SITE1, GS_SITE2,                    // resident site offsets
    GS_D1, GS_D2, GS_K1, GS_K2, GS_P1, GS_P2, GS_DE,
    GS_MSG1, GS_MSG2,                      // resident offsets of the bond's two messages
    GS_DIMS1, GS_DIMS2 = GS_DIMS1 + GT_KMAX, GS_N = GS_DIMS2 + GT_KMAX
};

// Writes a gate's results back into the resident arrays, one workgroup per gate: the new
// site tensors with the bond zero-padded from chi to its fixed dimension De, diag(S) padded
// into both messages of the bond (the working messages mw, and m64 when not null), and S
// into sout (stride s_stride, zero-padded).
template <typename TM>
__global__ void k_gate_scatter(const long long *__restrict__ rec, const int *__restrict__ chis,
                               const double *__restrict__ out, double *__restrict__ sites,
                               TM *__restrict__ mw, double *__restrict__ m64,
                               double *__restrict__ sout, int s_stride) {
    const long long *r = rec + (long long)blockIdx.x * GS_N;
    const int kind = (int)r[GS_KIND];
    const double *o = out + r[GS_OUT];
    if (kind == 1) {
        long long size = r[GS_D1];
        for (int j = 0; j < (int)r[GS_K1]; ++j) size *= r[GS_DIMS1 + j];
        for (long long i = threadIdx.x; i < size; i += blockDim.x) sites[r[GS_SITE1] + i] = o[i];
        return;
    }
    const int chi = chis[blockIdx.x], De = (int)r[GS_DE];
    for (int s = 0; s < 2; ++s) {
        const int k = (int)r[s ? GS_K2 : GS_K1], p = (int)r[s ? GS_P2 : GS_P1];
        const long long d = r[s ? GS_D2 : GS_D1];
        const long long *dims = r + (s ? GS_DIMS2 : GS_DIMS1);
        long long inner = 1, size = d;  // resident: port p has De; compact: chi
        for (int j = 0; j < k; ++j) size *= dims[j];
        for (int j = p + 1; j < k; ++j) inner *= dims[j];
        const double *src = o + (s ? r[GS_CAP1] : 0);
        double *dst = sites + r[s ? GS_SITE2 : GS_SITE1];
        for (long long i = threadIdx.x; i < size; i += blockDim.x) {
            const long long in_ = i % inner, b = (i / inner) % De, outer = i / (inner * De);
            dst[i] = b < chi ? src[(outer * chi + b) * inner + in_] : 0.0;
        }
    }
    const double *S = o + r[GS_CAP1] + r[GS_CAP2];
    for (long long i = threadIdx.x; i < (long long)De * De; i += blockDim.x) {
        const int a = (int)(i / De), b = (int)(i % De);
        const double v = a == b && a < chi ? S[a] : 0.0;
        mw[r[GS_MSG1] + i] = (TM)v;
        mw[r[GS_MSG2] + i] = (TM)v;
        if (m64) {
            m64[r[GS_MSG1] + i] = v;
            m64[r[GS_MSG2] + i] = v;
        }
    }
    if (sout)
        for (int i = threadIdx.x; i < s_stride; i += blockDim.x)
            sout[(long long)blockIdx.x * s_stride + i] = i < chi ? S[i] : 0.0;
}
