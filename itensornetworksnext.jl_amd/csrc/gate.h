// gate.h -- BP-gauged gate application: BPApplyGate / apply_gate_bp_nsite!
// (src/apply/apply_operators.jl:181-283), one workgroup per gate, FP64.
//
// Two-site gate on the bond between sites 1 and 2 (Val(2), :229-283):
//  1. every incoming message M on a port other than the shared one is factored
//     M = V diag(lam) V^T (cyclic Jacobi, one thread per message); the gauge is
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
        for (int j = 0; j < L.nmsg1; ++j) {
            const long long D = L.dims1[j];
            L.msg[0][j] = take(4 * D * D + D);
        }
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
    for (int s = 0; s < 2; ++s) {
        const int k = s ? L.k2 : L.k1, p = s ? L.p2 : L.p1;
        const int *dims = s ? L.dims2 : L.dims1;
        for (int j = 0; j < k; ++j) {
            const long long D = dims[j];
            L.msg[s][j] = j == p ? -1 : take(4 * D * D + D);
        }
    }
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

// ------------------------------------------------------------------ block helpers

struct GtDev {
    int tid, nt;
    double *red;   // [nt] reduction scratch
    int *flag;     // one int
    __device__ void sync() const { __syncthreads(); }
    __device__ double group_sum(double v, int gs) const {
        for (int o = gs >> 1; o > 0; o >>= 1) v += __shfl_xor(v, o, gs);
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

// serial cyclic Jacobi: W (D x D, symmetric, row-major) -> V diag(lam) V^T
GT_FN void gt_eigh_serial(double *W, double *V, double *lam, int D) {
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) V[i * D + j] = i == j ? 1.0 : 0.0;
    for (int sweep = 0; sweep < 60; ++sweep) {
        double off = 0.0, dia = 0.0;
        for (int p = 0; p < D; ++p) {
            dia += W[p * D + p] * W[p * D + p];
            for (int q = p + 1; q < D; ++q) off += W[p * D + q] * W[p * D + q];
        }
        if (off <= 1e-30 * (dia + off) || off == 0.0) break;
        for (int p = 0; p < D; ++p)
            for (int q = p + 1; q < D; ++q) {
                const double apq = W[p * D + q];
                if (apq == 0.0) continue;
                const double th = (W[q * D + q] - W[p * D + p]) / (2.0 * apq);
                const double t = (th >= 0.0 ? 1.0 : -1.0) / (fabs(th) + sqrt(th * th + 1.0));
                const double cs = 1.0 / sqrt(t * t + 1.0), sn = t * cs;
                for (int k = 0; k < D; ++k) {
                    const double wp = W[k * D + p], wq = W[k * D + q];
                    W[k * D + p] = cs * wp - sn * wq;
                    W[k * D + q] = sn * wp + cs * wq;
                }
                for (int k = 0; k < D; ++k) {
                    const double wp = W[p * D + k], wq = W[q * D + k];
                    W[p * D + k] = cs * wp - sn * wq;
                    W[q * D + k] = sn * wp + cs * wq;
                }
                for (int k = 0; k < D; ++k) {
                    const double vp = V[k * D + p], vq = V[k * D + q];
                    V[k * D + p] = cs * vp - sn * vq;
                    V[k * D + q] = sn * vp + cs * vq;
                }
            }
    }
    for (int i = 0; i < D; ++i) lam[i] = W[i * D + i];
}

// gauge of one message block at `e` (W, V, X, Xi, lam): M (D x D, [ket][bra]) ->
// X = V sqrt(lam), Xi = pinv(X)
GT_FN void gt_gauge_serial(const double *M, double *e, int D) {
    double *W = e, *V = e + D * D, *X = e + 2 * D * D, *Xi = e + 3 * D * D,
           *lam = e + 4 * D * D;
    for (int i = 0; i < D; ++i)
        for (int j = 0; j < D; ++j) W[i * D + j] = 0.5 * (M[i * D + j] + M[j * D + i]);
    gt_eigh_serial(W, V, lam, D);
    double lmax = 0.0;
    for (int i = 0; i < D; ++i) lmax = lam[i] > lmax ? lam[i] : lmax;
    for (int x = 0; x < D; ++x) {
        const double l = lam[x] > 1e-13 * lmax ? lam[x] : 0.0;
        const double sq = sqrt(l), inv = l > 0.0 ? 1.0 / sq : 0.0;
        for (int a = 0; a < D; ++a) {
            X[a * D + x] = V[a * D + x] * sq;
            Xi[x * D + a] = V[a * D + x] * inv;
        }
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
    for (long long idx = c.tid; idx < size; idx += c.nt) {
        const long long i = idx % inner, b = (idx / inner) % D, o = idx / (inner * D);
        const double *src = in + o * inner * D + i;
        double s = 0.0;
        for (int a = 0; a < D; ++a) s += src[a * inner] * Mx[a * D + b];
        out[idx] = s;
    }
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
    // to the Jacobi layout: tall (P >= n) columns (t, b), wide columns x
    double *J = ws + L.jac[s];
    const bool tall = P >= n;
    for (long long idx = c.tid; idx < size; idx += c.nt) {
        long long rem = idx, x = 0, xs = 1;
        int bb = 0;
        for (int j = k - 1; j >= 0; --j) {
            const int a = (int)(rem % dims[j]);
            rem /= dims[j];
            if (j == p) {
                bb = a;
            } else {
                x += a * xs;
                xs *= dims[j];
            }
        }
        const int t = (int)rem;
        const long long col = (long long)t * Db + bb;
        if (tall)
            J[col * P + x] = src[idx];
        else
            J[x * n + col] = src[idx];
    }
    c.sync();
    double *JV = ws + L.jv[s];
    if (tall)
        gt_jacobi(c, J, (int)P, n, JV);
    else
        gt_jacobi(c, J, n, (int)P, JV);
    // Q and R
    double *Q = ws + L.Q[s], *R = ws + L.R[s];
    if (tall) {
        // columns of J: U S (P x n); V = JV (n x n)
        for (int col = c.tid; col < n; col += c.nt) {
            double s2 = 0.0;
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
    // left factor (site 1) rows (i1, t1): U[:, c] = tall ? J col / sigma : TV col;
    // right factor (site 2) rows (i2, t2): V[:, c] = tall ? TV col : J col / sigma
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
        const int pc = (int)perm[cc];
        const double sq = sqrt(sig[cc]);  // sig holds the kept, sorted, scaled values
        double acc = 0.0;
        const double rs = ws[L.raw + pc];
        const double inv = rs > 0.0 ? 1.0 / rs : 0.0;  // raw sigma of the Jacobi column
        for (int i = 0; i < r; ++i) {
            const long long row = (long long)i * d + t;
            double f;
            if (s == 0)
                f = tall ? J[(long long)pc * m1 + row] * inv : TV[(long long)pc * m1 + row];
            else
                f = tall ? TV[(long long)pc * m2 + row] : J[(long long)pc * m2 + row] * inv;
            acc += Q[x * r + i] * f;
        }
        cur[idx] = acc * sq;
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
            // one thread per port; port j's message follows those of ports 0..j-1
            long long mo = 0;
            for (int j = 0; j < L.nmsg1; ++j) {
                if (j % c.nt == c.tid) gt_gauge_serial(M + mo, ws + L.msg[0][j], L.dims1[j]);
                mo += (long long)L.dims1[j] * L.dims1[j];
            }
            c.sync();
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
    // 1. gauges, one thread per message
    {
        int idx = 0;
        for (int s = 0; s < 2; ++s) {
            const int k = s ? L.k2 : L.k1, p = s ? L.p2 : L.p1;
            const int *dims = s ? L.dims2 : L.dims1;
            for (int j = 0; j < k; ++j) {
                if (j == p) continue;
                if (idx % c.nt == c.tid) gt_gauge_serial(mptr[s][j], ws + L.msg[s][j], dims[j]);
                ++idx;
            }
        }
    }
    c.sync();
    // 2-3. gauged matrices and their factorisations
    gt_site_qr(c, L, 0, A1, ws);
    gt_site_qr(c, L, 1, A2, ws);
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
    // 5. SVD, sort, truncate, normalise
    if (tall)
        gt_jacobi(c, J, m1, m2, ws + L.tv);
    else
        gt_jacobi(c, J, m2, m1, ws + L.tv);
    const int q = L.q, rows = tall ? m1 : m2;
    double *sig = ws + L.sig, *perm = ws + L.perm, *raw = ws + L.raw;
    for (int i = c.tid; i < q; i += c.nt) {
        double s2 = 0.0;
        for (int x = 0; x < rows; ++x) s2 += J[(long long)i * rows + x] * J[(long long)i * rows + x];
        raw[i] = sqrt(s2);  // sigma of Jacobi column i
    }
    c.sync();
    if (c.tid == 0) {
        // sort column indices by sigma, descending (insertion sort; q <= GT_QMAX)
        int idx[GT_QMAX];
        for (int i = 0; i < q; ++i) idx[i] = i;
        for (int i = 1; i < q; ++i) {
            const int v = idx[i];
            int j = i - 1;
            while (j >= 0 && raw[idx[j]] < raw[v]) {
                idx[j + 1] = idx[j];
                --j;
            }
            idx[j + 1] = v;
        }
        int chi = L.chimax;
        if (rtol >= 0.0)
            while (chi > 1 && raw[idx[chi - 1]] <= rtol * raw[idx[0]]) --chi;
        double nrm = 1.0;
        if (L.norm) {
            double s2 = 0.0;
            for (int i = 0; i < chi; ++i) s2 += raw[idx[i]] * raw[idx[i]];
            nrm = s2 > 0.0 ? sqrt(s2) : 1.0;
        }
        for (int i = 0; i < q; ++i) {
            sig[i] = i < chi ? raw[idx[i]] / nrm : 0.0;  // kept, sorted, scaled
            perm[i] = (double)idx[i];
        }
        *chi_out = chi;
        c.red[0] = (double)chi;
    }
    c.sync();
    const int chi = (int)c.red[0];
    c.sync();
    // 6. new site tensors and the singular values
    gt_site_rebuild(c, L, 0, chi, ws, out);
    gt_site_rebuild(c, L, 1, chi, ws, out + L.out1);
    for (int i = c.tid; i < chi; i += c.nt) out[L.out1 + L.out2 + i] = sig[i];
    c.sync();
    (void)r1;
    (void)r2;
}

// one workgroup per gate: desc (GD_N ints per gate), offs (GO_N per gate), chi (per gate)
__global__ __launch_bounds__(GT_NT) void k_gate(const int *__restrict__ desc,
                                                const long long *__restrict__ offs,
                                                const double *__restrict__ in,
                                                double *__restrict__ out,
                                                double *__restrict__ ws, int *__restrict__ chi,
                                                double rtol) {
    __shared__ GtLayout L;
    __shared__ double red[GT_NT];
    __shared__ int flag;
    const int g = blockIdx.x;
    if (threadIdx.x == 0) gt_layout(desc + (long long)g * GD_N, L);
    __syncthreads();
    GtDev c{(int)threadIdx.x, GT_NT, red, &flag};
    const long long *o = offs + (long long)g * GO_N;
    gt_run(c, L, in + o[GO_IN], out + o[GO_OUT], ws + o[GO_WS], chi + g, rtol);
}
