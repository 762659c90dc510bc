/*
 * bp_oracle.c -- FP64 CPU restatement of the reference belief-propagation path.
 *
 * TEST INFRASTRUCTURE ONLY (reference values).  Loaded by the test suite, the smoke check
 * and the benchmark's CPU-baseline leg; never by the product library.
 *
 * Reference sources restated here (ITensorNetworksNext.jl, paths relative to its root):
 *   src/beliefpropagation/beliefpropagation.jl:27-55    stopping criteria
 *   src/beliefpropagation/beliefpropagation.jl:200-210  sweep step (in-place updates)
 *   src/beliefpropagation/beliefpropagation.jl:242-257  SimpleMessageUpdate
 *   src/beliefpropagation/beliefpropagation.jl:261-267  iterate_diff
 *   src/beliefpropagation/messagecache.jl:124-131       incoming_messages
 *   src/beliefpropagation/messagecache.jl:139-201       vertex/edge scalars, bethe_free_energy
 *   src/normnetwork.jl:49-54                             vertex factor A (x) conj(A')
 *   src/contract_network.jl:32-37                        exact contraction
 *   src/AlgorithmsInterfaceExtensions/AlgorithmsInterfaceExtensions.jl:84-109  StopWhenConverged
 * The adjoint and energy functions have no reference counterpart (the reference has no
 * AD); they restate the implicit-function theorem for the same fixed point.
 */
#include "bp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXR 16 /* max tensor rank (phys + ports) */

/* ------------------------------------------------------------------ helpers */

static int vdeg(const orc_graph *g, int u) { return g->port_off[u + 1] - g->port_off[u]; }

static int port_owner(const orc_graph *g, int q) {
    /* binary search over port_off */
    int lo = 0, hi = g->nv - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) / 2;
        if (g->port_off[mid] <= q) lo = mid; else hi = mid - 1;
    }
    return lo;
}

/* dims of the site tensor of u: [phys, D_port0, ..., D_port(k-1)] */
static int site_dims(const orc_graph *g, int u, int *dims) {
    int k = vdeg(g, u);
    dims[0] = g->phys_dim[u];
    for (int j = 0; j < k; ++j) dims[1 + j] = g->edge_dim[g->port_edge[g->port_off[u] + j]];
    return k + 1;
}

static long long prod_range(const int *dims, int a, int b) {
    long long p = 1;
    for (int i = a; i < b; ++i) p *= dims[i];
    return p;
}

/* out[pre, b, post] = sum_a T[pre, a, post] * M[a, b]  (trans: M[b, a]); M is na x nb.
 * dims_out[ax] becomes nb. */
static void mode_product(const double *T, int r, const int *dims, int ax, const double *M,
                         int nb, int trans, double *out) {
    long long pre = prod_range(dims, 0, ax), post = prod_range(dims, ax + 1, r);
    int na = dims[ax];
    for (long long p = 0; p < pre; ++p) {
        const double *tin = T + p * na * post;
        double *tout = out + p * (long long)nb * post;
        memset(tout, 0, sizeof(double) * nb * post);
        for (int a = 0; a < na; ++a) {
            const double *row = tin + (long long)a * post;
            for (int b = 0; b < nb; ++b) {
                double m = trans ? M[(long long)b * na + a] : M[(long long)a * nb + b];
                if (m == 0.0) continue;
                double *o = tout + (long long)b * post;
                for (long long t = 0; t < post; ++t) o[t] += m * row[t];
            }
        }
    }
}

/* out[a, a'] = sum_{pre,post} T[pre, a, post] * B[pre, a', post] on axis ax */
static void open_contract(const double *T, const double *B, int r, const int *dims, int ax,
                          double *out) {
    long long pre = prod_range(dims, 0, ax), post = prod_range(dims, ax + 1, r);
    int n = dims[ax];
    memset(out, 0, sizeof(double) * n * n);
    for (long long p = 0; p < pre; ++p) {
        for (int a = 0; a < n; ++a) {
            const double *ta = T + (p * n + a) * post;
            for (int b = 0; b < n; ++b) {
                const double *tb = B + (p * n + b) * post;
                double s = 0.0;
                for (long long t = 0; t < post; ++t) s += ta[t] * tb[t];
                out[a * n + b] += s;
            }
        }
    }
}

static double dot(const double *a, const double *b, long long n) {
    double s = 0.0;
    for (long long i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static long long site_size(const orc_graph *g, int u) {
    return g->site_off[u + 1] - g->site_off[u];
}

static long long max_site_size(const orc_graph *g) {
    long long m = 1;
    for (int u = 0; u < g->nv; ++u)
        if (site_size(g, u) > m) m = site_size(g, u);
    return m;
}

/* Neumaier compensated sum */
typedef struct { double s, c; } ksum;
static void kadd(ksum *k, double x) {
    double t = k->s + x;
    if (fabs(k->s) >= fabs(x)) k->c += (k->s - t) + x; else k->c += (x - t) + k->s;
    k->s = t;
}
static double kval(const ksum *k) { return k->s + k->c; }

/* --------------------------------------------------- generic vertex contraction */

/* out[a,a'] = sum Aket[s,..a..,X] Abra[s,..a'..,X'] prod_{j != open} mats_j[x_j, x'_j].
 * work: 2 buffers of site size. */
static void vertex_open_w(const orc_graph *g, int u, const double *Aket, const double *Abra,
                          const double *const *mats, int open, double *out, double *w0,
                          double *w1) {
    int dims[MAXR];
    int r = site_dims(g, u, dims);
    int k = r - 1;
    long long n = prod_range(dims, 0, r);
    const double *cur = Aket;
    double *bufs[2] = {w0, w1};
    int which = 0;
    for (int j = 0; j < k; ++j) {
        if (j == open) continue;
        mode_product(cur, r, dims, 1 + j, mats[j], dims[1 + j], 0, bufs[which]);
        cur = bufs[which];
        which ^= 1;
    }
    if (open < 0) {
        out[0] = dot(cur, Abra, n);
    } else {
        open_contract(cur, Abra, r, dims, 1 + open, out);
    }
}

int orc_vertex_open(const orc_graph *g, int u, const double *Aket, const double *Abra,
                    const double *const *mats, int open, double *out) {
    long long n = site_size(g, u);
    double *w = (double *)malloc(sizeof(double) * 2 * n);
    if (!w) return 1;
    vertex_open_w(g, u, Aket, Abra, mats, open, out, w, w + n);
    free(w);
    return 0;
}

/* env[s, X] = sum_{X'} Abra[s, X'] prod_j mats_j[x_j, x'_j]  (bra side absorbs M^T) */
static void vertex_env_w(const orc_graph *g, int u, const double *Abra,
                         const double *const *mats, double *env, double *w0) {
    int dims[MAXR];
    int r = site_dims(g, u, dims);
    int k = r - 1;
    long long n = prod_range(dims, 0, r);
    if (k == 0) { memcpy(env, Abra, sizeof(double) * n); return; }
    /* ping-pong so that the last product lands in env */
    const double *cur = Abra;
    double *dst;
    for (int j = 0; j < k; ++j) {
        int remaining = k - 1 - j;
        dst = (remaining % 2 == 0) ? env : w0;
        mode_product(cur, r, dims, 1 + j, mats[j], dims[1 + j], 1, dst);
        cur = dst;
    }
}

int orc_vertex_env(const orc_graph *g, int u, const double *Abra, const double *const *mats,
                   double *env) {
    long long n = site_size(g, u);
    double *w = (double *)malloc(sizeof(double) * n);
    if (!w) return 1;
    vertex_env_w(g, u, Abra, mats, env, w);
    free(w);
    return 0;
}

/* ------------------------------------------------------------- message updates */

static void norm_message_w(const orc_graph *g, const double *sites, const double *msgs, int q,
                           double *out, double *w0, double *w1) {
    int u = port_owner(g, q);
    int base = g->port_off[u], k = vdeg(g, u);
    const double *mats[MAXR];
    for (int j = 0; j < k; ++j) mats[j] = msgs + g->msg_off[g->port_rev[base + j]];
    const double *A = sites + g->site_off[u];
    vertex_open_w(g, u, A, A, mats, q - base, out, w0, w1);
}

int orc_norm_message(const orc_graph *g, const double *sites, const double *msgs, int q,
                     double *out) {
    long long n = max_site_size(g);
    double *w = (double *)malloc(sizeof(double) * 2 * n);
    if (!w) return 1;
    norm_message_w(g, sites, msgs, q, out, w, w + n);
    free(w);
    return 0;
}

/* plain network: out[a] = sum A[a (port p), X] prod_{j != p} m_j[x_j]  (phys dim 1) */
static void plain_message_w(const orc_graph *g, const double *sites, const double *msgs, int q,
                            double *out, double *w0, double *w1) {
    int u = port_owner(g, q);
    int base = g->port_off[u], k = vdeg(g, u), p = q - base;
    int dims[MAXR];
    int r = site_dims(g, u, dims);
    const double *cur = sites + g->site_off[u];
    double *bufs[2] = {w0, w1};
    int which = 0;
    for (int j = 0; j < k; ++j) {
        if (j == p) continue;
        const double *m = msgs + g->msg_off[g->port_rev[base + j]];
        mode_product(cur, r, dims, 1 + j, m, 1, 0, bufs[which]);
        dims[1 + j] = 1;
        cur = bufs[which];
        which ^= 1;
    }
    long long n = prod_range(dims, 0, r);
    memcpy(out, cur, sizeof(double) * n); /* remaining dims: phys(=1) x D_p */
}

int orc_plain_message(const orc_graph *g, const double *sites, const double *msgs, int q,
                      double *out) {
    long long n = max_site_size(g);
    double *w = (double *)malloc(sizeof(double) * 2 * n);
    if (!w) return 1;
    plain_message_w(g, sites, msgs, q, out, w, w + n);
    free(w);
    return 0;
}

static long long msg_len(const orc_graph *g, int q) { return g->msg_off[q + 1] - g->msg_off[q]; }

/* sum normalisation as beliefpropagation.jl:248-253 */
static void normalize_sum(double *m, long long n) {
    double s = 0.0;
    for (long long i = 0; i < n; ++i) s += m[i];
    if (s != 0.0)
        for (long long i = 0; i < n; ++i) m[i] /= s;
}

static double fid_dist(const double *a, const double *b, long long n) {
    double na = sqrt(dot(a, a, n)), nb = sqrt(dot(b, b, n));
    double d = dot(a, b, n) / (na * nb);
    return 1.0 - d * d;
}

static double maxabs_diff(const double *a, const double *b, long long n) {
    double m = 0.0;
    for (long long i = 0; i < n; ++i) {
        double d = fabs(a[i] - b[i]);
        if (d > m || d != d) m = d;
    }
    return m;
}

int orc_sweep_jacobi(const orc_graph *g, const double *sites, const double *old_msgs,
                     double *new_msgs, int plain, int normalize, int nthreads,
                     double *res_fid, double *res_abs) {
    int nports = g->port_off[g->nv];
    long long wn = max_site_size(g);
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = 1;
#pragma omp parallel num_threads(nthreads)
#endif
    {
        double *w = (double *)malloc(sizeof(double) * 2 * wn);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 4)
#endif
        for (int q = 0; q < nports; ++q) {
            double *out = new_msgs + g->msg_off[q];
            if (plain) plain_message_w(g, sites, old_msgs, q, out, w, w + wn);
            else norm_message_w(g, sites, old_msgs, q, out, w, w + wn);
            long long n = msg_len(g, q);
            if (normalize) normalize_sum(out, n);
            const double *o = old_msgs + g->msg_off[q];
            if (res_fid) res_fid[q] = fid_dist(out, o, n);
            if (res_abs) res_abs[q] = maxabs_diff(out, o, n);
        }
        free(w);
    }
    return 0;
}

int orc_sweep_sequential(const orc_graph *g, const double *sites, double *msgs, int plain,
                         int normalize, const int *order, int norder) {
    long long wn = max_site_size(g);
    double *w = (double *)malloc(sizeof(double) * 2 * wn);
    double *tmp = (double *)malloc(sizeof(double) * wn);
    if (!w || !tmp) { free(w); free(tmp); return 1; }
    for (int i = 0; i < norder; ++i) {
        int q = order[i];
        if (plain) plain_message_w(g, sites, msgs, q, tmp, w, w + wn);
        else norm_message_w(g, sites, msgs, q, tmp, w, w + wn);
        long long n = msg_len(g, q);
        if (normalize) normalize_sum(tmp, n);
        memcpy(msgs + g->msg_off[q], tmp, sizeof(double) * n);
    }
    free(w);
    free(tmp);
    return 0;
}

double orc_iterate_diff(const orc_graph *g, const double *m1, const double *m2, int plain) {
    (void)plain;
    int nports = g->port_off[g->nv];
    double mx = -INFINITY;
    for (int q = 0; q < nports; ++q) {
        double d = fid_dist(m1 + g->msg_off[q], m2 + g->msg_off[q], msg_len(g, q));
        if (d > mx || d != d) mx = d;
    }
    return mx;
}

static double maxabs_all(const orc_graph *g, const double *m1, const double *m2) {
    int nports = g->port_off[g->nv];
    return maxabs_diff(m1, m2, g->msg_off[nports]);
}

int orc_bp(const orc_graph *g, const double *sites, double *msgs, int plain, int schedule,
           const int *order, int norder, int maxiter, double tol, int criterion,
           int nthreads, int *out_sweeps, double *out_delta, int *out_at_iteration) {
    int nports = g->port_off[g->nv];
    long long total = g->msg_off[nports];
    double *prev = (double *)malloc(sizeof(double) * total);
    double *tmp = (double *)malloc(sizeof(double) * total);
    if (!prev || !tmp) { free(prev); free(tmp); return 1; }
    memcpy(prev, msgs, sizeof(double) * total);
    int iteration = 0, at = -1;
    double delta = INFINITY;
    for (;;) {
        /* StopWhenAny: every criterion's is_finished! runs (StopWhenConverged copies the
         * previous iterate each time; its delta is skipped at iteration 0). */
        int stop = 0;
        if (maxiter > 0 && iteration >= maxiter) stop = 1;
        if (tol >= 0.0) {
            double d = criterion == 0 ? orc_iterate_diff(g, msgs, prev, plain)
                                      : maxabs_all(g, msgs, prev);
            memcpy(prev, msgs, sizeof(double) * total);
            if (iteration > 0) {
                delta = d;
                if (d < tol) { at = iteration; stop = 1; }
            }
        }
        if (stop) break;
        iteration += 1;
        if (schedule == 0) {
            orc_sweep_jacobi(g, sites, msgs, tmp, plain, 1, nthreads, NULL, NULL);
            memcpy(msgs, tmp, sizeof(double) * total);
        } else {
            orc_sweep_sequential(g, sites, msgs, plain, 1, order, norder);
        }
        if (maxiter <= 0 && tol < 0.0) break;
    }
    if (out_sweeps) *out_sweeps = iteration;
    if (out_delta) *out_delta = delta;
    if (out_at_iteration) *out_at_iteration = at;
    free(prev);
    free(tmp);
    return 0;
}

/* ------------------------------------------------------------------- scalars */

int orc_vertex_scalars(const orc_graph *g, const double *sites, const double *msgs, int plain,
                       double *zv) {
    long long wn = max_site_size(g);
    double *w = (double *)malloc(sizeof(double) * 2 * wn);
    if (!w) return 1;
    for (int u = 0; u < g->nv; ++u) {
        int base = g->port_off[u], k = vdeg(g, u);
        const double *A = sites + g->site_off[u];
        if (plain) {
            int dims[MAXR];
            int r = site_dims(g, u, dims);
            const double *cur = A;
            double *bufs[2] = {w, w + wn};
            int which = 0;
            for (int j = 0; j < k; ++j) {
                const double *m = msgs + g->msg_off[g->port_rev[base + j]];
                mode_product(cur, r, dims, 1 + j, m, 1, 0, bufs[which]);
                dims[1 + j] = 1;
                cur = bufs[which];
                which ^= 1;
            }
            double s = 0.0;
            for (long long i = 0; i < prod_range(dims, 0, r); ++i) s += cur[i];
            zv[u] = s;
        } else {
            const double *mats[MAXR];
            for (int j = 0; j < k; ++j) mats[j] = msgs + g->msg_off[g->port_rev[base + j]];
            vertex_open_w(g, u, A, A, mats, -1, zv + u, w, w + wn);
        }
    }
    free(w);
    return 0;
}

int orc_edge_scalars(const orc_graph *g, const double *msgs, int plain, double *ze) {
    (void)plain;
    int nports = g->port_off[g->nv];
    for (int e = 0; e < g->ne; ++e) ze[e] = NAN;
    for (int q = 0; q < nports; ++q) {
        int e = g->port_edge[q];
        int r = g->port_rev[q];
        if (q > r) continue; /* each undirected edge once */
        ze[e] = dot(msgs + g->msg_off[q], msgs + g->msg_off[r], msg_len(g, q));
    }
    return 0;
}

int orc_bethe_free_energy(const orc_graph *g, const double *sites, const double *msgs,
                          int plain, double *out_re, double *out_im) {
    double *zv = (double *)malloc(sizeof(double) * (g->nv + 1));
    double *ze = (double *)malloc(sizeof(double) * (g->ne + 1));
    if (!zv || !ze) { free(zv); free(ze); return 1; }
    orc_vertex_scalars(g, sites, msgs, plain, zv);
    orc_edge_scalars(g, msgs, plain, ze);
    int neg_v = 0, neg_e = 0, zero_e = 0;
    for (int u = 0; u < g->nv; ++u) if (zv[u] < 0) ++neg_v;
    for (int e = 0; e < g->ne; ++e) { if (ze[e] < 0) ++neg_e; if (ze[e] == 0) zero_e = 1; }
    if (zero_e) {
        *out_re = -INFINITY;
        *out_im = 0.0;
        free(zv); free(ze);
        return 0;
    }
    ksum kr = {0, 0};
    for (int u = 0; u < g->nv; ++u) kadd(&kr, log(fabs(zv[u])));
    for (int e = 0; e < g->ne; ++e) kadd(&kr, -log(fabs(ze[e])));
    /* complex promotion: log(x) = log|x| + i*pi for x < 0 */
    *out_re = kval(&kr);
    *out_im = M_PI * (double)(neg_v - neg_e);
    for (int u = 0; u < g->nv; ++u) if (zv[u] == 0) *out_re = -INFINITY;
    free(zv);
    free(ze);
    return 0;
}

/* --------------------------------------------------- labelled exact contraction */

int orc_contract_labeled(int nt, const int *ranks, const int *lab_off, const int *labels,
                         const int *label_dim, const double *data, const long long *data_off,
                         long long max_elems, int *out_rank, int *out_labels, double *out,
                         long long out_cap) {
    /* frontier */
    int flab[64];
    int fr = 0;
    double *F = (double *)malloc(sizeof(double));
    if (!F) return 1;
    F[0] = 1.0;
    for (int t = 0; t < nt; ++t) {
        int tr = ranks[t];
        const int *tl = labels + lab_off[t];
        const double *T = data + data_off[t];
        /* classify */
        int sh_f[64], sh_t[64], ns = 0;  /* positions of shared labels */
        int fo[64], nfo = 0, to[64], nto = 0;
        for (int i = 0; i < fr; ++i) {
            int found = -1;
            for (int j = 0; j < tr; ++j) if (tl[j] == flab[i]) found = j;
            if (found >= 0) { sh_f[ns] = i; sh_t[ns] = found; ++ns; } else fo[nfo++] = i;
        }
        for (int j = 0; j < tr; ++j) {
            int shared = 0;
            for (int i = 0; i < ns; ++i) if (sh_t[i] == j) shared = 1;
            if (!shared) to[nto++] = j;
        }
        if (nfo + nto > 64) { free(F); return 2; }
        /* strides */
        long long fstr[64], tstr[64];
        { long long s = 1; for (int i = fr - 1; i >= 0; --i) { fstr[i] = s; s *= label_dim[flab[i]]; } }
        { long long s = 1; for (int j = tr - 1; j >= 0; --j) { tstr[j] = s; s *= label_dim[tl[j]]; } }
        long long nFO = 1, nTO = 1, nSH = 1;
        for (int i = 0; i < nfo; ++i) nFO *= label_dim[flab[fo[i]]];
        for (int i = 0; i < nto; ++i) nTO *= label_dim[tl[to[i]]];
        for (int i = 0; i < ns; ++i) nSH *= label_dim[flab[sh_f[i]]];
        if (nFO * nTO > max_elems) { free(F); return 3; }
        /* offset tables */
        long long *offFO = (long long *)malloc(sizeof(long long) * nFO);
        long long *offTO = (long long *)malloc(sizeof(long long) * nTO);
        long long *offFS = (long long *)malloc(sizeof(long long) * nSH);
        long long *offTS = (long long *)malloc(sizeof(long long) * nSH);
        double *NF = (double *)malloc(sizeof(double) * (nFO * nTO > 0 ? nFO * nTO : 1));
        if (!offFO || !offTO || !offFS || !offTS || !NF) {
            free(offFO); free(offTO); free(offFS); free(offTS); free(NF); free(F);
            return 1;
        }
        for (long long x = 0; x < nFO; ++x) {
            long long rem = x, off = 0;
            for (int i = nfo - 1; i >= 0; --i) {
                int d = label_dim[flab[fo[i]]];
                off += (rem % d) * fstr[fo[i]];
                rem /= d;
            }
            offFO[x] = off;
        }
        for (long long x = 0; x < nTO; ++x) {
            long long rem = x, off = 0;
            for (int i = nto - 1; i >= 0; --i) {
                int d = label_dim[tl[to[i]]];
                off += (rem % d) * tstr[to[i]];
                rem /= d;
            }
            offTO[x] = off;
        }
        for (long long x = 0; x < nSH; ++x) {
            long long rem = x, of = 0, ot = 0;
            for (int i = ns - 1; i >= 0; --i) {
                int d = label_dim[flab[sh_f[i]]];
                long long c = rem % d;
                rem /= d;
                of += c * fstr[sh_f[i]];
                ot += c * tstr[sh_t[i]];
            }
            offFS[x] = of;
            offTS[x] = ot;
        }
        for (long long a = 0; a < nFO; ++a) {
            for (long long b = 0; b < nTO; ++b) {
                double s = 0.0;
                for (long long c = 0; c < nSH; ++c) s += F[offFO[a] + offFS[c]] * T[offTS[c] + offTO[b]];
                NF[a * nTO + b] = s;
            }
        }
        int nl[64], nr = 0;
        for (int i = 0; i < nfo; ++i) nl[nr++] = flab[fo[i]];
        for (int i = 0; i < nto; ++i) nl[nr++] = tl[to[i]];
        memcpy(flab, nl, sizeof(int) * nr);
        fr = nr;
        free(F);
        F = NF;
        free(offFO); free(offTO); free(offFS); free(offTS);
    }
    long long n = 1;
    for (int i = 0; i < fr; ++i) n *= label_dim[flab[i]];
    if (n > out_cap) { free(F); return 4; }
    memcpy(out, F, sizeof(double) * n);
    *out_rank = fr;
    for (int i = 0; i < fr; ++i) out_labels[i] = flab[i];
    free(F);
    return 0;
}

/* -------------------------------------------------------------------- adjoint */

static void transpose_sq(const double *m, int n, double *t) {
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) t[j * n + i] = m[i * n + j];
}

/* d<gbar, env-contraction>/dA for real A used on both layers:
 *   abar += env(Abra=A, mats) + env(A, mats^T)   (ket and bra derivatives) */
static void accum_site_grad(const orc_graph *g, int u, const double *Aket, const double *Abra,
                            const double *const *mats, double scale_ket, double scale_bra,
                            const double *phys_op_ket, double *abar, double *w0, double *w1,
                            double *mt_store) {
    int dims[MAXR];
    int r = site_dims(g, u, dims);
    int k = r - 1;
    long long n = prod_range(dims, 0, r);
    int d = dims[0];
    long long rest = n / d;
    /* ket derivative: d/dAket sum Aket * env(Abra, mats) = env(Abra, mats); if the ket is
     * op . A, chain rule gives op^T . env */
    if (scale_ket != 0.0) {
        vertex_env_w(g, u, Abra, mats, w1, w0);
        if (phys_op_ket) {
            for (int s = 0; s < d; ++s)
                for (long long x = 0; x < rest; ++x) {
                    double acc = 0.0;
                    for (int t = 0; t < d; ++t) acc += phys_op_ket[t * d + s] * w1[t * rest + x];
                    abar[s * rest + x] += scale_ket * acc;
                }
        } else {
            for (long long i = 0; i < n; ++i) abar[i] += scale_ket * w1[i];
        }
    }
    /* bra derivative: d/dAbra = env(Aket, mats^T) */
    if (scale_bra != 0.0) {
        const double *mt[MAXR];
        long long off = 0;
        for (int j = 0; j < k; ++j) {
            int D = dims[1 + j];
            transpose_sq(mats[j], D, mt_store + off);
            mt[j] = mt_store + off;
            off += (long long)D * D;
        }
        vertex_env_w(g, u, Aket, mt, w1, w0);
        for (long long i = 0; i < n; ++i) abar[i] += scale_bra * w1[i];
    }
}

static long long max_mat_store(const orc_graph *g) {
    long long m = 0;
    for (int u = 0; u < g->nv; ++u) {
        long long s = 0;
        for (int j = g->port_off[u]; j < g->port_off[u + 1]; ++j) {
            long long D = g->edge_dim[g->port_edge[j]];
            s += D * D;
        }
        if (s > m) m = s;
    }
    return m > 0 ? m : 1;
}

/* Threads used by orc_norm_vjp (and so by orc_fixed_point_gradient); 1 by default. */
static int g_vjp_threads = 1;

void orc_set_num_threads(int n) { g_vjp_threads = n > 0 ? n : 1; }

static long long max_msg_len(const orc_graph *g) {
    long long m = 1;
    int nports = g->port_off[g->nv];
    for (int q = 0; q < nports; ++q)
        if (msg_len(g, q) > m) m = msg_len(g, q);
    return m;
}

/* The vertex-u share of orc_norm_vjp.  Every slot written into mbar is an in-slot of u
 * (port_rev of a port of u), consumed by u alone, and abar is written at u's site only,
 * so different vertices never write the same entry. */
static void norm_vjp_vertex(const orc_graph *g, const double *sites, const double *msgs,
                            const double *gbar, double *mbar, double *abar, int u, double *w0,
                            double *w1, double *w2, double *mt, double *gg, double *tmp) {
    int base = g->port_off[u], k = vdeg(g, u);
    const double *A = sites + g->site_off[u];
    for (int p = 0; p < k; ++p) {
        int q = base + p;
        long long n = msg_len(g, q);
        /* recompute unnormalised g_q and its sum */
        const double *mats[MAXR];
        for (int j = 0; j < k; ++j) mats[j] = msgs + g->msg_off[g->port_rev[base + j]];
        vertex_open_w(g, u, A, A, mats, p, tmp, w0, w1);
        double S = 0.0;
        for (long long i = 0; i < n; ++i) S += tmp[i];
        const double *fb = gbar + g->msg_off[q];
        /* through normalisation f = g / S:  gbar_g = (fbar - <fbar, f>) / S */
        if (S != 0.0) {
            double fd = 0.0;
            for (long long i = 0; i < n; ++i) fd += fb[i] * tmp[i] / S;
            for (long long i = 0; i < n; ++i) gg[i] = (fb[i] - fd) / S;
        } else {
            memcpy(gg, fb, sizeof(double) * n);
        }
        /* port p carries gbar_g, the others carry incoming messages */
        mats[p] = gg;
        if (mbar) {
            for (int j = 0; j < k; ++j) {
                if (j == p) continue;
                vertex_open_w(g, u, A, A, mats, j, tmp, w0, w1);
                double *mb = mbar + g->msg_off[g->port_rev[base + j]];
                long long nj = msg_len(g, g->port_rev[base + j]);
                for (long long i = 0; i < nj; ++i) mb[i] += tmp[i];
            }
        }
        if (abar) {
            accum_site_grad(g, u, A, A, mats, 1.0, 1.0, NULL, abar + g->site_off[u], w0, w2, mt);
        }
    }
}

int orc_norm_vjp(const orc_graph *g, const double *sites, const double *msgs,
                 const double *gbar, double *mbar, double *abar) {
    long long wn = max_site_size(g);
    long long ms = max_mat_store(g);
    long long ml = max_msg_len(g);
    int fail = 0;
#pragma omp parallel num_threads(g_vjp_threads) reduction(| : fail)
    {
        double *w0 = (double *)malloc(sizeof(double) * wn);
        double *w1 = (double *)malloc(sizeof(double) * wn);
        double *w2 = (double *)malloc(sizeof(double) * wn);
        double *mt = (double *)malloc(sizeof(double) * ms);
        double *gg = (double *)malloc(sizeof(double) * ml);
        double *tmp = (double *)malloc(sizeof(double) * ml);
        if (!w0 || !w1 || !w2 || !mt || !gg || !tmp) {
            fail = 1;
        } else {
#pragma omp for schedule(dynamic, 4)
            for (int u = 0; u < g->nv; ++u)
                norm_vjp_vertex(g, sites, msgs, gbar, mbar, abar, u, w0, w1, w2, mt, gg, tmp);
        }
        free(w0); free(w1); free(w2); free(mt); free(gg); free(tmp);
    }
    return fail;
}

int orc_grad_lognorm_explicit(const orc_graph *g, const double *sites, const double *msgs,
                              double *grad) {
    long long wn = max_site_size(g);
    long long ms = max_mat_store(g);
    double *w0 = (double *)malloc(sizeof(double) * wn);
    double *w1 = (double *)malloc(sizeof(double) * wn);
    double *w2 = (double *)malloc(sizeof(double) * wn);
    double *mt = (double *)malloc(sizeof(double) * ms);
    if (!w0 || !w1 || !w2 || !mt) { free(w0); free(w1); free(w2); free(mt); return 1; }
    memset(grad, 0, sizeof(double) * g->site_off[g->nv]);
    for (int u = 0; u < g->nv; ++u) {
        int base = g->port_off[u], k = vdeg(g, u);
        const double *A = sites + g->site_off[u];
        const double *mats[MAXR];
        for (int j = 0; j < k; ++j) mats[j] = msgs + g->msg_off[g->port_rev[base + j]];
        double Z;
        vertex_open_w(g, u, A, A, mats, -1, &Z, w0, w1);
        accum_site_grad(g, u, A, A, mats, 1.0 / Z, 1.0 / Z, NULL, grad + g->site_off[u], w0, w2,
                        mt);
    }
    free(w0); free(w1); free(w2); free(mt);
    return 0;
}

/* apply a d x d operator on the physical index: out[s', X] = sum_s op[s', s] A[s, X] */
static void apply_phys(const double *op, const double *A, int d, long long rest, double *out) {
    for (int sp = 0; sp < d; ++sp)
        for (long long x = 0; x < rest; ++x) {
            double acc = 0.0;
            for (int s = 0; s < d; ++s) acc += op[sp * d + s] * A[s * rest + x];
            out[sp * rest + x] = acc;
        }
}

int orc_tfim_energy(const orc_graph *g, const double *sites, const double *msgs, double J,
                    double h, double *E, double *dE_dm, double *dE_dA) {
    static const double X2[4] = {0.0, 1.0, 1.0, 0.0};
    static const double Z2[4] = {1.0, 0.0, 0.0, -1.0};
    long long wn = max_site_size(g);
    long long ms = max_mat_store(g);
    int nports = g->port_off[g->nv];
    double *w0 = (double *)malloc(sizeof(double) * wn);
    double *w1 = (double *)malloc(sizeof(double) * wn);
    double *w2 = (double *)malloc(sizeof(double) * wn);
    double *opA = (double *)malloc(sizeof(double) * wn);
    double *opB = (double *)malloc(sizeof(double) * wn);
    double *mt = (double *)malloc(sizeof(double) * ms);
    double *gu = (double *)malloc(sizeof(double) * 4096);
    double *gv = (double *)malloc(sizeof(double) * 4096);
    double *gzu = (double *)malloc(sizeof(double) * 4096);
    double *gzv = (double *)malloc(sizeof(double) * 4096);
    double *tmp = (double *)malloc(sizeof(double) * 4096);
    if (!w0 || !w1 || !w2 || !opA || !opB || !mt || !gu || !gv || !gzu || !gzv || !tmp) return 1;
    if (dE_dm) memset(dE_dm, 0, sizeof(double) * g->msg_off[nports]);
    if (dE_dA) memset(dE_dA, 0, sizeof(double) * g->site_off[g->nv]);
    ksum acc = {0, 0};
    /* vertex terms: -h <X_v> = -h Zx_v / Z_v */
    for (int u = 0; u < g->nv; ++u) {
        int base = g->port_off[u], k = vdeg(g, u);
        const double *A = sites + g->site_off[u];
        int d = g->phys_dim[u];
        if (d != 2) return 5;
        long long rest = site_size(g, u) / d;
        const double *mats[MAXR];
        for (int j = 0; j < k; ++j) mats[j] = msgs + g->msg_off[g->port_rev[base + j]];
        apply_phys(X2, A, d, rest, opA);
        double Z, Zx;
        vertex_open_w(g, u, A, A, mats, -1, &Z, w0, w1);
        vertex_open_w(g, u, opA, A, mats, -1, &Zx, w0, w1);
        double ex = Zx / Z;
        kadd(&acc, -h * ex);
        /* d(-h Zx/Z) = -h (dZx - ex dZ) / Z */
        double c = -h / Z;
        if (dE_dm) {
            for (int j = 0; j < k; ++j) {
                int qin = g->port_rev[base + j];
                long long n = msg_len(g, qin);
                double *dm = dE_dm + g->msg_off[qin];
                vertex_open_w(g, u, opA, A, mats, j, tmp, w0, w1);
                for (long long i = 0; i < n; ++i) dm[i] += c * tmp[i];
                vertex_open_w(g, u, A, A, mats, j, tmp, w0, w1);
                for (long long i = 0; i < n; ++i) dm[i] -= c * ex * tmp[i];
            }
        }
        if (dE_dA) {
            double *da = dE_dA + g->site_off[u];
            /* Zx = <X.A, A>_M : ket derivative X^T env(A, M), bra derivative env(X.A, M^T) */
            accum_site_grad(g, u, opA, A, mats, c, 0.0, X2, da, w0, w2, mt);
            accum_site_grad(g, u, opA, A, mats, 0.0, c, NULL, da, w0, w2, mt);
            accum_site_grad(g, u, A, A, mats, -c * ex, -c * ex, NULL, da, w0, w2, mt);
        }
    }
    /* edge terms: -J <Z_u Z_v> = -J <gz_uv, gz_vu> / <g_uv, g_vu> */
    for (int q = 0; q < nports; ++q) {
        int r = g->port_rev[q];
        if (q > r) continue;
        int u = port_owner(g, q), v = port_owner(g, r);
        int bu = g->port_off[u], ku = vdeg(g, u), pu = q - bu;
        int bv = g->port_off[v], kv = vdeg(g, v), pv = r - bv;
        const double *Au = sites + g->site_off[u], *Av = sites + g->site_off[v];
        long long ru = site_size(g, u) / 2, rv = site_size(g, v) / 2;
        const double *mu[MAXR], *mv[MAXR];
        for (int j = 0; j < ku; ++j) mu[j] = msgs + g->msg_off[g->port_rev[bu + j]];
        for (int j = 0; j < kv; ++j) mv[j] = msgs + g->msg_off[g->port_rev[bv + j]];
        apply_phys(Z2, Au, 2, ru, opA);
        apply_phys(Z2, Av, 2, rv, opB);
        vertex_open_w(g, u, Au, Au, mu, pu, gu, w0, w1);
        vertex_open_w(g, v, Av, Av, mv, pv, gv, w0, w1);
        vertex_open_w(g, u, opA, Au, mu, pu, gzu, w0, w1);
        vertex_open_w(g, v, opB, Av, mv, pv, gzv, w0, w1);
        long long n = msg_len(g, q);
        double Dn = dot(gu, gv, n), Nn = dot(gzu, gzv, n);
        double ezz = Nn / Dn;
        kadd(&acc, -J * ezz);
        double c = -J / Dn;
        /* dN: port pu of u carries gzv; dD: port pu carries gv (and symmetric at v) */
        if (dE_dm || dE_dA) {
            const double *mu2[MAXR], *mv2[MAXR];
            for (int side = 0; side < 2; ++side) {
                int w = side == 0 ? u : v;
                int bw = side == 0 ? bu : bv, kw = side == 0 ? ku : kv, pw = side == 0 ? pu : pv;
                const double *Aw = side == 0 ? Au : Av;
                const double *opw = side == 0 ? opA : opB;
                const double **mw2 = side == 0 ? mu2 : mv2;
                const double *const *mw = side == 0 ? mu : mv;
                const double *gN = side == 0 ? gzv : gzu; /* other side's inserted env */
                const double *gD = side == 0 ? gv : gu;
                for (int j = 0; j < kw; ++j) mw2[j] = mw[j];
                if (dE_dm) {
                    for (int j = 0; j < kw; ++j) {
                        if (j == pw) continue;
                        int qin = g->port_rev[bw + j];
                        long long nj = msg_len(g, qin);
                        double *dm = dE_dm + g->msg_off[qin];
                        mw2[pw] = gN;
                        vertex_open_w(g, w, opw, Aw, mw2, j, tmp, w0, w1);
                        for (long long i = 0; i < nj; ++i) dm[i] += c * tmp[i];
                        mw2[pw] = gD;
                        vertex_open_w(g, w, Aw, Aw, mw2, j, tmp, w0, w1);
                        for (long long i = 0; i < nj; ++i) dm[i] -= c * ezz * tmp[i];
                    }
                }
                if (dE_dA) {
                    double *da = dE_dA + g->site_off[w];
                    mw2[pw] = gN;
                    accum_site_grad(g, w, opw, Aw, mw2, c, 0.0, Z2, da, w0, w2, mt);
                    accum_site_grad(g, w, opw, Aw, mw2, 0.0, c, NULL, da, w0, w2, mt);
                    mw2[pw] = gD;
                    accum_site_grad(g, w, Aw, Aw, mw2, -c * ezz, -c * ezz, NULL, da, w0, w2, mt);
                }
            }
        }
    }
    *E = kval(&acc);
    free(w0); free(w1); free(w2); free(opA); free(opB); free(mt);
    free(gu); free(gv); free(gzu); free(gzv); free(tmp);
    return 0;
}

/* top-left d x d block of an ld x ld operator */
static void op_block(const double *op, int ld, int d, double *out) {
    for (int a = 0; a < d; ++a)
        for (int b = 0; b < d; ++b) out[a * d + b] = op[a * ld + b];
}

int orc_expect_vertex(const orc_graph *g, const double *sites, const double *msgs,
                      const double *ops, int ld, double *out) {
    long long wn = max_site_size(g);
    double *w0 = (double *)malloc(sizeof(double) * wn);
    double *w1 = (double *)malloc(sizeof(double) * wn);
    double *opA = (double *)malloc(sizeof(double) * wn);
    double *ob = (double *)malloc(sizeof(double) * ld * ld);
    if (!w0 || !w1 || !opA || !ob) return 1;
    for (int u = 0; u < g->nv; ++u) {
        int base = g->port_off[u], k = vdeg(g, u), d = g->phys_dim[u];
        if (d > ld) return 5;
        const double *A = sites + g->site_off[u];
        const double *mats[MAXR];
        for (int j = 0; j < k; ++j) mats[j] = msgs + g->msg_off[g->port_rev[base + j]];
        op_block(ops + (long long)u * ld * ld, ld, d, ob);
        apply_phys(ob, A, d, site_size(g, u) / d, opA);
        double Z, Zo;
        vertex_open_w(g, u, A, A, mats, -1, &Z, w0, w1);
        vertex_open_w(g, u, opA, A, mats, -1, &Zo, w0, w1);
        out[u] = Zo / Z;
    }
    free(w0); free(w1); free(opA); free(ob);
    return 0;
}

int orc_expect_edge(const orc_graph *g, const double *sites, const double *msgs,
                    const double *ops_first, const double *ops_second, int ld, double *out) {
    long long wn = max_site_size(g);
    int nports = g->port_off[g->nv];
    double *w0 = (double *)malloc(sizeof(double) * wn);
    double *w1 = (double *)malloc(sizeof(double) * wn);
    double *opA = (double *)malloc(sizeof(double) * wn);
    double *opB = (double *)malloc(sizeof(double) * wn);
    double *ob = (double *)malloc(sizeof(double) * ld * ld);
    long long ml = 1;
    for (int q = 0; q < nports; ++q) ml = msg_len(g, q) > ml ? msg_len(g, q) : ml;
    double *gu = (double *)malloc(sizeof(double) * ml), *gv = (double *)malloc(sizeof(double) * ml);
    double *hu = (double *)malloc(sizeof(double) * ml), *hv = (double *)malloc(sizeof(double) * ml);
    if (!w0 || !w1 || !opA || !opB || !ob || !gu || !gv || !hu || !hv) return 1;
    for (int q = 0; q < nports; ++q) {
        int r = g->port_rev[q];
        if (q > r) continue;
        int e = g->port_edge[q];
        int u = port_owner(g, q), v = port_owner(g, r);
        int bu = g->port_off[u], ku = vdeg(g, u), bv = g->port_off[v], kv = vdeg(g, v);
        int du = g->phys_dim[u], dv = g->phys_dim[v];
        if (du > ld || dv > ld) return 5;
        const double *Au = sites + g->site_off[u], *Av = sites + g->site_off[v];
        const double *mu[MAXR], *mv[MAXR];
        for (int j = 0; j < ku; ++j) mu[j] = msgs + g->msg_off[g->port_rev[bu + j]];
        for (int j = 0; j < kv; ++j) mv[j] = msgs + g->msg_off[g->port_rev[bv + j]];
        op_block(ops_first + (long long)e * ld * ld, ld, du, ob);
        apply_phys(ob, Au, du, site_size(g, u) / du, opA);
        op_block(ops_second + (long long)e * ld * ld, ld, dv, ob);
        apply_phys(ob, Av, dv, site_size(g, v) / dv, opB);
        vertex_open_w(g, u, Au, Au, mu, q - bu, gu, w0, w1);
        vertex_open_w(g, v, Av, Av, mv, r - bv, gv, w0, w1);
        vertex_open_w(g, u, opA, Au, mu, q - bu, hu, w0, w1);
        vertex_open_w(g, v, opB, Av, mv, r - bv, hv, w0, w1);
        long long n = msg_len(g, q);
        out[e] = dot(hu, hv, n) / dot(gu, gv, n);
    }
    free(w0); free(w1); free(opA); free(opB); free(ob); free(gu); free(gv); free(hu); free(hv);
    return 0;
}

int orc_local_energy(const orc_graph *g, const double *sites, const double *msgs,
                     const double *ops_v, const double *ops_first, const double *ops_second,
                     int ld, double *E, double *dE_dm, double *dE_dA) {
    long long wn = max_site_size(g);
    long long ms = max_mat_store(g);
    int nports = g->port_off[g->nv];
    long long ml = 1;
    for (int q = 0; q < nports; ++q) ml = msg_len(g, q) > ml ? msg_len(g, q) : ml;
    double *w0 = (double *)malloc(sizeof(double) * wn);
    double *w1 = (double *)malloc(sizeof(double) * wn);
    double *w2 = (double *)malloc(sizeof(double) * wn);
    double *opA = (double *)malloc(sizeof(double) * wn);
    double *opB = (double *)malloc(sizeof(double) * wn);
    double *mt = (double *)malloc(sizeof(double) * ms);
    double *oa = (double *)malloc(sizeof(double) * ld * ld);
    double *ob = (double *)malloc(sizeof(double) * ld * ld);
    double *gu = (double *)malloc(sizeof(double) * ml), *gv = (double *)malloc(sizeof(double) * ml);
    double *hu = (double *)malloc(sizeof(double) * ml), *hv = (double *)malloc(sizeof(double) * ml);
    double *tmp = (double *)malloc(sizeof(double) * ml);
    if (!w0 || !w1 || !w2 || !opA || !opB || !mt || !oa || !ob || !gu || !gv || !hu || !hv || !tmp)
        return 1;
    if (dE_dm) memset(dE_dm, 0, sizeof(double) * g->msg_off[nports]);
    if (dE_dA) memset(dE_dA, 0, sizeof(double) * g->site_off[g->nv]);
    ksum acc = {0, 0};
    /* vertex terms <O_v> = Zo / Z */
    for (int u = 0; ops_v && u < g->nv; ++u) {
        int base = g->port_off[u], k = vdeg(g, u), d = g->phys_dim[u];
        if (d > ld) return 5;
        const double *A = sites + g->site_off[u];
        const double *mats[MAXR];
        for (int j = 0; j < k; ++j) mats[j] = msgs + g->msg_off[g->port_rev[base + j]];
        op_block(ops_v + (long long)u * ld * ld, ld, d, oa);
        apply_phys(oa, A, d, site_size(g, u) / d, opA);
        double Z, Zo;
        vertex_open_w(g, u, A, A, mats, -1, &Z, w0, w1);
        vertex_open_w(g, u, opA, A, mats, -1, &Zo, w0, w1);
        double ex = Zo / Z, c = 1.0 / Z;
        kadd(&acc, ex);
        if (dE_dm) {
            for (int j = 0; j < k; ++j) {
                int qin = g->port_rev[base + j];
                long long n = msg_len(g, qin);
                double *dm = dE_dm + g->msg_off[qin];
                vertex_open_w(g, u, opA, A, mats, j, tmp, w0, w1);
                for (long long i = 0; i < n; ++i) dm[i] += c * tmp[i];
                vertex_open_w(g, u, A, A, mats, j, tmp, w0, w1);
                for (long long i = 0; i < n; ++i) dm[i] -= c * ex * tmp[i];
            }
        }
        if (dE_dA) {
            double *da = dE_dA + g->site_off[u];
            accum_site_grad(g, u, opA, A, mats, c, 0.0, oa, da, w0, w2, mt);
            accum_site_grad(g, u, opA, A, mats, 0.0, c, NULL, da, w0, w2, mt);
            accum_site_grad(g, u, A, A, mats, -c * ex, -c * ex, NULL, da, w0, w2, mt);
        }
    }
    /* edge terms <A_e B_e> = <h_uv, h_vu> / <g_uv, g_vu> */
    for (int q = 0; ops_first && ops_second && q < nports; ++q) {
        int r = g->port_rev[q];
        if (q > r) continue;
        int e = g->port_edge[q];
        int u = port_owner(g, q), v = port_owner(g, r);
        int bu = g->port_off[u], ku = vdeg(g, u), pu = q - bu;
        int bv = g->port_off[v], kv = vdeg(g, v), pv = r - bv;
        int du = g->phys_dim[u], dv = g->phys_dim[v];
        if (du > ld || dv > ld) return 5;
        const double *Au = sites + g->site_off[u], *Av = sites + g->site_off[v];
        const double *mu[MAXR], *mv[MAXR];
        for (int j = 0; j < ku; ++j) mu[j] = msgs + g->msg_off[g->port_rev[bu + j]];
        for (int j = 0; j < kv; ++j) mv[j] = msgs + g->msg_off[g->port_rev[bv + j]];
        op_block(ops_first + (long long)e * ld * ld, ld, du, oa);
        op_block(ops_second + (long long)e * ld * ld, ld, dv, ob);
        apply_phys(oa, Au, du, site_size(g, u) / du, opA);
        apply_phys(ob, Av, dv, site_size(g, v) / dv, opB);
        vertex_open_w(g, u, Au, Au, mu, pu, gu, w0, w1);
        vertex_open_w(g, v, Av, Av, mv, pv, gv, w0, w1);
        vertex_open_w(g, u, opA, Au, mu, pu, hu, w0, w1);
        vertex_open_w(g, v, opB, Av, mv, pv, hv, w0, w1);
        long long n = msg_len(g, q);
        double Dn = dot(gu, gv, n), Nn = dot(hu, hv, n);
        double ee = Nn / Dn, c = 1.0 / Dn;
        kadd(&acc, ee);
        if (!dE_dm && !dE_dA) continue;
        const double *mw2[MAXR];
        for (int side = 0; side < 2; ++side) {
            int w = side == 0 ? u : v;
            int bw = side == 0 ? bu : bv, kw = side == 0 ? ku : kv, pw = side == 0 ? pu : pv;
            const double *Aw = side == 0 ? Au : Av;
            const double *opw = side == 0 ? opA : opB;
            const double *om = side == 0 ? oa : ob;
            const double *const *mw = side == 0 ? mu : mv;
            const double *gN = side == 0 ? hv : hu; /* other side's inserted env */
            const double *gD = side == 0 ? gv : gu;
            for (int j = 0; j < kw; ++j) mw2[j] = mw[j];
            if (dE_dm) {
                for (int j = 0; j < kw; ++j) {
                    if (j == pw) continue;
                    int qin = g->port_rev[bw + j];
                    long long nj = msg_len(g, qin);
                    double *dm = dE_dm + g->msg_off[qin];
                    mw2[pw] = gN;
                    vertex_open_w(g, w, opw, Aw, mw2, j, tmp, w0, w1);
                    for (long long i = 0; i < nj; ++i) dm[i] += c * tmp[i];
                    mw2[pw] = gD;
                    vertex_open_w(g, w, Aw, Aw, mw2, j, tmp, w0, w1);
                    for (long long i = 0; i < nj; ++i) dm[i] -= c * ee * tmp[i];
                }
            }
            if (dE_dA) {
                double *da = dE_dA + g->site_off[w];
                mw2[pw] = gN;
                accum_site_grad(g, w, opw, Aw, mw2, c, 0.0, om, da, w0, w2, mt);
                accum_site_grad(g, w, opw, Aw, mw2, 0.0, c, NULL, da, w0, w2, mt);
                mw2[pw] = gD;
                accum_site_grad(g, w, Aw, Aw, mw2, -c * ee, -c * ee, NULL, da, w0, w2, mt);
            }
        }
    }
    *E = kval(&acc);
    free(w0); free(w1); free(w2); free(opA); free(opB); free(mt); free(oa); free(ob);
    free(gu); free(gv); free(hu); free(hv); free(tmp);
    return 0;
}

int orc_fixed_point_gradient(const orc_graph *g, const double *sites, const double *msgs,
                             const double *dL_dm, const double *dL_dA, int maxiter, double tol,
                             double *grad, int *out_iters, double *out_res) {
    int nports = g->port_off[g->nv];
    long long nm = g->msg_off[nports], ns = g->site_off[g->nv];
    double *lam = (double *)calloc(nm, sizeof(double));
    double *nxt = (double *)calloc(nm, sizeof(double));
    if (!lam || !nxt) { free(lam); free(nxt); return 1; }
    if (dL_dm) memcpy(lam, dL_dm, sizeof(double) * nm);
    int it = 0;
    double res = INFINITY;
    while (it < maxiter) {
        if (dL_dm) memcpy(nxt, dL_dm, sizeof(double) * nm); else memset(nxt, 0, sizeof(double) * nm);
        orc_norm_vjp(g, sites, msgs, lam, nxt, NULL);
        {
            double sc = 0.0;
            for (long long i = 0; i < nm; ++i) sc = fmax(sc, fabs(nxt[i]));
            res = maxabs_diff(nxt, lam, nm);
            if (sc > 0.0) res /= sc; /* relative change of the adjoint iterate */
        }
        double *t = lam; lam = nxt; nxt = t;
        ++it;
        if (res < tol) break;
    }
    if (dL_dA) memcpy(grad, dL_dA, sizeof(double) * ns); else memset(grad, 0, sizeof(double) * ns);
    orc_norm_vjp(g, sites, msgs, lam, NULL, grad);
    if (out_iters) *out_iters = it;
    if (out_res) *out_res = res;
    free(lam);
    free(nxt);
    return 0;
}
