/*
 * bp_oracle.h -- FP64 CPU restatement of the reference belief-propagation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library gives the reference values for the HIP engine: only
 * the test suite, the smoke check and the CPU-baseline leg of the benchmark load
 * it.  The product library never links or calls it.
 *
 * Each function names the reference source it restates (paths relative to the
 * ITensorNetworksNext.jl source tree).
 *
 * Graph format (shared with the product C-ABI, include/bp_mi355x.h):
 *   - vertices 0..nv-1; each vertex u owns ports port_off[u]..port_off[u+1]-1;
 *   - port q of u points to neighbour port_nbr[q] over undirected edge port_edge[q];
 *     port_rev[q] is the neighbour's port pointing back at u;
 *   - the directed message u -> port_nbr[q] is stored at message slot q, so the
 *     message arriving at u through port q is slot port_rev[q];
 *   - site tensor of u is row-major [phys_dim[u]][D_port0]...[D_port(k-1)], ports in
 *     CSR order, at site_off[u];
 *   - norm-network messages are D_e x D_e row-major [ket][bra] at msg_off[q];
 *     plain-network messages are length-D_e vectors at msg_off[q].
 */
#ifndef BP_ORACLE_H
#define BP_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nv;
    int ne;
    const int *port_off;       /* nv + 1 */
    const int *port_nbr;       /* nports */
    const int *port_edge;      /* nports */
    const int *port_rev;       /* nports */
    const int *edge_dim;       /* ne */
    const int *phys_dim;       /* nv (1 for plain networks) */
    const long long *site_off; /* nv + 1 */
    const long long *msg_off;  /* nports + 1 */
} orc_graph;

/* ---- single message (norm network, double layer A (x) conj(A)) ----
 * messagecache.jl:124-131 (incoming_messages) + beliefpropagation.jl:242-257
 * (SimpleMessageUpdate): out = contract(A_u, conj A_u, m_{w->u} for w != v), unnormalised. */
int orc_norm_message(const orc_graph *g, const double *sites, const double *msgs, int q,
                     double *out);

/* ---- single message (plain network, vector messages) ---- same reference lines. */
int orc_plain_message(const orc_graph *g, const double *sites, const double *msgs, int q,
                      double *out);

/* ---- one Jacobi sweep: new[q] = normalise(update(old)) for every slot q.
 * beliefpropagation.jl:248-253 normalisation (skipped when the sum is zero).
 * res_fid[q] = 1 - |<new_hat, old_hat>|^2 (beliefpropagation.jl:261-267),
 * res_abs[q] = max |new - old|; either may be NULL.  plain != 0 selects vector messages. */
int orc_sweep_jacobi(const orc_graph *g, const double *sites, const double *old_msgs,
                     double *new_msgs, int plain, int normalize, int nthreads,
                     double *res_fid, double *res_abs);

/* ---- one sequential (in-place, Gauss-Seidel) sweep over `order` (message slots).
 * beliefpropagation.jl:200-210: each update is written into the cache immediately. */
int orc_sweep_sequential(const orc_graph *g, const double *sites, double *msgs, int plain,
                         int normalize, const int *order, int norder);

/* ---- the full solve with the reference stopping semantics:
 * StopAfterIteration(maxiter) | StopWhenConverged(tol)   (beliefpropagation.jl:27-55,
 * AlgorithmsInterfaceExtensions.jl:84-109: delta skipped at iteration 0).
 * schedule 0 = Jacobi, 1 = sequential over `order`.
 * criterion 0 = fidelity distance (reference iterate_diff), 1 = max-abs change.
 * maxiter <= 0 means "not given"; tol < 0 means "not given".
 * Outputs: sweeps done, last delta, converged-at iteration (-1 if never). */
int orc_bp(const orc_graph *g, const double *sites, double *msgs, int plain, int schedule,
           const int *order, int norder, int maxiter, double tol, int criterion,
           int nthreads, int *out_sweeps, double *out_delta, int *out_at_iteration);

/* iterate_diff (beliefpropagation.jl:261-267): max_q 1 - |<m1_hat, m2_hat>|^2. */
double orc_iterate_diff(const orc_graph *g, const double *m1, const double *m2, int plain);

/* vertex_scalars (messagecache.jl:139-151): Z_v = full contraction of factor v with all
 * incoming messages. */
int orc_vertex_scalars(const orc_graph *g, const double *sites, const double *msgs, int plain,
                       double *zv);

/* edge_scalars (messagecache.jl:153-178): one per undirected edge, Z_e = <m_uv, m_vu>
 * contracted on both shared names; edge order = undirected edge id. */
int orc_edge_scalars(const orc_graph *g, const double *msgs, int plain, double *ze);

/* bethe_free_energy (messagecache.jl:185-201): sum log Z_v - sum log Z_e with the
 * reference's complex promotion (any negative term -> complex log) and -Inf on a zero
 * edge term.  Neumaier-compensated FP64 sums.  Writes real and imaginary parts. */
int orc_bethe_free_energy(const orc_graph *g, const double *sites, const double *msgs,
                          int plain, double *out_re, double *out_im);

/* ---- exact contraction of a labelled tensor list (contract_network.jl:32-37).
 * Tensor t has rank ranks[t], labels labels[lab_off[t] .. lab_off[t]+ranks[t]-1], data at
 * data + data_off[t].  A label present in two tensors is summed; a label present once
 * stays open, ordered by first appearance.  Tensors are absorbed left to right into a
 * frontier tensor.  Returns the output rank in *out_rank, labels in out_labels and the
 * data in out (capacity out_cap elements).  Returns nonzero on error (e.g. frontier
 * larger than max_elems). */
int orc_contract_labeled(int nt, const int *ranks, const int *lab_off, const int *labels,
                         const int *label_dim, const double *data, const long long *data_off,
                         long long max_elems, int *out_rank, int *out_labels, double *out,
                         long long out_cap);

/* ---- generic vertex contraction building block used by the adjoint and the
 * energy estimator: out[a,a'] = sum Aket[s,..a..,X] Abra[s,..a'..,X'] prod_{j != open}
 * mats_j[x_j, x'_j]  (open = -1: full scalar written to out[0]).  mats is an array of
 * k pointers (one D_j x D_j matrix per port). */
int orc_vertex_open(const orc_graph *g, int u, const double *Aket, const double *Abra,
                    const double *const *mats, int open, double *out);

/* env[s, X] = sum_{X'} Abra[s, X'] prod_j mats_j[x_j, x'_j]  (every port absorbed on the
 * bra side), so that sum Aket * env = orc_vertex_open(..., open = -1). */
int orc_vertex_env(const orc_graph *g, int u, const double *Abra, const double *const *mats,
                   double *env);

/* ---- adjoint of the normalised Jacobi map at a fixed point.
 * Given cotangents gbar[q] of the outgoing (normalised) messages, accumulate
 *   mbar[q_in]  += (d F / d m)^T gbar   (into the incoming message slots)
 *   abar[site]  += (d F / d A)^T gbar   (real site tensors, ket and bra both)
 * where F is one Jacobi sweep followed by sum normalisation.  Either output may be NULL. */
/* Threads for orc_norm_vjp and orc_fixed_point_gradient (OpenMP over vertices; default 1). */
void orc_set_num_threads(int n);
int orc_norm_vjp(const orc_graph *g, const double *sites, const double *msgs,
                 const double *gbar, double *mbar, double *abar);

/* ---- gradient of the BP log-norm with respect to every site tensor at the fixed point
 * `msgs` (explicit partial; the message term vanishes by Bethe stationarity, and the
 * adjoint solve below confirms it).  grad has the layout of `sites`. */
int orc_grad_lognorm_explicit(const orc_graph *g, const double *sites, const double *msgs,
                              double *grad);

/* ---- BP-estimated transverse-field Ising energy
 *   E = sum_{<uv>} -J <Z_u Z_v>_BP + sum_v -h <X_v>_BP
 * with <.>_BP the ratio of the operator-inserted region contraction to the plain one
 * (edge region {u,v}: incoming boundary messages; vertex region {v}: incoming messages).
 * Writes E.  If dE_dm / dE_dA are non-NULL, also writes the explicit partials of E
 * with respect to the messages (message layout) and the site tensors (site layout). */
int orc_tfim_energy(const orc_graph *g, const double *sites, const double *msgs, double J,
                    double h, double *E, double *dE_dm, double *dE_dA);

/* ---- BP expectation values with arbitrary physical operators (the operator acts on the
 * ket layer: (O.A)[s', X] = sum_s O[s', s] A[s, X]).  Operators are ld x ld row-major
 * blocks; a vertex of physical dimension d uses the top-left d x d corner.
 *   vertex v:  out[v] = Z_v[O_v] / Z_v  (incoming messages on every port)
 *   edge e = {u, v} (u owns the lower port of e):
 *     out[e] = <g^{A_e}_{u->v}, g^{B_e}_{v->u}> / <g_{u->v}, g_{v->u}>
 *   with g^{O}_{u->v} the unnormalised message of u with O inserted, ops_first[e] = A_e on
 *   u, ops_second[e] = B_e on v.  Real arithmetic; no reference counterpart (the
 *   reference has no expectation API): the ratios restate its region scalars
 *   (messagecache.jl:139-151) with an operator inserted. */
int orc_expect_vertex(const orc_graph *g, const double *sites, const double *msgs,
                      const double *ops, int ld, double *out);
int orc_expect_edge(const orc_graph *g, const double *sites, const double *msgs,
                    const double *ops_first, const double *ops_second, int ld, double *out);

/* ---- BP-estimated local energy with arbitrary operators (the generalisation of the TFIM
 * energy): E = sum_v <O_v>_BP + sum_e <A_e B_e>_BP with the expectation values of
 * orc_expect_vertex / orc_expect_edge (ops_v, or the ops_first / ops_second pair, may be
 * NULL to drop that sum); optional explicit partials as orc_tfim_energy. */
int orc_local_energy(const orc_graph *g, const double *sites, const double *msgs,
                     const double *ops_v, const double *ops_first, const double *ops_second,
                     int ld, double *E, double *dE_dm, double *dE_dA);

/* ---- full implicit-differentiation gradient of an objective L(m*, A) at the fixed
 * point m*: solves lambda = J_m^T lambda + dL/dm (projected onto the sum-normalised
 * tangent space) by adjoint sweeps to tolerance tol on the relative change
 * max|lambda_new - lambda| / max|lambda_new|, then
 * grad = dL/dA + J_A^T lambda.  Writes the gradient (site layout), the sweep count and
 * the final adjoint residual. */
int orc_fixed_point_gradient(const orc_graph *g, const double *sites, const double *msgs,
                             const double *dL_dm, const double *dL_dA, int maxiter, double tol,
                             double *grad, int *out_iters, double *out_res);

#ifdef __cplusplus
}
#endif
#endif
