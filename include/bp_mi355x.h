/*
 * bp_mi355x.h -- C ABI of the MI355X belief-propagation engine (libbpmi355x.so).
 *
 * Plain C types only (pointers, sizes, status codes); no HIP or torch types cross the
 * boundary.  Every call returns an int status (BP_OK = 0); bp_last_error() gives the
 * message of the calling thread's last failure.  No C++ exception crosses the ABI.
 *
 * Reference interfaces these entry points replace (ITensorNetworksNext.jl, paths relative
 * to its source root):
 *   bp_sweep / bp_run        AI.solve over BeliefPropagationSweepAlgorithm with
 *                            message_update!(::SimpleMessageUpdate, ...)
 *                            src/beliefpropagation/beliefpropagation.jl:69-92, 200-210, 242-257
 *   bp_update_slot           one message_update!(alg, cache, factors, edge) call
 *                            src/beliefpropagation/beliefpropagation.jl:242-257
 *   bp_iterate_diff          AIE.iterate_diff(::MessageCache, ::MessageCache)
 *                            src/beliefpropagation/beliefpropagation.jl:261-267
 *   bp_vertex_scalars        vertex_scalars src/beliefpropagation/messagecache.jl:139-151
 *   bp_edge_scalars          edge_scalars   src/beliefpropagation/messagecache.jl:153-178
 *   bp_lognorm               bethe_free_energy src/beliefpropagation/messagecache.jl:185-201
 *   bp_init_messages         message_environment(f, nn) src/beliefpropagation/messagecache.jl:205-229
 *   bp_set_topology          MessageCache / NormNetwork graph structure
 *                            src/beliefpropagation/messagecache.jl:15-49, src/normnetwork.jl:13-28
 *   bp_upload_sites          NormNetwork vertex data lazy(A)*lazy(conj(A')) src/normnetwork.jl:49-54
 *   bp_contract_network,     contract_network(tn; alg = Exact(; order, order_alg)) and
 *   bp_cn_*                  contraction_order src/contract_network.jl:6-60
 *   bp_apply_gates           apply_operator / BPApplyGate / apply_gate_bp_nsite!
 *                            src/apply/apply_operators.jl:181-283
 *   bp_apply_gates_resident  the same on the resident network (no marshalling between gate
 *                            layers and BP runs)
 * The adjoint entry points (bp_grad_*) and the energy estimator have no reference
 * counterpart: the reference has no AD and no expectation-value API.
 *
 * Layout (row-major):
 *   - vertex u owns ports port_off[u] .. port_off[u+1]-1; port q points to port_nbr[q]
 *     over undirected edge port_edge[q]; port_rev[q] is the neighbour's port back to u;
 *   - message slot q holds the directed message u -> port_nbr[q] (q a port of u);
 *   - site tensor of u: [phys_dim[u]][D_port0]...[D_port(k-1)], concatenated over u;
 *   - norm-network message: D x D, [ket][bra]; plain-network message: length D.
 *   - a batch is a disjoint union of networks: vertex_net[u] gives u's network; the
 *     vertices (hence ports) of one network must be contiguous.
 * Host buffers are FP64 and copied; the caller keeps ownership.
 *
 * Threading: a context is used by one host thread at a time.  Beyond the HIP runtime's,
 * the library's only global mutable state is the per-device address of its static
 * scratch array (set once per device under a lock, by the first bp_create there) and the
 * array itself (bp_iterate_diff_noalloc).
 */
#ifndef BP_MI355X_H
#define BP_MI355X_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    BP_OK = 0,
    BP_ERR_ARG = 1,         /* invalid argument or shape */
    BP_ERR_HIP = 2,         /* HIP runtime error */
    BP_ERR_STATE = 3,       /* call out of order (e.g. no topology yet) */
    BP_ERR_UNSUPPORTED = 4, /* configuration outside what the kernels cover */
    BP_ERR_NOMEM = 5
};

/* contraction precision of a context */
enum {
    BP_F64 = 0,  /* FP64 storage and arithmetic */
    BP_F32 = 1,  /* FP32 storage and arithmetic (MFMA FP32 on the vertex path) */
    BP_BF16 = 2, /* BF16 MFMA operands, FP32 accumulation and FP32 messages */
    BP_C128 = 3  /* complex FP64 (ComplexF64, test_beliefpropagation.jl:154): generic kernel;
                    every host buffer of sites, messages and scalars holds interleaved
                    (re, im) doubles, twice the real length; BP sweeps (Jacobi and
                    sequential), scalars, log-norm and iterate_diff only */
};

/* message initialisations (message_environment(f, nn)) */
enum { BP_INIT_IDENTITY = 0, BP_INIT_ONES = 1, BP_INIT_UPLOAD = 2 };

/* convergence criterion of bp_run */
enum { BP_CRIT_FIDELITY = 0, /* reference iterate_diff: 1 - |<m1^, m2^>|^2 */
       BP_CRIT_MAXABS = 1 };  /* max |m_new - m_old| */

/* kernel path selection */
enum { BP_PATH_AUTO = 0, BP_PATH_GENERIC = 1, BP_PATH_VERTEX = 2,
       BP_PATH_MIXED = 3 /* reported by bp_get_path only: the vertex kernel for the networks
                            it covers (bond dimension 8, degree <= 4), the generic kernel for
                            the rest, in the same sweep */ };

typedef struct bp_context bp_context;

const char *bp_last_error(void);
int bp_version(void);

/* Lifecycle.  device: HIP device ordinal.  dtype: BP_F64 / BP_F32 / BP_BF16 / BP_C128. */
int bp_create(bp_context **out, int device, int dtype);
int bp_destroy(bp_context *ctx);
/* Run on a caller-provided hipStream_t (passed as void*); NULL restores the
 * context's own stream.  Streams are used at default priority. */
int bp_set_stream(bp_context *ctx, void *hip_stream);

/* Topology.  plain != 0: single-layer network with vector messages (phys_dim ignored).
 * vertex_net may be NULL (one network).  nnet = number of networks. */
int bp_set_topology(bp_context *ctx, int nv, int ne, const int *port_off, const int *port_nbr,
                    const int *port_edge, const int *port_rev, const int *edge_dim,
                    const int *phys_dim, const int *vertex_net, int nnet, int plain);

/* Site tensors (FP64 host buffer in the site layout); converted to the context dtype. */
int bp_upload_sites(bp_context *ctx, const double *sites);

/* Messages: kind = BP_INIT_*; `host` (FP64, message layout) only for BP_INIT_UPLOAD. */
int bp_init_messages(bp_context *ctx, int kind, const double *host);
int bp_download_messages(bp_context *ctx, double *out);

/* Messages of selected slots, packed in slot-list order into an FP64 host buffer (each
 * slot's D x D or length-D message after the previous one): the per-edge cache access of
 * message_update!(alg, cache, factors, edge) (beliefpropagation.jl:242-257 reads the
 * edge's incoming messages, messagecache.jl:124-131, and writes cache[edge]), so a
 * per-edge strategy moves O(degree * D^2) values per update instead of the whole cache.
 * Reads and writes the current messages (the FP64 ones after bp_polish_fp64). */
int bp_set_slot_messages(bp_context *ctx, int n, const int *slots, const double *data);
int bp_get_slot_messages(bp_context *ctx, int n, const int *slots, double *out);

/* Choose the kernel path (default BP_PATH_AUTO). */
int bp_set_path(bp_context *ctx, int path);
/* The path the next sweep will use (after AUTO resolution). */
int bp_get_path(bp_context *ctx, int *path);

/* One Jacobi sweep over every message of every not-yet-converged network:
 * new = normalise(update(old)) with per-slot residuals. */
int bp_sweep(bp_context *ctx);

/* Full solve, reference stopping semantics per network:
 *   StopAfterIteration(maxiter) | StopWhenConverged(tol), delta skipped at iteration 0.
 * maxiter <= 0: not given; tol < 0: not given.  Per-network outputs (length nnet, any may
 * be NULL): sweeps done, last delta, converged-at iteration (-1 if never).
 * use_graph != 0 captures the sweep sequence in a hipGraph and replays it. */
int bp_run(bp_context *ctx, int maxiter, double tol, int criterion, int use_graph,
           int *out_sweeps, double *out_delta, int *out_at_iteration);

/* Prepares a later bp_run(ctx, maxiter, ..., use_graph != 0): allocates what the sweeps
 * need and captures the sweep graphs that run replays (16-sweep chunks and the last,
 * maxiter % 16 sweeps, each from either message buffer), running no sweep.  No counterpart
 * in the reference (its solve has no setup step to hoist): for callers that time runs or
 * start them in latency-sensitive loops.  A later change that invalidates them (topology,
 * options, kernel choices) makes bp_run capture again. */
int bp_prepare_run(bp_context *ctx, int maxiter);

/* Sequential (Gauss-Seidel) update of one message slot in place, normalised. */
int bp_update_slot(bp_context *ctx, int slot);

/* Sequential solve with the reference's own schedule semantics: each sweep updates the
 * slots of `order` one after another in place (beliefpropagation.jl:200-210, the
 * default order being forest_cover_edge_sequence, beliefpropagation.jl:14), then the
 * sweep's delta against the previous sweep decides StopWhenConverged(tol); stopping and
 * outputs as bp_run (single network). */
int bp_run_sequential(bp_context *ctx, const int *order, int norder, int maxiter, double tol,
                      int criterion, int *out_sweeps, double *out_delta, int *out_at_iteration);

/* Per-slot residuals of the last sweep (FP64, length nports), either may be NULL. */
int bp_residuals(bp_context *ctx, double *fid, double *maxabs);

/* iterate_diff between the current messages and an FP64 host message buffer. */
int bp_iterate_diff(bp_context *ctx, const double *other, double *out);

/* bp_iterate_diff that allocates nothing: no counterpart in the reference.  For callers
 * that cannot allocate device memory per call (tight convergence-check loops, serving
 * loops, harnesses that forbid hipMalloc).  The same result as bp_iterate_diff; the input
 * is staged and the per-slot distances kept in the library's static per-device scratch
 * array (8 MiB, 1 Mi doubles), so 2 * nports + (message entries) + 1 must not exceed
 * 1048576 (BP_ERR_ARG otherwise; 32x32 D=8 needs 261889).  Calls that overlap on one GPU
 * share that array and race. */
int bp_iterate_diff_noalloc(bp_context *ctx, const double *other, double *out);

/* Vertex scalars Z_v (length nv) and edge scalars Z_e (length ne, edge order of the
 * topology), FP64 accumulation; BP_C128 contexts write interleaved complex values
 * (2 nv and 2 ne doubles). */
int bp_vertex_scalars(bp_context *ctx, double *zv);
int bp_edge_scalars(bp_context *ctx, double *ze);

/* BP log-norm per network (length nnet): real and imaginary parts (imaginary part is
 * pi * (#negative vertex terms - #negative edge terms), as the reference's complex
 * promotion gives; in BP_C128 contexts the sum of the terms' principal arguments);
 * -Inf when an edge term is zero.  Compensated FP64 sums. */
int bp_lognorm(bp_context *ctx, double *out_re, double *out_im);

/* FP64 polishing: run `nsweeps` FP64 Jacobi sweeps from the current messages on the FP64
 * copy of the sites, and make the FP64 result the context's messages for scalars and
 * download.  The sweeps run on an FP64 vertex kernel when one covers every network
 * (largest bond dimension 6..8, or 9..16 with the large-bond kernel), else on the generic
 * kernel.  Per-network last delta (max-abs) in out_delta (may be NULL). */
int bp_polish_fp64(bp_context *ctx, int nsweeps, double *out_delta);

/* Gradient of the BP log-norm with respect to every site tensor at the current fixed
 * point (FP64 host output in the site layout). */
int bp_grad_lognorm(bp_context *ctx, double *grad);

/* Transverse-field Ising energy E = sum_<uv> -J <Z_u Z_v>_BP + sum_v -h <X_v>_BP at the
 * current messages (phys dim 2), and optionally its gradient with respect to every site
 * tensor by fixed-point adjoint sweeps (grad may be NULL).  adj_tol / adj_maxiter bound
 * the adjoint solve; out_adj_iters / out_adj_res report it. */
int bp_tfim_energy(bp_context *ctx, double J, double h, double *energy, double *grad,
                   double adj_tol, int adj_maxiter, int *out_adj_iters, double *out_adj_res);

/* bp_grad_lognorm and bp_tfim_energy with the gradient left in device memory: `grad` is a
 * device buffer of nsite entries in the site layout, of the context's arithmetic type
 * (double for BP_F64 contexts, float for BP_F32 and BP_BF16), written on the context's
 * stream before the call returns.  No reference counterpart (the reference has no AD);
 * this is the pullback an rrule or torch.autograd.Function hands to a device optimiser
 * without a host round trip. */
int bp_grad_lognorm_device(bp_context *ctx, void *grad);
int bp_tfim_energy_device(bp_context *ctx, double J, double h, double *energy, void *grad,
                          double adj_tol, int adj_maxiter, int *out_adj_iters,
                          double *out_adj_res);

/* BP expectation values at the current messages, with arbitrary real operators on the
 * physical index (acting on the ket layer: (O.A)[s', X] = sum_s O[s', s] A[s, X]).
 * Operators are ld x ld row-major blocks (ld >= every physical dimension; a vertex of
 * physical dimension d uses the top-left d x d corner).
 *   bp_expect_vertex: ops[v] for every vertex v; out[v] = Z_v[O_v] / Z_v, the operator-
 *     inserted vertex scalar over the plain one (messagecache.jl:139-143 region scalar
 *     with an insertion).
 *   bp_expect_edge: per edge e (edge ids in the order of the topology's edges; the edge's
 *     first vertex owns its lower message slot) ops_first[e] on the first vertex and
 *     ops_second[e] on the second; out[e] = <g^A_{u->v}, g^B_{v->u}> / <g_{u->v}, g_{v->u}>
 *     with g^O the unnormalised message of the sender with O inserted (the two-site region
 *     {u, v} with its boundary messages, messagecache.jl:134-137 incoming_edge_data).
 * Norm networks only.  The reference has no expectation API (its region scalars are the
 * operator-free case); FP64 results, contractions in the context's arithmetic type. */
int bp_expect_vertex(bp_context *ctx, const double *ops, int ld, double *out);
int bp_expect_edge(bp_context *ctx, const double *ops_first, const double *ops_second, int ld,
                   double *out);

/* Local energy E = sum_v <O_v>_BP + sum_e <A_e B_e>_BP (the expectation values of
 * bp_expect_vertex / bp_expect_edge, same operator layout; ops_v, or the ops_first /
 * ops_second pair, may be NULL to drop that sum) at the current messages, and optionally
 * its gradient with respect to every site tensor by fixed-point adjoint sweeps (grad may be
 * NULL; FP64 host array in the site layout, or, for the _device variant, a device buffer
 * as bp_tfim_energy_device takes).  The TFIM energy is the case O_v = -h X,
 * A_e = -J Z, B_e = Z.  Generic (VALU) kernels in the context's arithmetic type.  No
 * reference counterpart (the reference has no AD or expectation API). */
int bp_local_energy(bp_context *ctx, const double *ops_v, const double *ops_first,
                    const double *ops_second, int ld, double *energy, double *grad,
                    double adj_tol, int adj_maxiter, int *out_adj_iters, double *out_adj_res);
int bp_local_energy_device(bp_context *ctx, const double *ops_v, const double *ops_first,
                           const double *ops_second, int ld, double *energy, void *grad,
                           double adj_tol, int adj_maxiter, int *out_adj_iters,
                           double *out_adj_res);

/* Kernel options (what the autotuner decides, settable by hand):
 *   "vertex_variant"   0: D <= 8 vertex sweep kernel (FP32 or BF16) with two LDS buffers
 *                      per wave, 1: one buffer per wave and two waves per SIMD (the default
 *                      in BF16, and in FP32 at physical dimension 1); FP32 only: 2: open
 *                      contractions on MFMA, bra in LDS (one wave per SIMD), 3: open
 *                      contractions on MFMA, bra rows from padded X/Y site copies in HBM
 *                      (two waves per SIMD),
 *                      4: as 1 with two of the four open contractions per slice on MFMA
 *                      (the partner message absorbed into the bra), 5: all four (the FP32
 *                      default at physical dimension >= 2; also BF16, with BF16 mode
 *                      products);
 *   "vertex16_waves"   waves per workgroup of the BF16 kernel for bond dimensions 9..16
 *                      (4, 8 or 16; default 16);
 *   "vertex_bond_mask" bit D set: networks whose largest bond dimension is D may run on the
 *                      vertex kernels (D 1..12 in FP32, 1..16 in BF16, default 0x1FFFE;
 *                      FP64: 6..8 on the D <= 8 kernel, 9..16 on the large-bond kernel,
 *                      default 0x1FFC0);
 *   "small_bond"       1 (default): FP32 networks whose bond dimensions are all <= 4 run on
 *                      the small-bond kernels; 0: padded to 8 on the D <= 8 kernels;
 *   "fold_bookkeeping" 1 (default): when one D <= 8 vertex launch covers every network, the
 *                      sweep kernels also do the per-network convergence bookkeeping;
 *                      0: a separate bookkeeping kernel after every sweep (same results);
 *   "normalize"        1 (default): every update divides the new message by its sum when
 *                      the sum is nonzero; 0: messages are kept unnormalised, as
 *                      SimpleMessageUpdate(; normalize = false) (beliefpropagation.jl:237-253);
 *                      sweeps then run on the generic kernel and gradients are refused.
 * Only "normalize" has a reference counterpart. */
int bp_set_option(bp_context *ctx, const char *name, int value);
/* The current value of an option above (after bp_autotune: the tuned choice).
 * "vertex_variant" reads -1 while no D <= 8 vertex plan exists (no sites or topology yet,
 * or the vertex kernels declined the network) and no variant was set.  No reference
 * counterpart. */
int bp_get_option(bp_context *ctx, const char *name, int *value);

/* Autotuner.  Times, on the context's own sites and messages, (a) for every largest bond
 * dimension D present that the vertex kernels take (<= 12 FP32, <= 16 BF16 and FP64; norm networks,
 * automatic path) whether its networks sweep faster on the vertex kernels (bonds padded to
 * 8, 12 or 16) or on the generic kernel, and
 * (b) the D <= 8 vertex sweep kernel variant ("vertex_variant" of bp_set_option) and the
 * waves per workgroup of the bond-dimension-16 kernel ("vertex16_waves"), and
 * (c) for every bucket of generic-kernel slots (owner's physical dimension and largest
 * bond dimension) the workgroup size (128/256/512) and whether intermediates live in LDS
 * (up to 64 KiB or up to 160 KiB) or in a global workspace.  mode 0: apply the choices found in `cache_path` only;
 * 1: also tune what the cache lacks and rewrite it; 2: retune everything and rewrite it.
 * cache_path may be NULL (no file).  The messages are restored afterwards.  out_tuned (may
 * be NULL) receives the number of decisions timed.  No reference counterpart (the
 * reference dispatches on Julia types, select_algorithm.jl:7-57, and has no kernels). */
int bp_autotune(bp_context *ctx, const char *cache_path, int mode, int *out_tuned);

/* Benchmark helper: launch one untimed and then `n` timed Jacobi sweeps back to back (no
 * convergence bookkeeping, messages ping-pong) and report the average device duration of
 * one sweep's kernel launches, timed with HIP events on the context's stream around each
 * sweep. */
int bp_time_sweeps(bp_context *ctx, int n, double *avg_ms);

/* Elapsed milliseconds of the last bp_run, measured with HIP events on the context's
 * stream around the sweep launches. */
int bp_last_run_ms(bp_context *ctx, double *ms);

/* ---------------------------------------------------------------- exact contraction
 * Exact contraction of a labelled tensor network (contract_network with the Exact
 * algorithm, contract_network.jl:12-37; the order choices of contraction_order,
 * :40-60).  Tensors t = 0..nt-1 carry ranks[t] labels, concatenated in `labels`; label l
 * has dimension label_dim[l] (0 <= l < n_labels).  A label on two tensors is summed over,
 * a label on one stays open; no label may appear twice on one tensor or on more than two
 * tensors.  The data of every tensor is row-major over its labels, the tensors
 * concatenated in order.
 *   order: NULL, or nt-1 pairwise steps (order[2s], order[2s+1]) over ids 0..nt-1 (the
 *     tensors) and nt+s' (the result of step s'), each id used once (Exact's `order`);
 *   order_alg (when order is NULL): BP_CN_ORDER_LEFT, ((T0 T1) T2)... (contraction_order
 *     with LeftAssociative); BP_CN_ORDER_GREEDY: repeatedly the pair of live operands
 *     sharing a label whose result grows the live size least (size(result) - size(a) -
 *     size(b); ties to the earlier pair, results appended at the end; the role of the
 *     reference's default Greedy() order_alg, whose algorithm lives in ITensorBase);
 *     BP_CN_ORDER_CHEAPEST: the tree with the fewest FLOPs (then the smallest largest
 *     intermediate) among the left-associative one and three greedy ones (costs: growth,
 *     result size, result size over the larger operand's size).  The contracted values do
 *     not depend on the order beyond roundoff; the result's label order does.
 * Every step is one FP64 product C[a's open labels, b's open labels] = sum over the shared
 * labels, so the result's labels are in that order (for the left-associative order: the
 * open labels in first-appearance order of the tensor list).  Intermediates (and the
 * gathered operands of transposes) may hold at most max_elems entries (<= 0: 2^31);
 * BP_ERR_UNSUPPORTED otherwise.  An empty network contracts to the scalar 1.
 *
 * A plan (order, layouts, offset tables, device workspace) is built once and executed on
 * any number of inputs of the same shapes: on host buffers (bp_cn_execute) or device
 * buffers of this library's HIP runtime (bp_cn_execute_device, result left in device
 * memory).  Both return after the result is written; bp_last_run_ms then gives the device
 * time of the contraction's launches.  bp_cn_plan_info (any output may be NULL): the
 * result's rank, labels and entries, the input entries, the products' FP64 FLOPs, their
 * algorithmic bytes (every operand read once, every result written once) and the device
 * workspace.  bp_contract_network plans, executes and frees in
 * one call (out_cap: capacity of `out` in entries, out_labels: room for 64 labels). */
typedef struct bp_cn_plan bp_cn_plan;
enum { BP_CN_ORDER_LEFT = 0, BP_CN_ORDER_GREEDY = 1, BP_CN_ORDER_CHEAPEST = 2 };
int bp_cn_plan_create(bp_context *ctx, int nt, const int *ranks, const int *labels,
                      int n_labels, const int *label_dim, const int *order, int order_alg,
                      long long max_elems, bp_cn_plan **out);
int bp_cn_plan_info(const bp_cn_plan *plan, int *out_rank, int *out_labels,
                    long long *out_size, long long *in_size, double *flops, double *bytes,
                    long long *workspace_bytes);
int bp_cn_execute(bp_context *ctx, bp_cn_plan *plan, const double *data, double *out);
int bp_cn_execute_device(bp_context *ctx, bp_cn_plan *plan, const void *data, void *out);
int bp_cn_plan_destroy(bp_cn_plan *plan);
int bp_contract_network(bp_context *ctx, int nt, const int *ranks, const int *labels,
                        int n_labels, const int *label_dim, const double *data,
                        const int *order, int order_alg, long long max_elems, int *out_rank,
                        int *out_labels, double *out, long long out_cap);

/* ---------------------------------------------------------------- gate application
 * BP-gauged gate application (BPApplyGate, apply_gate_bp_nsite!, src/apply/
 * apply_operators.jl:181-283) for a batch of gates, one workgroup per gate, FP64.  Gates of
 * one call act independently (no gate may touch a site tensor or message another gate of
 * the same call changes); a sequence of overlapping gates is a sequence of calls.
 *   desc: BP_GATE_DESC ints per gate: kind (1: one site, 2: two sites), d1, d2 (physical
 *     dimensions), k1, k2 (port counts, <= 8), p1, p2 (the port of each site on the gate's
 *     bond; -1 for one site), maxdim (<= 0: no rank cap), normalize (0/1), then the 8 port
 *     dimensions of site 1 and the 8 of site 2 (unused entries 0).
 *   in: per gate, packed in this order: site 1's tensor [d1][port 0]...[port k1-1]
 *     (row-major), site 2's tensor (two sites), the incoming messages (D x D, [ket][bra]) on
 *     every port of site 1 except p1 in port order (one site: on every port when
 *     normalize, none otherwise), then site 2's, then the gate: two sites
 *     [t1][t2][s1][s2] (out indices first), one site [t][s].
 *   out: per gate, at the capacity bp_gate_sizes reports: site 1's new tensor (the same
 *     layout with port p1's dimension replaced by chi, the first d1*P1*chi entries), site
 *     2's new tensor at its capacity, then the chi singular values (sorted, descending;
 *     normalised to unit norm when normalize).  The messages on the gate's bond become
 *     diag(S) both ways.  One-site gates write the new tensor only (chi = 0).
 *   chi: per gate, the kept rank: min(d1 min(P1, d1 Db), d2 min(P2, d2 Db)) capped by
 *     maxdim, then singular values <= rtol * S_max dropped (rtol < 0: none), at least 1.
 * The gauge of a message M is X = V sqrt(lam) from M = V diag(lam) V^T, its inverse the
 * pseudo-inverse (lam <= 1e-13 max(lam) dropped); the site factorisations and the gate's
 * SVD are one-sided Jacobi SVDs.  bp_last_run_ms gives the kernel's device time. */
enum { BP_GATE_DESC = 25 };
int bp_gate_sizes(int ngates, const int *desc, long long *in_size, long long *out_size);
int bp_apply_gates(bp_context *ctx, int ngates, const int *desc, const double *in,
                   double *out, int *chi, double rtol);

/* Gates on the context's resident network: the same algorithm applied in place to the
 * context's FP64 site tensors and current messages, with no host copy of either (the
 * network stays on the device between gate layers and BP runs; ITensorNetwork /
 * MessageCache are needed only to import and export, messagecache.jl:205-225,
 * tensornetwork.jl:138-173).  kind[g] = 2: a two-site gate on the bond of message slot
 * where[g] (site 1 = the slot's owner, site 2 = its neighbour); kind[g] = 1: a one-site gate
 * on vertex where[g].  gates: the gate matrices packed in order ([t1][t2][s1][s2], or [t][s]).
 * No two gates of one call may share a vertex.  Each bond keeps its dimension De: at most
 * min(De, maxdim) singular values are kept (maxdim <= 0: De; then rtol as in
 * bp_apply_gates) and the tensors and diag(S) are zero-padded to De; both messages of the bond become diag(S), in the working
 * messages and, after bp_polish_fp64, in the FP64 ones.  out_chi (per gate) and out_s
 * (ngates x s_stride, zero-padded; s_stride >= every gated bond's De) may be NULL.  Norm
 * networks only.  bp_last_run_ms gives the device time of the gather, gate and scatter
 * kernels; the working site copies are refreshed before return. */
int bp_apply_gates_resident(bp_context *ctx, int ngates, const int *kind, const int *where,
                            const double *gates, int maxdim, int normalize, double rtol,
                            int *out_chi, double *out_s, int s_stride);

/* The context's FP64 site tensors (site layout), e.g. after resident gates. */
int bp_download_sites(bp_context *ctx, double *sites);

#ifdef __cplusplus
}
#endif
#endif
