"""BP-gauged gate application on the device (bp_apply_gates; apply_operator /
apply_operators with BPApplyGate), checked as the reference's own tests check it
(test/test_apply_operator.jl), with the oracle's exact contraction (pinned in
test_oracle.py) as the reference and numpy's SVD for the global optimum:

  * untruncated one- and two-site gates on a 4-cycle are exact (:52-74);
  * a truncated two-site gate on a 4-path with converged BP messages equals the global
    optimal rank-k truncation of the gated state, and its singular values are proportional
    to the state's Schmidt values (:76-98);
  * apply_operators applies a sequence (:100-120).

Float64, and Float32 (computed in FP64, Float32 results); the ComplexF64 and U(1)-graded
variants of the reference test are outside the engine's real kernels (complex inputs raise
ArgumentError).  Tolerance: 1e-9 relative for Float64 (the reference allows eps^(1/3) =
6e-6), 1e-5 for Float32 (the reference allows eps(Float32)^(1/3) = 4.9e-3).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

TOL = 1e-9


def _say(msg):
    print(msg, flush=True)


def randn_operator(pkg, rng, domain):
    codomain = [pkg.uniquename(i) for i in domain]
    shape = [i.dim for i in codomain] + [i.dim for i in domain]
    return pkg.operator(rng.standard_normal(shape), codomain, domain)


def _ones(m):
    return np.ones_like(m)


def random_state(pkg, rng, g, site, nlayers, trunc):
    """test_apply_operator.jl:27-42: a product state with dimension-1 links, then layers of
    random two-site gates on every edge (applied by the engine, truncated)."""
    links = {frozenset(e): pkg.Index(1) for e in g.edges()}
    tn = pkg.tensornetwork(lambda v: pkg.ITensor(
        rng.standard_normal([site[v].dim] + [1] * g.degree(v)),
        [site[v]] + [links[frozenset(e)] for e in g.incident_edges(v)]), g.vertices())
    env = pkg.message_environment(_ones, pkg.normnetwork(tn))
    for _ in range(nlayers):
        for (a, b) in g.edges():
            gate = randn_operator(pkg, rng, [site[a], site[b]])
            tn, env = pkg.apply_operator(gate, tn, env, trunc=trunc)
    return tn


def _bp_env(pkg, tn):
    nn = pkg.normnetwork(tn)
    return pkg.beliefpropagation(nn, pkg.message_environment(_ones, nn),
                                 stopping_criterion={"maxiter": 100, "tol": 1e-13})


def _state(orc, tn, site_order):
    """prod(tn) as an array over `site_order` (exact contraction, the oracle)."""
    vs = tn.vertices()
    val, labs = orc.contract_labeled([tn[v].data for v in vs], [tn[v].inds for v in vs])
    return np.transpose(val, [labs.index(i) for i in site_order])


def _apply_full(op, psi, site_order):
    """ITB.apply(op, psi) on a full state array over site_order."""
    pos = [site_order.index(i) for i in op.domain]
    n = len(op.domain)
    out = np.tensordot(op.data, psi, axes=(list(range(n, 2 * n)), pos))
    # out axes: codomain (in domain order), then the remaining sites in order
    rest = [k for k in range(len(site_order)) if k not in pos]
    perm = [None] * len(site_order)
    for a, p in enumerate(pos):
        perm[p] = a
    for a, p in enumerate(rest):
        perm[p] = n + a
    return np.transpose(out, perm)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _cycle(pkg, n):
    g = pkg.named_path_graph(n)
    g.add_edge(n, 1)
    return g


def test_untruncated_gates_are_exact(pkg, orc):
    _say("case gates exact on a 4-cycle")
    rng = np.random.default_rng(123)
    g = _cycle(pkg, 4)
    site = {v: pkg.Index(3) for v in g.vertices()}
    tn = random_state(pkg, rng, g, site, 2, pkg.truncrank(4))
    env = _bp_env(pkg, tn)
    order = [site[v] for v in g.vertices()]
    psi = _state(orc, tn, order)
    for gate in (randn_operator(pkg, rng, [site[2]]),
                 randn_operator(pkg, rng, [site[2], site[3]])):
        gated, env2 = pkg.apply_operator(gate, tn, env)
        err = _rel(_state(orc, gated, order), _apply_full(gate, psi, order))
        _say(f"{len(gate.domain)}-site gate: rel err {err:.2e}")
        assert err < TOL
    # the two-site gate replaced the bond's messages by diag(S) both ways
    S = np.diag(env2[(2, 3)])
    np.testing.assert_array_equal(env2[(2, 3)], env2[(3, 2)])
    assert np.all(np.diff(S) <= 0) and S[-1] >= 0


@pytest.mark.parametrize("k", [1, 2, 3])
def test_truncated_gate_matches_global_svd(pkg, orc, k):
    _say(f"case truncated gate rank {k} on a 4-path")
    rng = np.random.default_rng(123)
    g = pkg.named_path_graph(4)
    site = {v: pkg.Index(3) for v in g.vertices()}
    tn = random_state(pkg, rng, g, site, 2, pkg.truncrank(4))
    env = _bp_env(pkg, tn)
    order = [site[v] for v in g.vertices()]
    gate = randn_operator(pkg, rng, [site[2], site[3]])
    full = _apply_full(gate, _state(orc, tn, order), order)
    U, S, Vt = np.linalg.svd(full.reshape(9, 9), full_matrices=False)
    best = (U[:, :k] * S[:k]) @ Vt[:k]
    gated, env2 = pkg.apply_operator(gate, tn, env, trunc=pkg.truncrank(k))
    got = _state(orc, gated, order).reshape(9, 9)
    err = _rel(got, best)
    # the bond's singular values are the Schmidt values up to one factor (BP messages are
    # normalised to unit sum, which scales the environment)
    Sd = np.diag(env2[(2, 3)])
    errS = float(np.max(np.abs(Sd / Sd[0] - S[:k] / S[0])))
    _say(f"rel err {err:.2e}, Schmidt values err {errS:.2e}")
    assert err < TOL
    assert errS < TOL


def test_apply_operators_sequence(pkg, orc):
    _say("case gate sequence on a 4-cycle")
    rng = np.random.default_rng(123)
    g = _cycle(pkg, 4)
    site = {v: pkg.Index(3) for v in g.vertices()}
    tn = random_state(pkg, rng, g, site, 2, pkg.truncrank(4))
    env = _bp_env(pkg, tn)
    order = [site[v] for v in g.vertices()]
    g1 = randn_operator(pkg, rng, [site[2], site[3]])
    g2 = randn_operator(pkg, rng, [site[3], site[4]])
    gated, _ = pkg.apply_operators([g1, g2], tn, env)
    want = _apply_full(g2, _apply_full(g1, _state(orc, tn, order), order), order)
    err = _rel(_state(orc, gated, order), want)
    _say(f"rel err {err:.2e}")
    assert err < TOL


def test_batched_gates_equal_sequential(pkg, orc):
    """A layer of gates on disjoint bonds of a 4x4 grid is one device batch; each gate's
    result equals applying it alone (bitwise: the gates run in separate workgroups), and the
    layer is exact untruncated."""
    _say("case batched layer on a 4x4 grid")
    rng = np.random.default_rng(7)
    g = pkg.named_grid((4, 4))
    site = {v: pkg.Index(2) for v in g.vertices()}
    tn = random_state(pkg, rng, g, site, 1, pkg.truncrank(3))
    env = _bp_env(pkg, tn)
    layer = [((i, j), (i, j + 1)) for i in range(1, 5) for j in (1, 3)]
    gates = [randn_operator(pkg, rng, [site[a], site[b]]) for a, b in layer]
    batched, envb = pkg.apply_operators(gates, tn, env)
    seq, envs = tn, env
    for gt in gates:
        seq, envs = pkg.apply_operator(gt, seq, envs)
    for (a, b) in layer:
        np.testing.assert_array_equal(envb[(a, b)], envs[(a, b)])
    order = [site[v] for v in g.vertices()]
    want = _state(orc, tn, order)
    for gt in gates:
        want = _apply_full(gt, want, order)
    err = _rel(_state(orc, batched, order), want)
    _say(f"layer rel err {err:.2e}")
    assert err < TOL


def test_normalize(pkg, orc):
    """normalize: two-site singular values of unit norm; a one-site gate divides the tensor
    by its norm with the square roots of its incoming messages (BP norm of the vertex)."""
    _say("case normalize")
    rng = np.random.default_rng(11)
    g = _cycle(pkg, 4)
    site = {v: pkg.Index(3) for v in g.vertices()}
    tn = random_state(pkg, rng, g, site, 1, pkg.truncrank(3))
    env = _bp_env(pkg, tn)
    gate2 = randn_operator(pkg, rng, [site[1], site[2]])
    _, env2 = pkg.apply_operator(gate2, tn, env, normalize=True, trunc=pkg.truncrank(2))
    S = np.diag(env2[(1, 2)])
    assert len(S) == 2 and np.linalg.norm(S) == pytest.approx(1.0, rel=1e-13)
    gate1 = randn_operator(pkg, rng, [site[3]])
    gated, _ = pkg.apply_operator(gate1, tn, env, normalize=True)
    # the BP norm of vertex 3 after the gate: sum over the ket and bra of the tensor with the
    # incoming messages, i.e. the vertex scalar, must be 1
    z = pkg.vertex_scalar(pkg.normnetwork(gated), env, 3)
    assert z == pytest.approx(1.0, rel=1e-10)


def test_float32_network_keeps_its_element_type(pkg, orc):
    """The reference's Float32 variant (test_apply_operator.jl:55-58): a Float32 network and
    Float32 gates give Float32 tensors (computed in FP64; the message cache keeps its FP64
    messages), and untruncated gates stay exact within the reference's eps(Float32)^(1/3);
    measured at Float32 rounding."""
    _say("case float32 gates exact on a 4-cycle")
    rng = np.random.default_rng(123)
    g = _cycle(pkg, 4)
    site = {v: pkg.Index(3) for v in g.vertices()}
    tn = random_state(pkg, rng, g, site, 2, pkg.truncrank(4))
    tn32 = tn.copy()
    for v in g.vertices():
        tn32[v] = pkg.ITensor(tn[v].data.astype(np.float32), tn[v].inds)
    env = _bp_env(pkg, tn32)
    order = [site[v] for v in g.vertices()]
    psi = _state(orc, tn32, order)
    for doms in ([site[2]], [site[2], site[3]]):
        gate = randn_operator(pkg, rng, doms)
        gate = pkg.operator(gate.data.astype(np.float32), gate.codomain, gate.domain)
        gated, env2 = pkg.apply_operator(gate, tn32, env)
        assert all(gated[v].data.dtype == np.float32 for v in g.vertices())
        err = _rel(_state(orc, gated, order), _apply_full(gate, psi, order))
        _say(f"{len(doms)}-site Float32 gate: rel err {err:.2e}")
        assert err < np.finfo(np.float32).eps ** (1 / 3) and err < 1e-5
