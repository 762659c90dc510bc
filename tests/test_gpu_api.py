"""The reference's own BP tests, ported to the reference-shaped API and run on the GPU
(test/test_beliefpropagation.jl:134-225 and the scalars of :83-115)."""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _links(pkg, g, dim):
    return {frozenset(e): pkg.Index(dim) for e in g.edges()}


def _plain_random(pkg, g, dim, rng):
    l = _links(pkg, g, dim)
    return pkg.tensornetwork(
        lambda v: pkg.random_itensor([l[frozenset(e)] for e in g.incident_edges(v)], rng),
        g.vertices()), l


def _ones_messages(tn, g):
    return {e: np.ones(int(np.prod([i.dim for i in tn.linkinds(e)]))) for e in g.all_edges()}


def _exact_plain(tn, g):
    letters = iter("abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ")
    name, subs, ops = {}, [], []
    for v in g.vertices():
        t = tn[v]
        subs.append("".join(name.setdefault(i, next(letters)) for i in t.inds))
        ops.append(t.data)
    return float(np.einsum(",".join(subs) + "->", *ops))


@pytest.mark.parametrize("dtype,rtol", [("f32", np.finfo(np.float32).eps ** (1 / 3)),
                                        ("f64", np.finfo(np.float64).eps ** (1 / 3))])
def test_algorithm_chain_and_tree_maxiter1(pkg, dtype, rtol):
    """BP with maxiter = 1 (one sequential forest-cover sweep) is exact on a chain and on a
    comb tree (test_beliefpropagation.jl:153-202)."""
    _say(f"case api chain/tree {dtype}")
    rng = np.random.default_rng(123)
    for g, dim in ((pkg.named_path_graph(2), 2), (pkg.named_comb_tree((4, 3)), 3)):
        tn, _ = _plain_random(pkg, g, dim, rng)
        cache = pkg.beliefpropagation(tn, _ones_messages(tn, g), stopping_criterion={"maxiter": 1},
                                      message_update_algorithm={"dtype": dtype})
        z_bp = np.exp(complex(pkg.bethe_free_energy(tn, cache)))
        z_exact = _exact_plain(tn, g)
        assert abs(z_bp.real / z_exact - 1) < rtol
        assert cache.state.iteration == 1


@pytest.mark.parametrize("n", [3, 4, 5])
def test_spin_ice_analytical(pkg, n):
    """exp(bethe_free_energy) == 1.5^(n^2) on the periodic n x n spin-ice network
    (test_beliefpropagation.jl:204-225), default strategy, maxiter 10, tol 1e-10; the
    reference's `≈` uses Julia's default rtol, sqrt(eps(Float64))."""
    _say(f"case api spin ice {n}")
    g = pkg.named_grid((n, n), periodic=True)
    l = _links(pkg, g, 2)
    t_data = np.zeros((2, 2, 2, 2))
    for idx in np.ndindex(2, 2, 2, 2):
        if sum(idx) == 2:
            t_data[idx] = 1.0
    tn = pkg.tensornetwork(
        lambda v: pkg.ITensor(t_data, [l[frozenset(e)] for e in g.incident_edges(v)]), g.vertices())
    rng = np.random.default_rng(7)
    msgs = {e: rng.random(2) for e in g.all_edges()}
    cache = pkg.beliefpropagation(tn, msgs, stopping_criterion={"maxiter": 10, "tol": 1e-10})
    assert math.exp(pkg.bethe_free_energy(tn, cache)) == pytest.approx(
        1.5 ** (n * n), rel=math.sqrt(np.finfo(np.float64).eps))


def _peps(pkg, L, D, seed):
    g = pkg.named_grid((L, L))
    l = _links(pkg, g, D)
    s = {v: pkg.Index(2) for v in g.vertices()}
    rng = np.random.default_rng(seed)
    tn = pkg.tensornetwork(lambda v: pkg.random_itensor(
        [s[v]] + [l[frozenset(e)] for e in g.incident_edges(v)], rng), g.vertices())
    return g, tn, pkg.normnetwork(tn)


def test_norm_network_bp_strategies_agree(pkg, orc):
    """The reference-default sequential strategy and the Jacobi device strategy reach the
    same fixed point and BP log-norm on a 4x4 D=2 PEPS norm network; both match the FP64
    oracle, and the scalars obey region_scalar == prod(vertex_scalars)."""
    _say("case api norm network strategies")
    g, tn, nn = _peps(pkg, 4, 2, 11)
    init = pkg.message_environment(pkg.one, nn)
    seq = pkg.beliefpropagation(nn, init, stopping_criterion={"maxiter": 200, "tol": 1e-14})
    jac = pkg.beliefpropagation(nn, init, stopping_criterion={"maxiter": 400, "tol": 1e-14},
                                message_update_algorithm=pkg.JacobiGPUUpdate(dtype="f64"))
    assert seq.state.at_iteration > 0 and jac.state.at_iteration > 0
    # the reference's criterion is a squared-angle fidelity: tol 1e-14 certifies messages to
    # about sqrt(1e-14) = 1e-7
    for e in seq.keys():
        np.testing.assert_allclose(seq[e], jac[e], atol=1e-6)
    ln_s = pkg.bethe_free_energy(nn, seq)
    ln_j = pkg.bethe_free_energy(nn, jac)
    assert ln_s == pytest.approx(ln_j, abs=1e-9)
    fl = pkg.to_topology(nn)
    want, _, _, _ = orc.bp(fl.topo, fl.sites, init.to_flat(fl.topo), maxiter=400, tol=1e-14)
    assert ln_j == pytest.approx(orc.bethe_free_energy(fl.topo, fl.sites, want), abs=1e-10)
    zv = pkg.vertex_scalars(nn, jac)
    region = g.vertices()[:3]
    assert pkg.region_scalar(nn, jac, region) == pytest.approx(
        float(np.prod(pkg.vertex_scalars(nn, jac, region))), rel=1e-12)
    assert len(pkg.edge_scalars(jac)) == g.ne()
    assert pkg.iterate_diff(jac, jac.copy()) == pytest.approx(0.0, abs=10 * np.finfo(float).eps)
    assert zv.shape == (g.nv(),)


def test_single_message_update_and_custom_strategy(pkg, orc):
    _say("case api message_update_")
    g, tn, nn = _peps(pkg, 3, 3, 12)
    cache = pkg.message_environment(pkg.one, nn)
    fl = pkg.to_topology(nn)
    e = ((2, 2), (2, 3))
    before = cache.to_flat(fl.topo)
    pkg.message_update_(cache, nn, e)
    q = fl.slot(e)
    want = orc.message(fl.topo, fl.sites, before, q)
    want = want / want.sum()
    np.testing.assert_allclose(cache[e].ravel(), want, rtol=1e-12)

    class Delegating(pkg.MessageUpdateAlgorithm):
        calls = 0

        def update(self, cache, factors, edge):
            Delegating.calls += 1
            return pkg.SimpleMessageUpdate().update(cache, factors, edge)

    g2, tn2, nn2 = _peps(pkg, 2, 2, 13)
    out = pkg.beliefpropagation(nn2, pkg.message_environment(pkg.one, nn2),
                                stopping_criterion={"maxiter": 2},
                                message_update_algorithm=Delegating())
    assert Delegating.calls == 2 * 2 * g2.ne()
    ref = pkg.beliefpropagation(nn2, pkg.message_environment(pkg.one, nn2),
                                stopping_criterion={"maxiter": 2})
    for k in out.keys():
        np.testing.assert_allclose(out[k], ref[k], rtol=1e-12)
