"""Host-side logic of the reference-shaped API (CPU only, no device calls).

Mirrors test/test_algorithmsinterfaceextensions.jl, test/test_tensornetwork.jl and
test/test_normnetwork.jl name-map semantics, plus select_algorithm's rules
(src/select_algorithm.jl:16-51) and the stopping-criterion shorthand
(src/beliefpropagation/beliefpropagation.jl:16-55)."""
import numpy as np
import pytest


def test_select_algorithm_rules(pkg):
    f = pkg.message_update_
    default = pkg.select_algorithm(f, None, ())
    assert isinstance(default, pkg.SimpleMessageUpdate)
    inst = pkg.JacobiGPUUpdate(dtype="f64")
    assert pkg.select_algorithm(f, inst, ()) is inst
    with pytest.raises(pkg.ArgumentError):
        pkg.select_algorithm(f, inst, (), dtype="f32")
    via_dict = pkg.select_algorithm(f, {"dtype": "f32"}, ())
    assert isinstance(via_dict, pkg.SimpleMessageUpdate) and via_dict.dtype == "f32"
    with pytest.raises(pkg.ArgumentError):
        pkg.select_algorithm(f, {"dtype": "f32"}, (), normalize=True)
    with pytest.raises(pkg.MethodError):
        pkg.default_algorithm(len, ())


def test_stopping_criterion_shorthand(pkg):
    sel = pkg.select_beliefpropagation_stopping_criterion
    c = sel({"maxiter": 10})
    assert isinstance(c, pkg.StopAfterIteration) and c.limits() == (10, None)
    c = sel({"tol": 1e-10})
    assert isinstance(c, pkg.StopWhenConverged) and c.limits() == (None, 1e-10)
    c = sel({"maxiter": 10, "tol": 1e-10})
    assert isinstance(c, pkg.StopWhenAny) and c.limits() == (10, 1e-10)
    explicit = pkg.StopAfterIteration(5) | pkg.StopWhenConverged(1e-3)
    assert sel(explicit) is explicit
    with pytest.raises(pkg.ArgumentError, match="must be specified"):
        sel(None)
    with pytest.raises(pkg.ArgumentError, match="Unrecognized"):
        sel({"maxiter": 3, "bogus": 1})
    with pytest.raises(pkg.ArgumentError, match="At least one"):
        sel({})


def test_edge_inference_and_errors(pkg):
    i, j, k = pkg.Index(2), pkg.Index(3), pkg.Index(4)
    tn = pkg.ITensorNetwork({1: pkg.random_itensor([i, j]), 2: pkg.random_itensor([j, k]),
                             3: pkg.random_itensor([k])})
    assert set(map(frozenset, tn.edges())) == {frozenset((1, 2)), frozenset((2, 3))}
    assert tn.linkinds((1, 2)) == [j]
    assert tn.siteinds(1) == [i]
    with pytest.raises(ValueError, match="at most one existing tensor"):
        tn[4] = pkg.random_itensor([j])


def test_normnetwork_name_map(pkg):
    g = pkg.named_path_graph(3)
    l = {frozenset(e): pkg.Index(2) for e in g.edges()}
    s = {v: pkg.Index(2) for v in g.vertices()}
    tn = pkg.tensornetwork(lambda v: pkg.random_itensor(
        [s[v]] + [l[frozenset(e)] for e in g.incident_edges(v)]), g.vertices())
    nn = pkg.normnetwork(tn)
    for v in g.vertices():
        assert nn.bra_index(s[v]) == s[v]  # site names map to themselves
    for e in g.edges():
        li = l[frozenset(e)]
        assert nn.bra_index(li) != li and nn.bra_index(li).dim == li.dim
    b = nn.conj_bratensor(2)
    assert b.inds[0] == s[2] and all(x != y for x, y in zip(b.inds[1:], tn[2].inds[1:]))


def test_flattening_preserves_norm(pkg, orc):
    """to_topology permutes named tensors to the port layout; <psi|psi> is unchanged (the
    reference's test_normnetwork.jl:148-166 pins <psi|psi> = ||prod||^2)."""
    g = pkg.named_grid((2, 3))
    rng = np.random.default_rng(3)
    l = {frozenset(e): pkg.Index(2) for e in g.edges()}
    s = {v: pkg.Index(2) for v in g.vertices()}
    # deliberately scrambled index order inside each tensor
    def make(v):
        inds = [s[v]] + [l[frozenset(e)] for e in g.incident_edges(v)]
        rng.shuffle(inds)
        return pkg.random_itensor(inds, rng)
    tn = pkg.tensornetwork(make, g.vertices())
    fl = pkg.to_topology(pkg.normnetwork(tn))
    # ||psi||^2 via a dense contraction of the named tensors
    letters = iter("abcdefghijklmnopqrstuvwxyz")
    name = {}
    subs, ops = [], []
    for v in g.vertices():
        t = tn[v]
        subs.append("".join(name.setdefault(i, next(letters)) for i in t.inds))
        ops.append(t.data)
    psi = np.einsum(",".join(subs) + "->" + "".join(name[s[v]] for v in g.vertices()), *ops)
    assert orc.norm_exact(fl.topo, fl.sites) == pytest.approx(float(np.sum(psi ** 2)), rel=1e-12)


def test_messagecache_basics(pkg):
    g = pkg.named_grid((3, 3))
    c = pkg.messagecache((e, np.full((2, 2), k)) for k, e in enumerate(g.all_edges()))
    assert len(c) == 2 * g.ne()
    inc = pkg.incoming_messages(c, ((2, 2), (2, 3)))
    assert len(inc) == g.degree((2, 2)) - 1
    assert not any(np.array_equal(m, c[((2, 3), (2, 2))]) for m in inc)
    region = [(1, 1), (1, 2)]
    assert len(pkg.incoming_edge_data(c, region)) == g.degree((1, 1)) + g.degree((1, 2)) - 2
    t = pkg.build_topology(g, 2, 2)
    back = pkg.MessageCache.from_flat(t, c.to_flat(t))
    for e in c.keys():
        np.testing.assert_array_equal(back[e], c[e])


def test_message_environment_identity(pkg):
    g = pkg.named_grid((2, 2))
    l = {frozenset(e): pkg.Index(3) for e in g.edges()}
    s = {v: pkg.Index(2) for v in g.vertices()}
    tn = pkg.tensornetwork(lambda v: pkg.random_itensor(
        [s[v]] + [l[frozenset(e)] for e in g.incident_edges(v)]), g.vertices())
    env = pkg.message_environment(pkg.one, pkg.normnetwork(tn))
    assert len(env) == 2 * g.ne()
    for m in env.values():
        np.testing.assert_array_equal(m, np.eye(3))


def test_beliefpropagation_requires_stopping_criterion(pkg):
    g = pkg.named_path_graph(2)
    i = pkg.Index(2)
    tn = pkg.ITensorNetwork({1: pkg.random_itensor([i]), 2: pkg.random_itensor([i])})
    with pytest.raises(pkg.ArgumentError, match="must be specified"):
        pkg.beliefpropagation(tn, {(1, 2): np.ones(2), (2, 1): np.ones(2)})
