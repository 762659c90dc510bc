"""Exact contraction on the device (bp_contract_network / bp_cn_*, the contract_network
counterpart) against the FP64 oracle's labelled frontier contraction, whose restatement is
pinned in test_oracle.py (ABCD, order independence, Ising closed forms, <psi|psi> by
numpy einsum).

Tolerances: 1e3 * eps relative between orders and against the oracle for closed networks
(the reference's own bound, test_contract_network.jl:40-41); 1e-12 relative to the largest
entry for open results and random networks; the ABCD case is exact (every product and sum
is representable).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

EPS = np.finfo(float).eps


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    a, b = np.asarray(a, dtype=float), np.asarray(b, dtype=float)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


def _align(val, labs, want_labs):
    """Permute a result with labels `labs` to the label order `want_labs`."""
    if not want_labs:
        return val
    assert sorted(map(repr, labs)) == sorted(map(repr, want_labs))
    return np.transpose(val, [labs.index(l) for l in want_labs])


def test_contract_abcd(pkg):
    """test_contract_network.jl:13-24: every order gives 15.5 * D exactly."""
    _say("case contract ABCD")
    i, j, k = pkg.Index(2), pkg.Index(2), pkg.Index(5)
    A = pkg.ITensor(np.array([[1.0, 1.0], [0.5, 1.0]]), [i, j])
    B = pkg.ITensor(np.array([2.0, 1.0]), [i])
    Cv = pkg.ITensor(np.array([5.0, 1.0]), [j])
    Dv = np.array([-2.0, 3.0, 4.0, 5.0, 1.0])
    D = pkg.ITensor(Dv, [k])
    for alg in (pkg.LeftAssociative(), pkg.Greedy(), pkg.Cheapest(), pkg.Flat()):
        r = pkg.contract_network([A, B, Cv, D], alg=pkg.Exact(order_alg=alg))
        assert list(r.inds) == [k]
        np.testing.assert_array_equal(r.data, 15.5 * Dv)
    r = pkg.contract_network([A, B, Cv, D], alg=pkg.Exact(order=((3, 2), (0, 1))))
    np.testing.assert_array_equal(r.data, 15.5 * Dv)


def test_contract_grid_orders(pkg, orc):
    """test_contract_network.jl:26-45: a random 4x4 grid network contracts to the same
    scalar in every order, within 1e3 eps, and equals the oracle's."""
    _say("case contract 4x4 grid orders")
    g = pkg.named_grid((4, 4))
    links = {frozenset(e): pkg.Index(2) for e in g.edges()}
    rng = np.random.default_rng(3)
    tn = pkg.tensornetwork(lambda v: pkg.random_itensor(
        [links[frozenset(e)] for e in g.incident_edges(v)], rng), g.vertices())
    zs = [pkg.contract_network(tn, alg=pkg.Exact(order_alg=a))[()]
          for a in (pkg.LeftAssociative(), pkg.Greedy(), pkg.Cheapest())]
    want, _ = orc.contract_labeled([tn[v].data for v in tn.vertices()],
                                   [tn[v].inds for v in tn.vertices()])
    _say(f"z = {zs}, oracle {want}")
    for z in zs:
        assert abs(z - zs[0]) / abs(zs[0]) <= 1e3 * EPS
        assert abs(z - want) / abs(want) <= 1e3 * EPS


def _random_network(rng, nt, nl, maxd):
    labels = [[] for _ in range(nt)]
    dims = {}
    for l in range(nl):
        dims[l] = int(rng.integers(1, maxd + 1))
        k = 2 if (nt > 1 and rng.random() < 0.7) else 1
        for t in rng.choice(nt, size=k, replace=False):
            labels[t].append(l)
    for ls in labels:
        rng.shuffle(ls)
    tensors = [rng.standard_normal(tuple(dims[l] for l in ls)) for ls in labels]
    return tensors, [tuple(ls) for ls in labels]


def _random_tree(rng, nt):
    ids, pairs, nxt = list(range(nt)), [], nt
    while len(ids) > 1:
        i, j = rng.choice(len(ids), 2, replace=False)
        a, b = ids[i], ids[j]
        pairs.append((a, b))
        ids = [x for x in ids if x not in (a, b)] + [nxt]
        nxt += 1
    return pairs


def test_contract_random_networks(pkg, orc):
    """Random networks (open and summed labels, disconnected parts, scalars, shuffled label
    orders, transposes that need gathers): every order against the oracle; the
    left-associative order also reproduces the oracle's label order."""
    _say("case contract random networks")
    rng = np.random.default_rng(11)
    worst = 0.0
    for trial in range(60):
        nt = int(rng.integers(1, 7))
        tensors, labels = _random_network(rng, nt, int(rng.integers(0, 9)), 4)
        want, wl = orc.contract_labeled(tensors, labels)
        for alg in ("left_associative", "greedy", "cheapest"):
            got, gl = pkg.contract_labeled(tensors, labels, order_alg=alg)
            if alg == "left_associative":
                assert gl == wl
            worst = max(worst, _rel(_align(got, gl, wl), want))
        if nt >= 2:
            got, gl = pkg.contract_labeled(tensors, labels, order=_random_tree(rng, nt))
            worst = max(worst, _rel(_align(got, gl, wl), want))
    _say(f"worst relative error {worst:.2e}")
    assert worst < 1e-12


def test_contract_large_shapes(pkg, orc):
    """Products large enough for many tiles, split K (long shared dimension with few output
    tiles), a dot product, and a transpose of both operands."""
    _say("case contract large shapes")
    rng = np.random.default_rng(12)
    cases = [
        ([rng.standard_normal((3, 4000)), rng.standard_normal((4000, 2))],
         [("i", "k"), ("k", "j")]),
        ([rng.standard_normal(200000), rng.standard_normal(200000)], [("k",), ("k",)]),
        ([rng.standard_normal((130, 70, 5)), rng.standard_normal((5, 70, 90))],
         [("a", "b", "c"), ("c", "b", "d")]),
        ([rng.standard_normal((6, 7, 8, 9)), rng.standard_normal((8, 6, 10, 9))],
         [("a", "b", "c", "d"), ("c", "a", "e", "d")]),
        ([rng.standard_normal((300, 300)), rng.standard_normal((300, 300)),
          rng.standard_normal((300, 300))], [("i", "j"), ("j", "k"), ("k", "l")]),
    ]
    for tensors, labels in cases:
        want, wl = orc.contract_labeled(tensors, labels)
        got, gl = pkg.contract_labeled(tensors, labels, order_alg="left_associative")
        err = _rel(got, want)
        _say(f"shapes {[t.shape for t in tensors]} err {err:.2e}")
        assert gl == wl and err < 1e-12


def _ising_itensors(pkg, g, beta):
    """ising_network (src/ITensorNetworkGenerators/ising_network.jl:7-51, h = 0): a delta
    tensor per vertex with sqrt of the bond Boltzmann matrix on every leg."""
    M = np.array([[math.exp(beta), math.exp(-beta)], [math.exp(-beta), math.exp(beta)]])
    w, V = np.linalg.eigh(M)
    S = V @ np.diag(np.sqrt(w)) @ V.T
    links = {frozenset(e): pkg.Index(2) for e in g.edges()}

    def site(v):
        inds = [links[frozenset(e)] for e in g.incident_edges(v)]
        k = len(inds)
        T = np.zeros((2,) * k)
        for s in range(2):
            T[(s,) * k] = 1.0
        for j in range(k):
            T = np.moveaxis(np.tensordot(T, S, axes=([j], [0])), -1, j)
        return pkg.ITensor(T, inds)

    return pkg.tensornetwork(site, g.vertices())


@pytest.mark.parametrize("periodic", [False, True])
def test_contract_ising_1d(pkg, periodic):
    """test_itensornetworkgenerators.jl:27-53 with the closed forms of test/utils.jl:12-39."""
    _say(f"case contract 1D Ising periodic={periodic}")
    n, beta = 4, 0.4
    g = pkg.named_path_graph(n)
    if periodic:
        g.add_edge(n, 1)
    z = pkg.contract_network(_ising_itensors(pkg, g, beta))[()]
    if periodic:
        want = (2 * math.cosh(beta)) ** n + (2 * math.sinh(beta)) ** n
    else:
        want = 2 * (2 * math.cosh(beta)) ** (n - 1)
    assert z == pytest.approx(want, rel=1e-13)


def test_contract_ising_2d(pkg, orc):
    """test_itensornetworkgenerators.jl:54-72: 4x4 periodic grid at the critical inverse
    temperature; free energy density within 10% of Onsager's infinite-volume value, and the
    partition function equal to the oracle's."""
    _say("case contract 2D Ising")
    from scipy.integrate import quad

    beta = math.log(1 + math.sqrt(2)) / 2
    g = pkg.named_grid((4, 4), periodic=True)
    tn = _ising_itensors(pkg, g, beta)
    z = pkg.contract_network(tn)[()]
    kappa = 2 * math.sinh(2 * beta) / math.cosh(2 * beta) ** 2
    integral, _ = quad(lambda th: math.log((1 + math.sqrt(abs(1 - (kappa * math.sin(th)) ** 2)))
                                           / 2), 0, math.pi)
    f_inf = (-math.log(2 * math.cosh(2 * beta)) - integral / (2 * math.pi)) / beta
    f = -math.log(z) / (beta * 16)
    want, _ = orc.contract_labeled([tn[v].data for v in tn.vertices()],
                                   [tn[v].inds for v in tn.vertices()])
    _say(f"Z = {z}, oracle {want}, f = {f}, f_inf = {f_inf}")
    assert f == pytest.approx(f_inf, rel=0.1)
    assert abs(z - want) / want <= 1e3 * EPS


def test_contract_norm_networks(pkg, orc):
    """<psi|psi> of NormNetworks: a normalised single tensor gives 1
    (test_normnetwork.jl:137-166), a renamed bra gives the same value, and 3x3 / 4x4
    PEPS norms equal the oracle's norm_exact on the flattened network."""
    _say("case contract norm networks")
    s = pkg.Index(3)
    v = np.random.default_rng(1).standard_normal(3)
    tn1 = pkg.ITensorNetwork({1: pkg.ITensor(v / np.linalg.norm(v), [s])})
    assert pkg.contract_network(pkg.normnetwork(tn1))[()] == pytest.approx(1.0, rel=1e-14)
    for dims, D in (((3, 3), 2), ((4, 4), 3)):
        g = pkg.named_grid(dims)
        links = {frozenset(e): pkg.Index(D) for e in g.edges()}
        sites = {u: pkg.Index(2) for u in g.vertices()}
        rng = np.random.default_rng(5)
        tn = pkg.tensornetwork(lambda u: pkg.random_itensor(
            [sites[u]] + [links[frozenset(e)] for e in g.incident_edges(u)], rng),
            g.vertices())
        nn = pkg.normnetwork(tn)
        z = pkg.contract_network(nn)[()]
        custom = {i: pkg.uniquename(i) for i in links.values()}
        z2 = pkg.contract_network(pkg.normnetwork(tn, custom))[()]
        fl = pkg.to_topology(nn)
        want = orc.norm_exact(fl.topo, fl.sites)
        _say(f"{dims} D={D}: z={z} oracle={want}")
        assert abs(z - want) / want <= 1e3 * EPS
        assert z2 == pytest.approx(z, rel=1e3 * EPS)


def test_contract_open_result_itensor(pkg, orc):
    """An open network: the ITensor result carries the open indices, in whatever order the
    tree leaves them, with the oracle's values."""
    _say("case contract open result")
    rng = np.random.default_rng(6)
    a, b, c, d, e = (pkg.Index(n) for n in (2, 3, 4, 5, 3))
    T1 = pkg.random_itensor([a, b, c], rng)
    T2 = pkg.random_itensor([c, d], rng)
    T3 = pkg.random_itensor([e, b, d], rng)
    r = pkg.contract_network([T1, T2, T3])
    want, wl = orc.contract_labeled([T1.data, T2.data, T3.data], [T1.inds, T2.inds, T3.inds])
    assert set(r.inds) == {a, e}
    assert _rel(r.permute(wl), want) < 1e-12


def test_contraction_plan_reuse_and_device(pkg, orc):
    """One plan on two inputs of the same shapes; the device-buffer execution equals the
    host-buffer one; the empty network is 1 and a single tensor is itself."""
    _say("case contract plan reuse")
    rng = np.random.default_rng(7)
    shapes = [(4, 5, 6), (6, 5, 7), (7, 3)]
    labels = [("a", "b", "c"), ("c", "b", "d"), ("d", "e")]
    with pkg.ContractionPlan(shapes, labels, order_alg="greedy") as p:
        for _ in range(2):
            ts = [rng.standard_normal(s) for s in shapes]
            want, wl = orc.contract_labeled(ts, labels)
            got = p(ts)
            assert _rel(_align(got, p.out_labels, wl), want) < 1e-12
        flat = np.concatenate([t.ravel() for t in ts])
        din = pkg._lib.DeviceArray(flat.size, np.float64)
        dout = pkg._lib.DeviceArray(p.out_size, np.float64)
        din.upload(flat)
        p.execute_device(din, dout)
        # the same launches in the same order: bitwise the host-buffer result
        np.testing.assert_array_equal(dout.numpy(), np.asarray(got).ravel())
        din.free()
        dout.free()
        assert p.last_ms > 0.0
    assert pkg.contract_labeled([], [])[0] == 1.0
    T = rng.standard_normal((3, 4))
    got, gl = pkg.contract_labeled([T], [("x", "y")])
    np.testing.assert_array_equal(got, T)
    assert gl == ["x", "y"]


def test_contract_errors(pkg):
    """Invalid networks and orders are rejected before any launch."""
    _say("case contract errors")
    BPError = pkg._lib.BPError
    x = np.ones((2, 2))
    with pytest.raises(BPError):  # a label on three tensors
        pkg.contract_labeled([x, x, x], [("i", "j"), ("i", "k"), ("i", "l")])
    with pytest.raises(BPError):  # a label twice on one tensor
        pkg.contract_labeled([x], [("i", "i")])
    with pytest.raises(BPError):  # an intermediate beyond max_elems
        pkg.contract_labeled([np.ones((50, 50)), np.ones((50, 50))], [("i", "j"), ("k", "l")],
                             max_elems=1000)
    with pytest.raises(BPError):  # an id used twice
        pkg.contract_labeled([x, x, x], [("i", "j"), ("j", "k"), ("k", "l")],
                             order=[(0, 1), (0, 2)])
    with pytest.raises(pkg.ArgumentError):
        pkg.contract_network([pkg.random_itensor([pkg.Index(2)])], alg="nope")
