"""Reference semantics on the device, through the C ABI: loop counts of
StopAfterIteration | StopWhenConverged, iterate_diff closed forms, the edge cases of
SimpleMessageUpdate and bethe_free_energy, per-network masks with a poisoned network, the
nested solver with a user strategy, and SimpleMessageUpdate(normalize = false).

Reference lines: AlgorithmsInterfaceExtensions.jl:84-109 (stopping), beliefpropagation.jl:
237-267 (update, zero-sum skip, iterate_diff), messagecache.jl:185-201 (complex promotion,
-Inf).  The loop counts are fixed by the path length (exact after N - 1 Jacobi sweeps).
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("N", [2, 4, 7])
@pytest.mark.parametrize("dtype,path", [("f64", "generic"), ("f32", "auto")])
def test_path_loop_counts(pkg, orc, N, dtype, path):
    _say(f"case path loop counts N={N} {dtype} {path}")
    t = pkg.build_topology(pkg.named_path_graph(N), 3, 2)
    s = pkg.peps.random_sites(t, 5)
    m0 = pkg.peps.ones_messages(t)
    with pkg.Engine(t, s, dtype=dtype, path=path) as e:
        e.set_messages(m0)
        r = e.run(maxiter=N - 1)
        assert (r.sweeps[0], r.at_iteration[0]) == (N - 1, -1)
        e.set_messages(m0)
        r = e.run(maxiter=100, tol=1e-300, criterion="maxabs")
        # sweep N repeats sweep N - 1 bit for bit: delta exactly 0 at iteration N
        assert (r.sweeps[0], r.at_iteration[0], r.delta[0]) == (N, N, 0.0)
        got = e.messages()
        e.set_messages(m0)
        r = e.run(tol=1e300)  # every delta passes; iteration 0 is still skipped
        assert (r.sweeps[0], r.at_iteration[0]) == (1, 1)
    want, _, _, _ = orc.bp(t, s, m0, maxiter=N)
    tol = 1e-13 if dtype == "f64" else 1e-5
    assert np.max(np.abs(got - want)) / np.max(np.abs(want)) < tol


@pytest.mark.parametrize("a,b,want", [
    (np.eye(2), np.diag([1.0, 0.0]), 0.5),
    (np.array([[1.0, 2.0], [3.0, 4.0]]), -np.array([[1.0, 2.0], [3.0, 4.0]]), 0.0),
    (np.diag([1.0, 0.0]), np.diag([0.0, 1.0]), 1.0),
])
def test_iterate_diff_closed_forms_on_device(pkg, a, b, want):
    _say(f"case iterate_diff closed form {want}")
    t = pkg.build_topology(pkg.named_path_graph(2), 2, 2)
    s = pkg.peps.random_sites(t, 1)
    # slot 0 carries the case, slot 1 identical messages (distance 0): the max is `want`
    mine = np.concatenate([a.ravel(), np.eye(2).ravel()])
    other = np.concatenate([b.ravel(), np.eye(2).ravel()])
    with pkg.Engine(t, s, dtype="f64") as e:
        e.set_messages(mine)
        assert e.iterate_diff(other) == pytest.approx(want, abs=1e-15)


@pytest.mark.parametrize("D,dtype,path", [(2, "f64", "generic"), (8, "f32", "vertex"),
                                          (8, "f64", "generic"), (8, "bf16", "vertex")])
def test_zero_message_is_not_normalised(pkg, orc, D, dtype, path):
    """A vertex whose tensor vanishes sends all-zero messages: their sum is exactly zero in
    any arithmetic, the division is skipped (beliefpropagation.jl:249-252) and the messages
    stay 0 instead of 0/0; Z_v = 0 makes the log-norm -Inf."""
    _say(f"case zero message D={D} {dtype} {path}")
    t = pkg.build_topology(pkg.named_grid((3, 3)), D, 2)
    u = t.vertices.index((2, 2))
    s = pkg.peps.random_sites(t, 3, scale=pkg.peps.peps_scale(2, D))
    s[t.site_off[u]:t.site_off[u + 1]] = 0.0
    m0 = pkg.peps.identity_messages(t)
    want, _, _ = orc.sweep_jacobi(t, s, m0)
    zero = np.zeros(want.size, dtype=bool)
    for q in range(int(t.port_off[u]), int(t.port_off[u + 1])):
        zero[int(t.msg_off[q]):int(t.msg_off[q + 1])] = True
    assert np.all(want[zero] == 0.0)  # the oracle's own skip
    assert orc.bethe_free_energy(t, s, want) == -math.inf
    with pkg.Engine(t, s, dtype=dtype, path=path) as e:
        e.set_messages(m0)
        e.sweep()
        got = e.messages()
        ln = e.lognorm()
    assert np.all(got[zero] == 0.0)
    tol = {"f64": 1e-13, "f32": 1e-5, "bf16": 5e-2}[dtype]
    err = np.max(np.abs(got[~zero] - want[~zero])) / np.max(np.abs(want[~zero]))
    _say(f"other slots rel err {err:.2e}")
    assert err < tol
    assert ln == -math.inf


def test_lognorm_minus_inf_on_a_zero_edge_term(pkg, orc):
    _say("case lognorm -inf")
    t = pkg.build_topology(pkg.named_grid((2, 2)), 2, 2)
    s = pkg.peps.random_sites(t, 2)
    m = pkg.peps.identity_messages(t)
    q = 0
    r = int(t.port_rev[q])
    m[t.msg_off[q]:t.msg_off[q + 1]] = np.diag([1.0, 0.0]).ravel()
    m[t.msg_off[r]:t.msg_off[r + 1]] = np.diag([0.0, 1.0]).ravel()
    assert orc.edge_scalars(t, m)[t.port_edge[q]] == 0.0
    assert orc.bethe_free_energy(t, s, m) == -math.inf
    with pkg.Engine(t, s, dtype="f64") as e:
        e.set_messages(m)
        assert e.lognorm() == -math.inf


def test_lognorm_complex_promotion(pkg, orc):
    """Negative vertex or edge terms make the log-norm complex: imaginary part
    pi * (#negative Z_v - #negative Z_e), real part sum log|Z_v| - sum log|Z_e|."""
    _say("case lognorm complex promotion")
    t = pkg.build_topology(pkg.named_grid((3, 3)), 2, 2)
    s = pkg.peps.random_sites(t, 4)
    for seed in range(40):
        m = pkg.peps.random_messages(t, seed) - 0.5  # entries of both signs
        zv = orc.vertex_scalars(t, s, m)
        ze = orc.edge_scalars(t, m)
        nv, ne = int(np.sum(zv < 0)), int(np.sum(ze < 0))
        if nv != ne and nv > 0 and ne > 0:
            break
    else:
        pytest.fail("no seed with negative vertex and edge terms")
    want = complex(np.sum(np.log(np.abs(zv))) - np.sum(np.log(np.abs(ze))), math.pi * (nv - ne))
    assert orc.bethe_free_energy(t, s, m) == pytest.approx(want, rel=1e-13)
    with pkg.Engine(t, s, dtype="f64") as e:
        e.set_messages(m)
        got = e.lognorm()
    assert isinstance(got, complex)
    assert got.real == pytest.approx(want.real, rel=1e-12)
    assert got.imag == pytest.approx(want.imag, rel=1e-15)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_nan_network_does_not_touch_the_others(pkg, dtype):
    """Per-network masks: a network whose sites hold a NaN never converges (a NaN
    residual stops nothing) and leaves the other networks' results bit for bit as in
    their own runs."""
    _say(f"case nan network in batch {dtype}")
    tops = [pkg.build_topology(pkg.named_grid((L, L)), 8, 2) for L in (3, 4, 3)]
    sites = [pkg.peps.random_sites(t, 60 + i, scale=pkg.peps.peps_scale(2, 8))
             for i, t in enumerate(tops)]
    bad = sites[1].copy()
    bad[5] = np.nan
    b = pkg.concat_topologies(tops)
    with pkg.Engine(b, np.concatenate([sites[0], bad, sites[2]]), dtype=dtype) as e:
        e.init_messages("identity")
        r = e.run(maxiter=60, tol=1e-10)
        mb = e.messages()
    assert r.at_iteration[1] == -1 and r.sweeps[1] == 60
    off = 0
    for i, (t, s) in enumerate(zip(tops, sites)):
        n = int(t.msg_off[-1])
        if i != 1:
            with pkg.Engine(t, s, dtype=dtype) as e:
                e.init_messages("identity")
                ri = e.run(maxiter=60, tol=1e-10)
                mi = e.messages()
            assert ri.at_iteration[0] > 0
            assert (r.sweeps[i], r.at_iteration[i]) == (ri.sweeps[0], ri.at_iteration[0])
            np.testing.assert_array_equal(mb[off:off + n], mi)
        else:
            assert np.isnan(mb[off:off + n]).any()
        off += n


def test_unnormalised_updates_match_oracle(pkg, orc):
    """SimpleMessageUpdate(normalize = false): sequential and Jacobi sweeps keep the raw
    contraction (beliefpropagation.jl:248-253 skipped)."""
    _say("case normalize=false")
    g = pkg.named_grid((3, 3))
    t = pkg.build_topology(g, 2, 2)
    s = pkg.peps.random_sites(t, 8)
    m0 = pkg.peps.random_messages(t, 8)
    order = [t.slot(a, b) for a, b in pkg.forest_cover_edge_sequence(g)]
    want_seq = orc.sweep_sequential(t, s, m0, order, normalize=False)
    want_jac, _, _ = orc.sweep_jacobi(t, s, m0, normalize=False)
    with pkg.Engine(t, s, dtype="f64") as e:
        e.set_option("normalize", 0)
        assert e.get_option("normalize") == 0
        assert e.path() == "generic"
        e.set_messages(m0)
        e.sweep()
        np.testing.assert_allclose(e.messages(), want_jac, rtol=1e-12, atol=1e-14)
        e.set_messages(m0)
        e.run_sequential(order, maxiter=1)
        np.testing.assert_allclose(e.messages(), want_seq, rtol=1e-12, atol=1e-14)
        with pytest.raises(pkg._lib.BPError):
            e.grad_lognorm()
    # the API strategy
    nn = _norm_network(pkg, g, t, s)
    fl = pkg.to_topology(nn)
    cache = pkg.MessageCache.from_flat(fl.topo, m0)
    out = pkg.beliefpropagation(nn, cache, stopping_criterion={"maxiter": 1},
                                message_update_algorithm=pkg.SimpleMessageUpdate(normalize=False))
    want = orc.sweep_sequential(fl.topo, fl.sites, m0,
                                [fl.slot(ed) for ed in pkg.default_beliefpropagation_edges(nn)],
                                normalize=False)
    np.testing.assert_allclose(out.to_flat(fl.topo), want, rtol=1e-12, atol=1e-14)


def _norm_network(pkg, g, t, s):
    """A NormNetwork whose flattening is exactly topology t with sites s."""
    l = {frozenset(e): pkg.Index(int(t.edge_dim[t.port_edge[t.slot(*e)]])) for e in g.edges()}
    si = {v: pkg.Index(2) for v in g.vertices()}
    tensors = {}
    for ui, v in enumerate(t.vertices):
        inds = [si[v]] + [l[frozenset((v, t.vertices[t.port_nbr[q]]))]
                          for q in range(t.port_off[ui], t.port_off[ui + 1])]
        A = s[t.site_off[ui]:t.site_off[ui + 1]].reshape([i.dim for i in inds])
        tensors[v] = pkg.ITensor(A, inds)
    return pkg.NormNetwork(pkg.ITensorNetwork(tensors))


def test_nested_solver_with_user_strategy_and_sweep_token(pkg, orc):
    """The reference's nested loop with a user MessageUpdateAlgorithm (one device update
    per edge, O(degree D^2) transfer) and with whole device sweeps behind JacobiSweep
    entries: both reach the oracle's fixed point and report the oracle's loop counts."""
    _say("case nested solver")
    g = pkg.named_grid((3, 3))
    t = pkg.build_topology(g, 2, 2)
    s = pkg.peps.random_sites(t, 9, scale=pkg.peps.peps_scale(2, 2))
    nn = _norm_network(pkg, g, t, s)
    fl = pkg.to_topology(nn)
    m0 = pkg.peps.identity_messages(fl.topo)
    cache = pkg.MessageCache.from_flat(fl.topo, m0)

    class Counting(pkg.MessageUpdateAlgorithm):
        def __init__(self):
            self.calls = 0
            self.inner = pkg.SimpleMessageUpdate()

        def update(self, c, factors, edge):
            self.calls += 1
            return self.inner.update(c, factors, edge)

    alg = Counting()
    out = pkg.beliefpropagation(nn, cache, stopping_criterion={"maxiter": 50, "tol": 1e-12},
                                message_update_algorithm=alg)
    order = [fl.slot(ed) for ed in pkg.default_beliefpropagation_edges(nn)]
    want, sw, dl, at = orc.bp(fl.topo, fl.sites, m0, maxiter=50, tol=1e-12,
                              schedule="sequential", order=order)
    assert (out.state.iteration, out.state.at_iteration) == (sw, at)
    assert alg.calls == sw * len(order)
    np.testing.assert_allclose(out.to_flat(fl.topo), want, rtol=0, atol=1e-12)
    # whole device sweeps through the same nested loop
    out = pkg.beliefpropagation(nn, cache, edges=[pkg.JACOBI_SWEEP],
                                stopping_criterion={"maxiter": 50, "tol": 1e-12},
                                message_update_algorithm=pkg.JacobiGPUUpdate(dtype="f64"))
    want, sw, dl, at = orc.bp(fl.topo, fl.sites, m0, maxiter=50, tol=1e-12)
    assert (out.state.iteration, out.state.at_iteration) == (sw, at)
    np.testing.assert_allclose(out.to_flat(fl.topo), want, rtol=0, atol=1e-12)
