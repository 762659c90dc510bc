"""Gates on a device-resident network (bp_apply_gates_resident via DeviceNetwork) against the
marshalled path (apply_operators, itself checked against exact contraction and the global
SVD in test_gpu_apply.py), with the oracle's exact contraction comparing states:

  * one layer from the same sites and messages: the same singular values and the same state
    (the algorithm runs on the same inputs; only the bond index object differs);
  * a simple-update loop (BP, layer, BP, layer) kept on the device equals the same loop
    through ITensorNetwork / MessageCache round trips;
  * bonds wider than the kept rank are zero-padded and represent the same state;
  * one-site gates with normalize; an FP32 BP context; overlapping gates rejected.

Tolerances: 1e-10 relative for one layer from the same inputs (the two paths may order a
vertex's ports differently, which reorders the Jacobi sweeps), 1e-9 after BP runs
(converged to 1e-13), 1e-5 for FP32 messages.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _ones(m):
    return np.ones_like(m)


def _net(pkg, g, D, d, seed):
    rng = np.random.default_rng(seed)
    links = {frozenset(e): pkg.Index(D) for e in g.edges()}
    site = {v: pkg.Index(d) for v in g.vertices()}
    tn = pkg.tensornetwork(lambda v: pkg.random_itensor(
        [site[v]] + [links[frozenset(e)] for e in g.incident_edges(v)], rng), g.vertices())
    return tn, site, rng


def _gate(pkg, rng, dom):
    cod = [pkg.uniquename(i) for i in dom]
    shape = [i.dim for i in cod] + [i.dim for i in dom]
    return pkg.operator(rng.standard_normal(shape), cod, dom)


def _bp(pkg, tn):
    nn = pkg.normnetwork(tn)
    return pkg.beliefpropagation(nn, pkg.message_environment(_ones, nn),
                                 stopping_criterion={"maxiter": 300, "tol": 1e-13})


def _state(orc, tn, order):
    vs = tn.vertices()
    val, labs = orc.contract_labeled([tn[v].data for v in vs], [tn[v].inds for v in vs])
    return np.transpose(val, [labs.index(i) for i in order])


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _layer(pkg, rng, g, site, horizontal=True):
    L = max(v[0] for v in g.vertices())
    pairs = [((i, j), (i, j + 1)) for i in range(1, L + 1) for j in range(1, L, 2)] \
        if horizontal else [((i, j), (i + 1, j)) for j in range(1, L + 1) for i in range(1, L, 2)]
    return [_gate(pkg, rng, [site[a], site[b]]) for a, b in pairs], pairs


def test_resident_layer_matches_marshalled(pkg, orc):
    _say("case resident layer vs marshalled")
    g = pkg.named_grid((4, 4))
    tn, site, rng = _net(pkg, g, 3, 2, 31)
    env = _bp(pkg, tn)
    gates, pairs = _layer(pkg, rng, g, site)
    want_tn, want_env = pkg.apply_operators(gates, tn, env, trunc=pkg.truncrank(3))
    with pkg.DeviceNetwork(tn, env, dtype="f64") as dn:
        S = dn.apply_operators(gates, trunc=pkg.truncrank(3))
        got_tn, got_env = dn.itensornetwork(), dn.messagecache()
    order = [site[v] for v in g.vertices()]
    err = _rel(_state(orc, got_tn, order), _state(orc, want_tn, order))
    for (a, b), s in zip(pairs, S):
        np.testing.assert_allclose(s, np.diag(want_env[(a, b)]), rtol=1e-10, atol=0)
        np.testing.assert_allclose(got_env[(a, b)], want_env[(a, b)], rtol=1e-10, atol=1e-300)
    _say(f"state rel err {err:.2e}")
    assert err < 1e-10


def test_resident_simple_update_loop(pkg, orc):
    """BP -> layer -> BP -> layer on the device, against the same loop through host
    containers."""
    _say("case resident simple-update loop")
    g = pkg.named_grid((3, 3))
    tn, site, rng = _net(pkg, g, 2, 2, 32)
    l1, _ = _layer(pkg, rng, g, site, True)
    l2, _ = _layer(pkg, rng, g, site, False)
    # host loop: flatten, BP, download, gates, rebuild the ITensorNetwork, every round.  Both
    # loops converge Jacobi BP to max |dm| <= 1e-14 (a truncated gate depends on the
    # messages; a fidelity stop would leave them correct to only ~sqrt(tol))
    t_h = tn
    for layer in (l1, l2):
        fl = pkg.to_topology(pkg.normnetwork(t_h))
        with pkg.Engine(fl.topo, fl.sites, dtype="f64") as e:
            e.init_messages("ones")
            assert e.run(maxiter=3000, tol=1e-14, criterion="maxabs").converged[0]
            env = pkg.MessageCache.from_flat(fl.topo, e.messages())
        t_h, _ = pkg.apply_operators(layer, t_h, env, trunc=pkg.truncrank(2))
    # device loop: nothing leaves the device until the end
    with pkg.DeviceNetwork(tn, dtype="f64") as dn:
        for layer in (l1, l2):
            dn.engine.init_messages("ones")
            r = dn.beliefpropagation(maxiter=3000, tol=1e-14, criterion="maxabs")
            assert r.converged[0]
            dn.apply_operators(layer, trunc=pkg.truncrank(2))
        t_d = dn.itensornetwork()
    order = [site[v] for v in g.vertices()]
    err = _rel(_state(orc, t_d, order), _state(orc, t_h, order))
    _say(f"state rel err {err:.2e}")
    assert err < 1e-9


def test_resident_padding(pkg, orc):
    """A 4-path with bond dimension 4: the end bonds can hold rank 2 only, so the resident
    tensors are zero-padded; the state equals the marshalled one (whose bonds shrink)."""
    _say("case resident padding")
    g = pkg.named_path_graph(4)
    tn, site, rng = _net(pkg, g, 4, 2, 33)
    env = _bp(pkg, tn)
    gate = _gate(pkg, rng, [site[1], site[2]])
    want_tn, want_env = pkg.apply_operator(gate, tn, env)
    assert np.diag(want_env[(1, 2)]).size == 2
    with pkg.DeviceNetwork(tn, env, dtype="f64") as dn:
        S = dn.apply_operator(gate)
        got = dn.itensornetwork()
        m = dn.messagecache()[(1, 2)]
    assert S.size == 2 and got[1].data.shape[-1] == 4
    np.testing.assert_array_equal(m[2:, :], 0.0)
    order = [site[v] for v in g.vertices()]
    err = _rel(_state(orc, got, order), _state(orc, want_tn, order))
    _say(f"state rel err {err:.2e}")
    assert err < 1e-10


def test_resident_one_site_and_fp32(pkg, orc):
    _say("case resident one-site normalize and FP32 context")
    g = pkg.named_grid((3, 3))
    tn, site, rng = _net(pkg, g, 2, 3, 34)
    env = _bp(pkg, tn)
    o1 = _gate(pkg, rng, [site[(2, 2)]])
    o2 = _gate(pkg, rng, [site[(1, 1)], site[(1, 2)]])
    # a resident bond keeps its dimension (2): compare with the marshalled gate truncated to it
    want_tn, _ = pkg.apply_operators([o1, o2], tn, env, normalize=True, trunc=pkg.truncrank(2))
    order = [site[v] for v in g.vertices()]
    want = _state(orc, want_tn, order)
    with pkg.DeviceNetwork(tn, env, dtype="f64") as dn:
        dn.apply_operators([o1, o2], normalize=True)
        err64 = _rel(_state(orc, dn.itensornetwork(), order), want)
    # FP32 messages (converted on the fly); sites stay FP64
    with pkg.DeviceNetwork(tn, env, dtype="f32") as dn:
        dn.apply_operators([o1, o2], normalize=True)
        err32 = _rel(_state(orc, dn.itensornetwork(), order), want)
    _say(f"FP64 ctx rel err {err64:.2e}, FP32 ctx rel err {err32:.2e}")
    assert err64 < 1e-10
    assert err32 < 1e-5


def test_resident_rejects_overlapping_gates(pkg):
    _say("case resident overlap rejected")
    g = pkg.named_path_graph(3)
    tn, site, rng = _net(pkg, g, 2, 2, 35)
    with pkg.DeviceNetwork(tn, dtype="f64") as dn:
        t = dn.topo
        q12 = dn._slot(0, 1)
        q23 = dn._slot(1, 2)
        G = np.eye(4)
        with pytest.raises(pkg._lib.BPError, match="earlier gate"):
            dn.engine.apply_gates([2, 2], [q12, q23], [G, G])


def test_resident_layer_against_oracle(pkg, orc):
    """One resident layer checked directly against the oracle: every gate against
    orc.apply_gate2 on the same site tensors and converged messages (the two-site product
    and the singular values), then BP on the gated network against orc.bp on the oracle's
    gated sites (the log-norm, which does not depend on the bond gauge of the factors)."""
    _say("case resident layer vs oracle apply_gate2 + bp")
    g = pkg.named_grid((4, 4))
    D = 3
    tn, site, rng = _net(pkg, g, D, 2, 33)
    gates, pairs = _layer(pkg, rng, g, site)
    with pkg.DeviceNetwork(tn, dtype="f64") as dn:
        assert dn.beliefpropagation(maxiter=3000, tol=1e-14, criterion="maxabs").converged[0]
        t = dn.topo
        sites0, msgs = dn.engine.sites(), dn.engine.messages()

        def tensor(sites, u):
            dims = [int(t.phys_dim[u])] + [int(t.edge_dim[t.port_edge[q]])
                                          for q in range(t.port_off[u], t.port_off[u + 1])]
            return sites[t.site_off[u]:t.site_off[u + 1]].reshape(dims)

        def incoming(u, w):
            out = {}
            for j, q in enumerate(range(t.port_off[u], t.port_off[u + 1])):
                if int(t.port_nbr[q]) == w:
                    continue
                sl = int(t.port_rev[q])
                Dq = int(t.edge_dim[t.port_edge[q]])
                out[j] = msgs[t.msg_off[sl]:t.msg_off[sl + 1]].reshape(Dq, Dq)
            return out

        def port(u, w):
            return [int(t.port_nbr[q]) for q in range(t.port_off[u], t.port_off[u + 1])].index(w)

        want = {}
        osites = sites0.copy()
        for op in gates:
            kind, slot, mat, vs = dn._gate(op)
            u1, u2 = dn._vidx[vs[0]], dn._vidx[vs[1]]
            p1, p2 = port(u1, u2), port(u2, u1)
            B1, B2, S = orc.apply_gate2(tensor(sites0, u1), tensor(sites0, u2), p1, p2,
                                        incoming(u1, u2), incoming(u2, u1), mat, maxdim=D)
            assert len(S) == D
            want[(u1, u2)] = (p1, p2, B1, B2, S)
            osites[t.site_off[u1]:t.site_off[u1 + 1]] = B1.ravel()
            osites[t.site_off[u2]:t.site_off[u2 + 1]] = B2.ravel()
        S_dev = dn.apply_operators(gates, trunc=pkg.truncrank(D))
        sites1 = dn.engine.sites()
        for ((u1, u2), (p1, p2, B1, B2, S)), s_dev in zip(want.items(), S_dev):
            np.testing.assert_allclose(s_dev, S, rtol=1e-10, atol=0)
            got = np.tensordot(tensor(sites1, u1), tensor(sites1, u2), axes=([1 + p1], [1 + p2]))
            exp = np.tensordot(B1, B2, axes=([1 + p1], [1 + p2]))
            assert _rel(got, exp) < 1e-10
        dn.engine.init_messages("identity")
        assert dn.beliefpropagation(maxiter=3000, tol=1e-14, criterion="maxabs").converged[0]
        ln_dev = dn.engine.lognorm()
    m, sw, dl, at = orc.bp(t, osites, pkg.peps.identity_messages(t), maxiter=3000, tol=1e-14,
                           criterion="maxabs")
    assert at > 0
    ln = orc.bethe_free_energy(t, osites, m)
    _say(f"gated log-norm device={ln_dev:.12f} oracle={ln:.12f}")
    assert abs(ln_dev - ln) < 1e-10 * max(1.0, abs(ln))
