"""BP expectation values and local energies with arbitrary operators on the device
(bp_expect_vertex / bp_expect_edge / bp_local_energy) against the FP64 oracle, whose
restatement is pinned by exact tree expectations, identity and linearity checks, the TFIM
energy and finite differences of the gradient (test_oracle.py).

Tolerances: expectation values, FP64 contexts 1e-11 relative to the oracle at the same
messages, FP32 2e-5 (FP32 site tensors and contractions); TFIM energy assembled from the
expectation values equals the engine's own within 1e-9 (FP64) / 1e-6 (FP32); local-energy
gradients within 1e-7 (FP64) / 2e-3 (FP32) of the oracle's largest entry.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

X = np.array([[0.0, 1.0], [1.0, 0.0]])
Z = np.array([[1.0, 0.0], [0.0, -1.0]])


def _say(msg):
    print(msg, flush=True)


def _converged(pkg, orc, L, D, d, seed, dtype):
    t = pkg.build_topology(pkg.named_grid((L, L)), D, d)
    s = pkg.peps.random_sites(t, seed, scale=pkg.peps.peps_scale(d, D))
    e = pkg.Engine(t, s, dtype=dtype)
    e.init_messages("identity")
    r = e.run(maxiter=500, tol=1e-12, criterion="fidelity")
    assert r.converged[0]
    return t, s, e


@pytest.mark.parametrize("dtype,d,tol", [("f64", 2, 1e-11), ("f64", 3, 1e-11), ("f32", 2, 2e-5)])
def test_expect_matches_oracle(pkg, orc, dtype, d, tol):
    _say(f"case expect vs oracle {dtype} d={d}")
    t, s, e = _converged(pkg, orc, 4, 3, d, 71, dtype)
    m = e.messages()
    rng = np.random.default_rng(9)
    ops = rng.standard_normal((t.nv, d, d))
    a = rng.standard_normal((t.ne, d, d))
    b = rng.standard_normal((t.ne, d, d))
    gv = e.expect_vertex(ops)
    ge = e.expect_edge(a, b)
    e.close()
    s_used = s.astype(np.float32).astype(np.float64) if dtype == "f32" else s
    wv = orc.expect_vertex(t, s_used, m, ops)
    we = orc.expect_edge(t, s_used, m, a, b)
    errv = float(np.max(np.abs(gv - wv)) / np.max(np.abs(wv)))
    erre = float(np.max(np.abs(ge - we)) / np.max(np.abs(we)))
    _say(f"vertex rel err {errv:.2e}, edge rel err {erre:.2e}")
    assert errv < tol and erre < tol


@pytest.mark.parametrize("dtype,path,tol", [("f64", "auto", 1e-9), ("f32", "auto", 1e-6),
                                            ("f32", "generic", 1e-6)])
def test_expect_assembles_tfim_energy(pkg, orc, dtype, path, tol):
    """sum_e -J <Z Z> + sum_v -h <X> from the expectation API equals the engine's TFIM
    energy (on the FP32 vertex path the energy comes from the MFMA kernels)."""
    _say(f"case expect tfim {dtype} {path}")
    t = pkg.build_topology(pkg.named_grid((5, 5)), 4, 2)
    s = pkg.peps.random_sites(t, 72, scale=pkg.peps.peps_scale(2, 4))
    with pkg.Engine(t, s, dtype=dtype, path=path) as e:
        e.init_messages("identity")
        e.run(maxiter=500, tol=1e-12, criterion="fidelity")
        ev = e.expect_vertex(np.broadcast_to(X, (t.nv, 2, 2)))
        ee = e.expect_edge(np.broadcast_to(Z, (t.ne, 2, 2)), np.broadcast_to(Z, (t.ne, 2, 2)))
        E = e.tfim_energy(1.0, 3.0)
    got = -3.0 * ev.sum() - ee.sum()
    _say(f"assembled {got:.12f} engine {E:.12f}")
    assert got == pytest.approx(E, rel=tol)


def test_expect_host_api(pkg, orc):
    """expect / expect2 on a NormNetwork built from ITensors, against the oracle on the
    same flattened network; identity insertions give 1."""
    _say("case expect host api")
    g = pkg.named_grid((3, 3))
    links = {frozenset(e): pkg.Index(2) for e in g.edges()}
    sites = {v: pkg.Index(2) for v in g.vertices()}
    rng = np.random.default_rng(10)
    tn = pkg.tensornetwork(lambda v: pkg.random_itensor(
        [sites[v]] + [links[frozenset(e)] for e in g.incident_edges(v)], rng), g.vertices())
    nn = pkg.normnetwork(tn)
    cache = pkg.beliefpropagation(nn, pkg.message_environment(pkg.one, nn),
                                  stopping_criterion={"maxiter": 200, "tol": 1e-13})
    v0 = g.vertices()[4]
    O = rng.standard_normal((2, 2))
    got = pkg.expect(nn, cache, {v0: O, g.vertices()[0]: np.eye(2)})
    assert got[g.vertices()[0]] == pytest.approx(1.0, abs=1e-12)
    e0 = list(g.edges())[3]
    A, B = rng.standard_normal((2, 2, 2))
    got2 = pkg.expect2(nn, cache, {e0: (A, B), (e0[1], e0[0]): (B, A)})
    assert got2[e0] == pytest.approx(got2[(e0[1], e0[0])], rel=1e-12)
    # the oracle on the same flattened network
    fl = pkg.to_topology(nn)
    t = fl.topo
    m = cache.to_flat(t)
    ops = np.broadcast_to(np.eye(2), (t.nv, 2, 2)).copy()
    ops[t.vertices.index(v0)] = O
    want = orc.expect_vertex(t, fl.sites, m, ops)[t.vertices.index(v0)]
    assert got[v0] == pytest.approx(want, rel=1e-11)
    own = t.port_owner()
    iu, iv = t.vertices.index(e0[0]), t.vertices.index(e0[1])
    a = np.broadcast_to(np.eye(2), (t.ne, 2, 2)).copy()
    b = a.copy()
    for q in range(t.nports):
        if int(own[q]) == iu and int(t.port_nbr[q]) == iv:
            k = int(t.port_edge[q])
            if q < int(t.port_rev[q]):
                a[k], b[k] = A, B
            else:
                a[k], b[k] = B, A
    want2 = orc.expect_edge(t, fl.sites, m, a, b)[k]
    assert got2[e0] == pytest.approx(want2, rel=1e-11)


# ---------------------------------------------------------------- local energies + gradients

def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("dtype,gtol", [("f64", 1e-7), ("f32", 2e-3)])
def test_local_energy_gradient_vs_oracle(pkg, orc, dtype, gtol):
    """Random non-symmetric one- and two-site operators: energy and fixed-point gradient on
    the device against the oracle's (explicit partials + adjoint solve, FD-pinned)."""
    _say(f"case local energy gradient {dtype}")
    t = pkg.build_topology(pkg.named_grid((4, 4)), 3, 2)
    s = pkg.peps.random_sites(t, 73, scale=pkg.peps.peps_scale(2, 3))
    rng = np.random.default_rng(13)
    ov = rng.standard_normal((t.nv, 2, 2))
    oa = rng.standard_normal((t.ne, 2, 2))
    ob = rng.standard_normal((t.ne, 2, 2))
    with pkg.Engine(t, s, dtype=dtype) as e:
        e.init_messages("identity")
        if dtype == "f64":
            r = e.run(maxiter=2000, tol=1e-14, criterion="maxabs")
        else:
            r = e.run(maxiter=500, tol=1e-12, criterion="fidelity")
        assert r.converged[0]
        m = e.messages()
        E, g, it, res = e.local_energy(ov, oa, ob, grad=True, adj_tol=1e-10, adj_maxiter=2000)
        Ed, gd, _, _ = e.local_energy(ov, oa, ob, grad=True, adj_tol=1e-10, adj_maxiter=2000,
                                      device=True)
        gdev = gd.numpy().astype(np.float64)
        gd.free()
    s_used = s.astype(np.float32).astype(np.float64) if dtype == "f32" else s
    m64, _, _, at = orc.bp(t, s_used, m, maxiter=3000, tol=1e-15, criterion="maxabs")
    Eo, dm, da = orc.local_energy(t, s_used, m64, ov, oa, ob, partials=True)
    want, _, _ = orc.fixed_point_gradient(t, s_used, m64, dm, da, maxiter=5000, tol=1e-13)
    Em = orc.local_energy(t, s_used, m, ov, oa, ob)  # the oracle at the engine's messages
    _say(f"E={E:.12f} oracle={Em:.12f} (at its own fixed point {Eo:.12f}) sweeps={it} "
         f"res={res:.1e} grad err={_rel(g, want):.2e}")
    assert E == pytest.approx(Em, rel=1e-12 if dtype == "f64" else 1e-5)
    assert Ed == E
    assert _rel(g, want) < gtol
    assert _rel(gdev, g) < (1e-12 if dtype == "f64" else 1e-6)


@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_local_energy_reduces_to_tfim(pkg, dtype):
    """O_v = -h X, A_e = -J Z, B_e = Z: energy and gradient equal the TFIM entry point's
    (on FP32 contexts the TFIM runs on the MFMA kernels, the local energy on the generic
    ones)."""
    _say(f"case local energy = tfim {dtype}")
    t = pkg.build_topology(pkg.named_grid((5, 5)), 4, 2)
    s = pkg.peps.random_sites(t, 74, scale=pkg.peps.peps_scale(2, 4))
    J, h = 1.0, 3.0
    with pkg.Engine(t, s, dtype=dtype) as e:
        e.init_messages("identity")
        e.run(maxiter=500, tol=1e-12, criterion="fidelity")
        E0, g0, _, _ = e.tfim_energy(J, h, grad=True, adj_tol=1e-9, adj_maxiter=2000)
        E1, g1, _, _ = e.local_energy(np.broadcast_to(-h * X, (t.nv, 2, 2)),
                                      np.broadcast_to(-J * Z, (t.ne, 2, 2)),
                                      np.broadcast_to(Z, (t.ne, 2, 2)), grad=True,
                                      adj_tol=1e-9, adj_maxiter=2000)
    _say(f"E tfim={E0:.12f} local={E1:.12f} grad rel diff={_rel(g1, g0):.2e}")
    assert E1 == pytest.approx(E0, rel=1e-10 if dtype == "f64" else 1e-6)
    assert _rel(g1, g0) < (1e-8 if dtype == "f64" else 2e-4)
