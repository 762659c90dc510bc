"""Pins of the FP64 CPU oracle against the reference's own known answers
and against what the mathematics fixes.  CPU only.

Reference tests mirrored (ITensorNetworksNext.jl test/):
  test_beliefpropagation.jl:153-202  BP on a chain / comb tree equals exact contraction
  test_beliefpropagation.jl:204-225  spin ice: exp(bethe_free_energy) == 1.5^(n^2)
  test_beliefpropagation.jl:134-150  iterate_diff(c, copy(c)) ~ 0
  test_contract_network.jl:13-24     ABCD == 15.5 * [-2, 3, 4, 5, 1]
  test_contract_network.jl:26-45     contraction order independence on a 4x4 grid
  test_normnetwork.jl:137-166        <psi|psi> == ||psi||^2, single normalised tensor == 1
  test_itensornetworkgenerators.jl:29-54 (with test/utils.jl:12-39) 1D Ising closed forms
"""
import math

import numpy as np
import pytest


def _norm_tree(pkg, dims, D, seed):
    g = pkg.named_comb_tree(dims) if dims[1] > 1 else pkg.named_path_graph(dims[0])
    t = pkg.build_topology(g, D, 2)
    s = pkg.peps.random_sites(t, seed)
    return g, t, s


@pytest.mark.parametrize("dims,D", [((2, 1), 2), ((4, 3), 3)])
def test_norm_tree_bp_is_exact(pkg, orc, dims, D):
    g, t, s = _norm_tree(pkg, dims, D, 11)
    exact = orc.norm_exact(t, s)
    # Jacobi to convergence
    m, sw, dl, at = orc.bp(t, s, pkg.peps.ones_messages(t), maxiter=60, tol=1e-15,
                           criterion="maxabs")
    assert at > 0
    z = math.exp(orc.bethe_free_energy(t, s, m))
    assert abs(z / exact - 1) < 1e-12
    # one sequential forest-cover sweep (the reference runs maxiter = 1)
    order = [t.slot(a, b) for a, b in pkg.forest_cover_edge_sequence(g)]
    ms = orc.sweep_sequential(t, s, pkg.peps.ones_messages(t), order)
    assert abs(math.exp(orc.bethe_free_energy(t, s, ms)) / exact - 1) < 1e-12


def test_plain_tree_bp_is_exact(pkg, orc):
    g = pkg.named_comb_tree((4, 3))
    t = pkg.build_topology(g, 3, plain=True)
    s = pkg.peps.random_sites(t, 3)
    order = [t.slot(a, b) for a, b in pkg.forest_cover_edge_sequence(g)]
    ms = orc.sweep_sequential(t, s, pkg.peps.ones_messages(t), order)
    z = orc.bethe_free_energy(t, s, ms)
    exact = orc.plain_exact(t, s)
    assert abs(np.exp(complex(z)) / exact - 1) < 1e-12


@pytest.mark.parametrize("n", [3, 4, 5])
def test_spin_ice_pauling(pkg, orc, n):
    g = pkg.named_grid((n, n), periodic=True)
    t = pkg.build_topology(g, 2, plain=True)
    s = pkg.peps.spin_ice_sites(t)
    order = [t.slot(a, b) for a, b in pkg.forest_cover_edge_sequence(g)]
    m, sw, dl, at = orc.bp(t, s, pkg.peps.random_messages(t, 123), maxiter=10, tol=1e-10,
                           schedule="sequential", order=order)
    z = math.exp(orc.bethe_free_energy(t, s, m))
    assert z == pytest.approx(1.5 ** (n * n), rel=1e-8)


def test_product_state_exact(pkg, orc):
    g = pkg.named_grid((3, 3))
    t = pkg.build_topology(g, 1, 2)
    s = pkg.peps.random_sites(t, 5)
    m, _, _, _ = orc.bp(t, s, pkg.peps.identity_messages(t), maxiter=3)
    z = math.exp(orc.bethe_free_energy(t, s, m))
    prod = np.prod([np.sum(s[t.site_off[u]:t.site_off[u + 1]] ** 2) for u in range(t.nv)])
    assert z == pytest.approx(prod, rel=1e-13)
    assert orc.norm_exact(t, s) == pytest.approx(prod, rel=1e-13)


def test_contract_abcd(orc):
    A = np.array([[1.0, 1.0], [0.5, 1.0]])
    B = np.array([2.0, 1.0])
    Cv = np.array([5.0, 1.0])
    Dv = np.array([-2.0, 3.0, 4.0, 5.0, 1.0])
    val, labs = orc.contract_labeled([A, B, Cv, Dv], [("i", "j"), ("i",), ("j",), ("k",)])
    assert labs == ["k"]
    np.testing.assert_array_equal(val, 15.5 * Dv)
    val2, _ = orc.contract_labeled([Dv, Cv, A, B], [("k",), ("j",), ("i", "j"), ("i",)])
    np.testing.assert_array_equal(val2, 15.5 * Dv)


def test_contract_order_independence(pkg, orc):
    g = pkg.named_grid((4, 4))
    t = pkg.build_topology(g, 2, plain=True)
    s = pkg.peps.random_sites(t, 9)
    tensors, labels = [], []
    for u in range(t.nv):
        shp = t.site_shape(u)[1:]
        tensors.append(s[t.site_off[u]:t.site_off[u + 1]].reshape(shp))
        labels.append(tuple(int(t.port_edge[q]) for q in range(t.port_off[u], t.port_off[u + 1])))
    z1, _ = orc.contract_labeled(tensors, labels)
    perm = np.random.default_rng(0).permutation(len(tensors))
    z2, _ = orc.contract_labeled([tensors[i] for i in perm], [labels[i] for i in perm])
    assert abs(z1 - z2) / abs(z1) <= 1e3 * np.finfo(float).eps


def test_norm_network_physics(pkg, orc):
    # single normalised tensor contracts to 1
    g = pkg.NamedGraph(vertices=[1])
    t = pkg.build_topology(g, 1, 3)
    v = np.random.default_rng(1).standard_normal(3)
    assert orc.norm_exact(t, v / np.linalg.norm(v)) == pytest.approx(1.0, rel=1e-14)
    # 2x2 grid: <psi|psi> = ||psi||^2 with psi built by numpy einsum (a library routine)
    g = pkg.named_grid((2, 2))
    t = pkg.build_topology(g, 2, 2)
    s = pkg.peps.random_sites(t, 2)
    A = [s[t.site_off[u]:t.site_off[u + 1]].reshape(t.site_shape(u)) for u in range(t.nv)]
    lab = "abcd"
    subs = []
    for u in range(t.nv):
        subs.append("spqr"[u] + "".join(lab[t.port_edge[q]] for q in range(t.port_off[u], t.port_off[u + 1])))
    psi = np.einsum(",".join(subs) + "->spqr", *A)
    assert orc.norm_exact(t, s) == pytest.approx(float(np.sum(psi ** 2)), rel=1e-13)


def test_iterate_diff_identical(pkg, orc):
    t = pkg.build_topology(pkg.named_grid((2,)), 2, 2)
    m = pkg.peps.ones_messages(t)
    assert orc.iterate_diff(t, m, m.copy()) == pytest.approx(0.0, abs=10 * np.finfo(float).eps)


def _ising_chain(pkg, n, beta, periodic):
    """1D Ising plain network: each bond matrix split as sqrt(M) onto its two vertices
    (the construction of src/ITensorNetworkGenerators/ising_network.jl:7-51 with h = 0)."""
    g = pkg.named_path_graph(n)
    if periodic:
        g.add_edge(n, 1)
    t = pkg.build_topology(g, 2, plain=True)
    M = np.array([[math.exp(beta), math.exp(-beta)], [math.exp(-beta), math.exp(beta)]])
    w, V = np.linalg.eigh(M)
    S = V @ np.diag(np.sqrt(w)) @ V.T
    s = np.zeros(int(t.site_off[-1]))
    for u in range(t.nv):
        k = t.degree(u)
        T = np.zeros((2,) * k)
        for spin in range(2):
            T[(spin,) * k] = 1.0
        for j in range(k):
            T = np.moveaxis(np.tensordot(T, S, axes=([j], [0])), -1, j)
        s[t.site_off[u]:t.site_off[u + 1]] = T.ravel()
    return g, t, s


@pytest.mark.parametrize("periodic", [False, True])
def test_ising_1d_closed_form(pkg, orc, periodic):
    n, beta = 4, 0.4
    g, t, s = _ising_chain(pkg, n, beta, periodic)
    Z = orc.plain_exact(t, s)
    if periodic:
        lp, lm = 2 * math.cosh(beta), 2 * math.sinh(beta)
        want = lp ** n + lm ** n
    else:
        want = 2 * (2 * math.cosh(beta)) ** (n - 1)
    assert Z == pytest.approx(want, rel=1e-13)
    if not periodic:  # a path is a tree: BP exact
        order = [t.slot(a, b) for a, b in pkg.forest_cover_edge_sequence(g)]
        ms = orc.sweep_sequential(t, s, pkg.peps.ones_messages(t), order)
        assert math.exp(orc.bethe_free_energy(t, s, ms)) == pytest.approx(want, rel=1e-12)


def test_jacobi_and_sequential_share_fixed_point(pkg, orc):
    g = pkg.named_grid((4, 4))
    t = pkg.build_topology(g, 2, 2)
    s = pkg.peps.random_sites(t, 1234)
    mj, _, _, atj = orc.bp(t, s, pkg.peps.identity_messages(t), maxiter=400, tol=1e-14,
                           criterion="maxabs")
    order = [t.slot(a, b) for a, b in pkg.forest_cover_edge_sequence(g)]
    ms, _, _, ats = orc.bp(t, s, pkg.peps.identity_messages(t), maxiter=400, tol=1e-14,
                           criterion="maxabs", schedule="sequential", order=order)
    assert atj > 0 and ats > 0
    assert np.max(np.abs(mj - ms)) < 1e-11
    assert orc.bethe_free_energy(t, s, mj) == pytest.approx(orc.bethe_free_energy(t, s, ms),
                                                            abs=1e-11)


def test_bethe_message_scale_invariance(pkg, orc):
    t = pkg.build_topology(pkg.named_grid((3, 3)), 2, 2)
    s = pkg.peps.random_sites(t, 4)
    m, _, _, _ = orc.bp(t, s, pkg.peps.identity_messages(t), maxiter=50)
    scaled = m.copy()
    rng = np.random.default_rng(0)
    for q in range(t.nports):
        scaled[t.msg_off[q]:t.msg_off[q + 1]] *= rng.uniform(0.1, 10.0)
    assert orc.bethe_free_energy(t, s, scaled) == pytest.approx(
        orc.bethe_free_energy(t, s, m), abs=1e-12)


def test_message_update_matches_einsum(pkg, orc):
    """One interior message of a 3x3 lattice with random (non-symmetric) incoming
    messages, against numpy.einsum of the defining contraction."""
    t = pkg.build_topology(pkg.named_grid((3, 3)), 3, 2)
    s = pkg.peps.random_sites(t, 8)
    m = pkg.peps.random_messages(t, 8)
    u = t.vertices.index((2, 2))
    A = s[t.site_off[u]:t.site_off[u + 1]].reshape(t.site_shape(u))
    ports = list(range(t.port_off[u], t.port_off[u + 1]))
    M = [m[t.msg_off[t.port_rev[q]]:t.msg_off[t.port_rev[q] + 1]].reshape(3, 3) for q in ports]
    for p, q in enumerate(ports):
        got = orc.message(t, s, m, q).reshape(3, 3)
        ket, bra = "abcd", "ABCD"
        terms = ["s" + ket, "s" + bra]
        for j in range(4):
            if j != p:
                terms.append(ket[j] + bra[j])
        ops = [A, A] + [M[j] for j in range(4) if j != p]
        want = np.einsum(",".join(terms) + "->" + ket[p] + bra[p], *ops)
        np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-14)


# ------------------------------------------------------------------ adjoint pins


def _converged(pkg, orc, L, D, seed, scale=True):
    t = pkg.build_topology(pkg.named_grid((L, L)), D, 2)
    sc = pkg.peps.peps_scale(2, D) if scale else None
    s = pkg.peps.random_sites(t, seed, scale=sc)
    m, _, _, at = orc.bp(t, s, pkg.peps.identity_messages(t), maxiter=2000, tol=1e-15,
                         criterion="maxabs")
    assert at > 0
    return t, s, m


def _fd(pkg, orc, t, s, m, f, idx, h):
    out = []
    for i in idx:
        vals = []
        for sgn in (1, -1):
            s2 = s.copy()
            s2[i] += sgn * h
            m2, _, _, _ = orc.bp(t, s2, m, maxiter=2000, tol=1e-15, criterion="maxabs")
            vals.append(f(t, s2, m2))
        out.append((vals[0] - vals[1]) / (2 * h))
    return np.array(out)


def test_grad_lognorm_closed_form_vs_fd(pkg, orc):
    t, s, m = _converged(pkg, orc, 3, 2, 21, scale=False)
    g = orc.grad_lognorm_explicit(t, s, m)
    idx = np.random.default_rng(3).choice(s.size, 8, replace=False)
    fd = _fd(pkg, orc, t, s, m, lambda t_, s_, m_: orc.bethe_free_energy(t_, s_, m_), idx, 1e-5)
    np.testing.assert_allclose(g[idx], fd, rtol=1e-6, atol=1e-8)
    # the full implicit-function gradient agrees: the message term vanishes at the fixed
    # point (Bethe stationarity)
    full, it, res = orc.fixed_point_gradient(t, s, m, None, g, tol=1e-13)
    np.testing.assert_allclose(full, g, rtol=1e-9, atol=1e-11)


def test_tfim_gradient_vs_fd(pkg, orc):
    t, s, m = _converged(pkg, orc, 3, 2, 22, scale=False)
    E, dm, da = orc.tfim_energy(t, s, m, J=1.0, h=3.0, partials=True)
    grad, it, res = orc.fixed_point_gradient(t, s, m, dm, da, maxiter=5000, tol=1e-14)
    assert res < 1e-12
    idx = np.random.default_rng(4).choice(s.size, 8, replace=False)
    fd = _fd(pkg, orc, t, s, m, lambda t_, s_, m_: orc.tfim_energy(t_, s_, m_, 1.0, 3.0), idx,
             1e-5)
    np.testing.assert_allclose(grad[idx], fd, rtol=1e-6, atol=1e-8)
    # dropping the adjoint (message) term is visibly wrong: the adjoint matters here
    assert np.max(np.abs(da[idx] - fd)) > 1e-4 * np.max(np.abs(fd))


def test_tfim_product_state_closed_form(pkg, orc):
    """D = 1: BP is exact and <O> factorises into single-site expectations."""
    g = pkg.named_grid((3, 3))
    t = pkg.build_topology(g, 1, 2)
    s = pkg.peps.random_sites(t, 6)
    m, _, _, _ = orc.bp(t, s, pkg.peps.identity_messages(t), maxiter=3)
    J, h = 1.0, 3.0
    ex, ez = [], []
    for u in range(t.nv):
        a = s[t.site_off[u]:t.site_off[u + 1]]
        n2 = a @ a
        ex.append(2 * a[0] * a[1] / n2)
        ez.append((a[0] ** 2 - a[1] ** 2) / n2)
    want = -h * sum(ex)
    for (x, y) in t.edges:
        want -= J * ez[t.vertices.index(x)] * ez[t.vertices.index(y)]
    assert orc.tfim_energy(t, s, m, J, h) == pytest.approx(want, rel=1e-13)


# ---------------------------------------------------------------- expectation values

def _state_vector(t, s):
    """psi[s_0, ..., s_{nv-1}] of a small network by direct contraction of its sites."""
    args = []
    for u in range(t.nv):
        k = t.degree(u)
        ports = [int(t.port_off[u]) + j for j in range(k)]
        dims = [int(t.phys_dim[u])] + [int(t.edge_dim[t.port_edge[q]]) for q in ports]
        args += [s[t.site_off[u]:t.site_off[u + 1]].reshape(dims),
                 [u] + [t.nv + int(t.port_edge[q]) for q in ports]]
    return np.einsum(*args, list(range(t.nv)))


def _apply(op, psi, axis):
    return np.moveaxis(np.tensordot(op, psi, axes=([1], [axis])), 0, axis)


def _edge_ends(t):
    """(first, second) vertex of every topology edge: first owns the edge's lower port."""
    own = t.port_owner()
    ends = {}
    for q in range(t.nports):
        r = int(t.port_rev[q])
        if q < r:
            ends[int(t.port_edge[q])] = (int(own[q]), int(own[r]))
    return [ends[e] for e in range(t.ne)]


@pytest.mark.parametrize("kind", ["path", "comb"])
def test_expect_tree_exact(pkg, orc, kind):
    """On a tree BP is exact: <O_v>_BP and <A_e B_e>_BP (adjacent sites) equal the exact
    <psi|O|psi>/<psi|psi> for random, non-symmetric operators."""
    g = pkg.named_path_graph(5) if kind == "path" else pkg.named_comb_tree((3, 2))
    t = pkg.build_topology(g, 2, 2)
    s = pkg.peps.random_sites(t, 31)
    m, _, _, at = orc.bp(t, s, pkg.peps.identity_messages(t), maxiter=200, tol=1e-15,
                         criterion="maxabs")
    assert at > 0
    rng = np.random.default_rng(5)
    psi = _state_vector(t, s)
    nn = float(np.sum(psi * psi))
    ops = rng.standard_normal((t.nv, 2, 2))
    got = orc.expect_vertex(t, s, m, ops)
    for u in range(t.nv):
        want = float(np.sum(psi * _apply(ops[u], psi, u))) / nn
        assert got[u] == pytest.approx(want, rel=1e-11, abs=1e-13)
    a = rng.standard_normal((t.ne, 2, 2))
    b = rng.standard_normal((t.ne, 2, 2))
    got_e = orc.expect_edge(t, s, m, a, b)
    for e, (u, v) in enumerate(_edge_ends(t)):
        want = float(np.sum(psi * _apply(a[e], _apply(b[e], psi, v), u))) / nn
        assert got_e[e] == pytest.approx(want, rel=1e-11, abs=1e-13)


def test_expect_identity_and_linearity(pkg, orc):
    """Identity operators give exactly 1; the estimates are linear in the operator."""
    t, s, m = _converged(pkg, orc, 4, 3, 32)
    one = np.broadcast_to(np.eye(2), (t.nv, 2, 2))
    assert np.allclose(orc.expect_vertex(t, s, m, one), 1.0, rtol=0, atol=1e-13)
    onee = np.broadcast_to(np.eye(2), (t.ne, 2, 2))
    assert np.allclose(orc.expect_edge(t, s, m, onee, onee), 1.0, rtol=0, atol=1e-13)
    rng = np.random.default_rng(6)
    o1, o2 = rng.standard_normal((2, t.nv, 2, 2))
    lhs = orc.expect_vertex(t, s, m, 2.0 * o1 - 3.0 * o2)
    rhs = 2.0 * orc.expect_vertex(t, s, m, o1) - 3.0 * orc.expect_vertex(t, s, m, o2)
    assert np.allclose(lhs, rhs, rtol=1e-12, atol=1e-12)


def test_expect_reproduces_tfim_energy(pkg, orc):
    """sum_e -J <Z Z>_BP + sum_v -h <X>_BP equals the (separately pinned) TFIM energy."""
    t, s, m = _converged(pkg, orc, 4, 3, 33)
    X = np.array([[0.0, 1.0], [1.0, 0.0]])
    Z = np.array([[1.0, 0.0], [0.0, -1.0]])
    J, h = 1.0, 3.0
    ev = orc.expect_vertex(t, s, m, np.broadcast_to(X, (t.nv, 2, 2)))
    ee = orc.expect_edge(t, s, m, np.broadcast_to(Z, (t.ne, 2, 2)),
                         np.broadcast_to(Z, (t.ne, 2, 2)))
    assert -h * ev.sum() - J * ee.sum() == pytest.approx(orc.tfim_energy(t, s, m, J, h),
                                                          rel=1e-13)


def _tfim_ops(t, J, h):
    X = np.array([[0.0, 1.0], [1.0, 0.0]])
    Z = np.array([[1.0, 0.0], [0.0, -1.0]])
    return (np.broadcast_to(-h * X, (t.nv, 2, 2)), np.broadcast_to(-J * Z, (t.ne, 2, 2)),
            np.broadcast_to(Z, (t.ne, 2, 2)))


def test_local_energy_reduces_to_tfim(pkg, orc):
    """With O_v = -h X, A_e = -J Z, B_e = Z the local energy and both explicit partials are
    the (FD-pinned) TFIM energy's."""
    t, s, m = _converged(pkg, orc, 3, 3, 34)
    E0, dm0, da0 = orc.tfim_energy(t, s, m, J=0.7, h=2.5, partials=True)
    ov, oa, ob = _tfim_ops(t, 0.7, 2.5)
    E, dm, da = orc.local_energy(t, s, m, ov, oa, ob, partials=True)
    assert E == pytest.approx(E0, rel=1e-13)
    np.testing.assert_allclose(dm, dm0, rtol=1e-11, atol=1e-13 * np.max(np.abs(dm0)))
    np.testing.assert_allclose(da, da0, rtol=1e-11, atol=1e-13 * np.max(np.abs(da0)))


def test_local_energy_gradient_vs_fd(pkg, orc):
    """Random non-symmetric one- and two-site operators: the implicit gradient (explicit
    partials + adjoint solve) matches central finite differences of re-converged BP."""
    t, s, m = _converged(pkg, orc, 3, 2, 35, scale=False)
    rng = np.random.default_rng(12)
    ov = rng.standard_normal((t.nv, 2, 2))
    oa = rng.standard_normal((t.ne, 2, 2))
    ob = rng.standard_normal((t.ne, 2, 2))
    E, dm, da = orc.local_energy(t, s, m, ov, oa, ob, partials=True)
    grad, it, res = orc.fixed_point_gradient(t, s, m, dm, da, maxiter=5000, tol=1e-14)
    assert res < 1e-12
    idx = rng.choice(s.size, 8, replace=False)
    fd = _fd(pkg, orc, t, s, m,
             lambda t_, s_, m_: orc.local_energy(t_, s_, m_, ov, oa, ob), idx, 1e-5)
    np.testing.assert_allclose(grad[idx], fd, rtol=1e-6, atol=1e-8)
    # the value is the sum of the expectation values
    want = (orc.expect_vertex(t, s, m, ov).sum() + orc.expect_edge(t, s, m, oa, ob).sum())
    assert E == pytest.approx(want, rel=1e-13)


# ------------------------------------------------------------ gate application oracle

def _two_site(A1, A2, p1, p2):
    return np.tensordot(A1, A2, axes=([1 + p1], [1 + p2]))


def test_gate_oracle_untruncated_exact(orc):
    """test_apply_operator.jl:52-74: with invertible message square roots an untruncated
    two-site gate reproduces G applied to the two-site tensor (any gauge)."""
    rng = np.random.default_rng(21)
    for trial in range(10):
        d1, d2, D = 3, 2, 3
        dims1, dims2 = [2, D, 3], [D, 4]
        A1 = rng.standard_normal([d1] + dims1)
        A2 = rng.standard_normal([d2] + dims2)
        psd = lambda n: (lambda B: B @ B.T + 0.1 * np.eye(n))(rng.standard_normal((n, n)))
        m1 = {0: psd(2), 2: psd(3)}
        m2 = {1: psd(4)}
        G = rng.standard_normal((d1 * d2, d1 * d2))
        B1, B2, S = orc.apply_gate2(A1, A2, 1, 0, m1, m2, G)
        got = _two_site(B1, B2, 1, 0)
        want = np.tensordot(G.reshape(d1, d2, d1, d2), _two_site(A1, A2, 1, 0),
                            axes=([2, 3], [0, 3]))
        want = np.moveaxis(want, 1, 3)  # [t1, a, c, t2, e]
        assert np.max(np.abs(got - want)) < 1e-12 * np.max(np.abs(want))
        assert np.all(np.diff(S) <= 0)


def test_gate_oracle_truncation_is_optimal_on_a_tree(pkg, orc):
    """test_apply_operator.jl:76-98: on a path with the oracle's converged (exact) BP
    environments, a rank-k two-site gate equals the global optimal truncation, and its
    singular values are proportional to the Schmidt values."""
    g = pkg.named_path_graph(4)
    t = pkg.build_topology(g, 3, 2)
    s = pkg.peps.random_sites(t, 5)
    m, _, _, at = orc.bp(t, s, pkg.peps.identity_messages(t), maxiter=400, tol=1e-15,
                         criterion="maxabs")
    assert at >= 0
    A = [s[t.site_off[u]:t.site_off[u + 1]].reshape(t.site_shape(u)) for u in range(t.nv)]

    def incoming(u, skip):
        out = {}
        for j, q in enumerate(range(t.port_off[u], t.port_off[u + 1])):
            w = int(t.port_nbr[q])
            if w == skip:
                continue
            sl = int(t.port_rev[q])
            D = int(t.edge_dim[t.port_edge[q]])
            out[j] = m[t.msg_off[sl]:t.msg_off[sl + 1]].reshape(D, D)
        return out

    u1, u2 = 1, 2
    p1 = [int(t.port_nbr[q]) for q in range(t.port_off[u1], t.port_off[u1 + 1])].index(u2)
    p2 = [int(t.port_nbr[q]) for q in range(t.port_off[u2], t.port_off[u2 + 1])].index(u1)
    rng = np.random.default_rng(3)
    G = rng.standard_normal((4, 4))

    def full(tensors):
        labels = []
        for u in range(t.nv):
            ports = range(t.port_off[u], t.port_off[u + 1])
            labels.append((("s", u),) + tuple(("e", int(t.port_edge[q])) for q in ports))
        val, labs = orc.contract_labeled(tensors, labels)
        return np.transpose(val, [labs.index(("s", u)) for u in range(t.nv)])

    psi = full(A)
    gated = np.moveaxis(np.tensordot(G.reshape(2, 2, 2, 2), psi, axes=([2, 3], [1, 2])),
                        [0, 1], [1, 2])
    U, S, Vt = np.linalg.svd(gated.reshape(4, 4), full_matrices=False)
    for k in (1, 2, 3):
        B1, B2, Sk = orc.apply_gate2(A[u1], A[u2], p1, p2, incoming(u1, u2), incoming(u2, u1),
                                     G, maxdim=k)
        T = list(A)
        T[u1], T[u2] = B1, B2
        best = (U[:, :k] * S[:k]) @ Vt[:k]
        assert np.max(np.abs(full(T).reshape(4, 4) - best)) < 1e-10 * np.max(np.abs(best))
        assert np.max(np.abs(Sk / Sk[0] - S[:k] / S[0])) < 1e-10


# ------------------------------------------------------------ iterate_diff closed forms
# beliefpropagation.jl:261-267: max over directed edges of 1 - |<m1^, m2^>|^2 with
# m^ = m / ||m||_F.  For I_2 against diag(1, 0): <I/sqrt2, diag(1,0)> = 1/sqrt2, so the
# distance is 1 - 1/2 = 0.5 (without the square it would be 1 - 0.707 = 0.293); m against
# -m or 3m: 0 (phase and scale drop out); orthogonal messages: 1.

def _two_vertex_topo(pkg):
    t = pkg.build_topology(pkg.named_path_graph(2), 2, 2)
    assert t.nports == 2
    return t


def _pack(t, mats):
    out = np.zeros(int(t.msg_off[-1]))
    for q, m in enumerate(mats):
        out[t.msg_off[q]:t.msg_off[q + 1]] = np.asarray(m, dtype=float).ravel()
    return out


@pytest.mark.parametrize("a,b,want", [
    (np.eye(2), np.diag([1.0, 0.0]), 0.5),
    (np.array([[1.0, 2.0], [3.0, 4.0]]), -np.array([[1.0, 2.0], [3.0, 4.0]]), 0.0),
    (np.array([[1.0, 2.0], [3.0, 4.0]]), 3.0 * np.array([[1.0, 2.0], [3.0, 4.0]]), 0.0),
    (np.diag([1.0, 0.0]), np.diag([0.0, 1.0]), 1.0),
    (np.array([[1.0, 1.0], [0.0, 0.0]]), np.array([[1.0, 0.0], [0.0, 0.0]]), 0.5),
])
def test_iterate_diff_closed_forms(pkg, orc, a, b, want):
    t = _two_vertex_topo(pkg)
    m1 = _pack(t, [a, a])
    m2 = _pack(t, [b, b])
    assert orc.iterate_diff(t, m1, m2) == pytest.approx(want, abs=1e-15)


def test_iterate_diff_is_the_maximum_over_edges(pkg, orc):
    t = _two_vertex_topo(pkg)
    m1 = _pack(t, [np.eye(2), np.eye(2)])
    m2 = _pack(t, [np.eye(2), np.diag([1.0, 0.0])])
    assert orc.iterate_diff(t, m1, m2) == pytest.approx(0.5, abs=1e-15)
    assert orc.iterate_diff(t, m2, m1) == pytest.approx(0.5, abs=1e-15)


# ------------------------------------------------------------ loop-count semantics
# StopAfterIteration(maxiter) | StopWhenConverged(tol) (AlgorithmsInterfaceExtensions.jl:
# 84-109).  On an N-vertex path, Jacobi message m_{k -> k+1} depends on the k - 1 messages
# behind it, so every message is exact after N - 1 sweeps and not before; sweep N repeats
# sweep N - 1 bit for bit (same inputs), so its max-abs change is exactly 0 and the run
# stops at iteration N.  These counts are fixed by the path length, not by the oracle.

@pytest.mark.parametrize("N", [2, 4, 7])
def test_path_jacobi_exact_after_n_minus_1_sweeps(pkg, orc, N):
    t = pkg.build_topology(pkg.named_path_graph(N), 3, 2)
    s = pkg.peps.random_sites(t, 5)
    m0 = pkg.peps.ones_messages(t)
    exact = math.log(orc.norm_exact(t, s))
    fixed, _, _, _ = orc.bp(t, s, m0, maxiter=3 * N)
    m, sw, dl, at = orc.bp(t, s, m0, maxiter=N - 1)
    assert (sw, at) == (N - 1, -1)
    np.testing.assert_allclose(m, fixed, rtol=0, atol=1e-14)
    assert orc.bethe_free_energy(t, s, m) == pytest.approx(exact, abs=1e-12)
    if N > 2:
        # the messages m_{1->2} ... m_{N-1->N} chain: one of them is still off
        m, _, _, _ = orc.bp(t, s, m0, maxiter=N - 2)
        assert np.max(np.abs(m - fixed)) > 1e-6
    m, sw, dl, at = orc.bp(t, s, m0, maxiter=100, tol=1e-300, criterion="maxabs")
    assert (sw, at, dl) == (N, N, 0.0)


def test_convergence_check_skips_iteration_zero(pkg, orc):
    """With a tolerance every delta passes, the run still sweeps once: the check at
    iteration 0 (where delta would be 0) is skipped (AIE.jl:98-99)."""
    t = pkg.build_topology(pkg.named_grid((3, 3)), 2, 2)
    s = pkg.peps.random_sites(t, 6)
    m0 = pkg.peps.ones_messages(t)
    for crit in ("fidelity", "maxabs"):
        _, sw, dl, at = orc.bp(t, s, m0, tol=1e300, criterion=crit)
        assert (sw, at) == (1, 1)
        assert dl > 0.0
    _, sw, dl, at = orc.bp(t, s, m0, maxiter=5)
    assert (sw, at) == (5, -1) and dl == math.inf
