"""Pins of the complex (ComplexF64) restatement in oracle/complex_oracle.py.  CPU only.

Reference tests mirrored (ITensorNetworksNext.jl test/):
  test_beliefpropagation.jl:153-202, T = ComplexF64  BP after one sweep on a chain / comb
                                                      tree equals the exact contraction
  test_normnetwork.jl:148-166, T = ComplexF64        <psi|psi> = ||psi||^2, real and positive;
                                                      a normalised tensor contracts to 1
What the mathematics fixes besides: on real data the restatement equals the pinned real
oracle (bp_oracle.c); the bra layer must be conjugated (dropping it breaks tree exactness);
iterate_diff conjugates its first argument and ignores a global phase; converged norm-network
messages are Hermitian.
"""
import cmath
import math

import numpy as np
import pytest

from oracle import complex_oracle as co


def _csites(t, seed, scale=1.0):
    rng = np.random.default_rng(seed)
    n = int(t.site_off[-1])
    return scale * (rng.standard_normal(n) + 1j * rng.standard_normal(n))


def _tree(pkg, dims):
    return pkg.named_comb_tree(dims) if dims[1] > 1 else pkg.named_path_graph(dims[0])


def _order(pkg, g, t):
    return [t.slot(a, b) for a, b in pkg.forest_cover_edge_sequence(g)]


# ---------------------------------------------------------------- equal to the real oracle

@pytest.mark.parametrize("plain", [False, True])
def test_real_data_matches_real_oracle(pkg, orc, plain):
    g = pkg.named_grid((3, 3))
    t = pkg.build_topology(g, 2, plain=True) if plain else pkg.build_topology(g, 3, 2)
    s = pkg.peps.random_sites(t, 4)
    m0 = pkg.peps.random_messages(t, 7)
    for q in range(t.nports):
        a = co.message(t, s, m0, q, normalize=False)
        b = orc.message(t, s, m0, q)
        assert np.allclose(a.real, b, rtol=1e-12, atol=1e-14) and not np.any(a.imag)
    jr, _, _ = orc.sweep_jacobi(t, s, m0)
    assert np.allclose(co.sweep_jacobi(t, s, m0).real, jr, rtol=1e-12, atol=1e-14)
    order = _order(pkg, g, t)
    assert np.allclose(co.sweep_sequential(t, s, m0, order).real,
                       orc.sweep_sequential(t, s, m0, order), rtol=1e-11, atol=1e-13)
    mr, swr, dlr, atr = orc.bp(t, s, m0, maxiter=40, tol=1e-12)
    mc, swc, dlc, atc = co.bp(t, s, m0, maxiter=40, tol=1e-12)
    assert (swc, atc) == (swr, atr)
    assert dlc == pytest.approx(dlr, rel=1e-6, abs=1e-15)
    assert np.allclose(mc.real, mr, rtol=1e-10, atol=1e-13)
    assert np.allclose(co.vertex_scalars(t, s, mr).real, orc.vertex_scalars(t, s, mr), rtol=1e-11)
    assert np.allclose(co.edge_scalars(t, mr).real, orc.edge_scalars(t, mr), rtol=1e-11)
    assert co.iterate_diff(t, m0, mr) == pytest.approx(orc.iterate_diff(t, m0, mr), rel=1e-10)
    # random non-PSD messages give negative terms: the principal logarithm of a negative
    # real is log|x| + i pi, which is the real oracle's complex promotion
    fr = complex(orc.bethe_free_energy(t, s, m0))
    fc = co.bethe_free_energy(t, s, m0)
    assert fc.real == pytest.approx(fr.real, rel=1e-11)
    assert cmath.exp(1j * fc.imag) == pytest.approx(cmath.exp(1j * fr.imag), abs=1e-9)


# ---------------------------------------------------------------- tree exactness, complex

@pytest.mark.parametrize("dims,D", [((2, 1), 2), ((4, 3), 3)])
def test_plain_complex_tree_exact(pkg, dims, D):
    """test_beliefpropagation.jl:153-202 with T = ComplexF64: ones messages, one sweep."""
    g = _tree(pkg, dims)
    t = pkg.build_topology(g, D, plain=True)
    s = _csites(t, 21)
    m = co.sweep_sequential(t, s, pkg.peps.ones_messages(t), _order(pkg, g, t))
    z_bp = cmath.exp(co.bethe_free_energy(t, s, m))
    z = co.plain_exact(t, s)
    assert abs(z_bp - z) <= 1e-11 * abs(z)
    # and through the loop of beliefpropagation with maxiter = 1
    mb, sw, _, _ = co.bp(t, s, pkg.peps.ones_messages(t), maxiter=1, schedule="sequential",
                         order=_order(pkg, g, t))
    assert sw == 1 and np.allclose(mb, m, rtol=0, atol=0)


@pytest.mark.parametrize("dims,D", [((2, 1), 2), ((4, 3), 3)])
def test_norm_complex_tree_exact(pkg, dims, D):
    g = _tree(pkg, dims)
    t = pkg.build_topology(g, D, 2)
    s = _csites(t, 5)
    exact = co.norm_exact(t, s)
    m = co.sweep_sequential(t, s, pkg.peps.ones_messages(t), _order(pkg, g, t))
    f = co.bethe_free_energy(t, s, m)
    assert abs(math.exp(f.real) / exact - 1) < 1e-11
    assert abs(cmath.exp(1j * f.imag) - 1) < 1e-9  # Z_BP is real and positive
    # Jacobi to convergence reaches the same value, with Hermitian messages
    mj, _, _, at = co.bp(t, s, pkg.peps.identity_messages(t), maxiter=60, tol=1e-15,
                         criterion="maxabs")
    assert at > 0
    assert abs(math.exp(co.bethe_free_energy(t, s, mj).real) / exact - 1) < 1e-11
    for q in range(t.nports):
        M = mj[t.msg_off[q]:t.msg_off[q + 1]].reshape(t.msg_shape(q))
        assert np.allclose(M, M.conj().T, atol=1e-12)


def test_dropping_the_bra_conjugate_breaks_exactness(pkg, monkeypatch):
    """A planted mistake in the oracle (bra = A instead of conj(A)) must fail the tree pin."""
    g = pkg.named_comb_tree((3, 2))
    t = pkg.build_topology(g, 2, 2)
    s = _csites(t, 9)
    exact = co.norm_exact(t, s)
    order = _order(pkg, g, t)
    monkeypatch.setattr(co.np, "conj", lambda x: x)
    m = co.sweep_sequential(t, s, pkg.peps.ones_messages(t), order)
    f = co.bethe_free_energy(t, s, m)
    monkeypatch.undo()
    assert abs(cmath.exp(f) / exact - 1) > 1e-3


def test_norm_exact_pins(pkg):
    """test_normnetwork.jl:148-166, ComplexF64: <psi|psi> = ||prod||^2 (here the state vector
    contracted by pairwise tensordots in a fixed order), real and positive; a normalised
    single tensor gives 1."""
    g = pkg.named_grid((2, 2))
    t = pkg.build_topology(g, 2, 2)
    s = _csites(t, 3)
    A = [co._site(t, s, u) for u in range(4)]
    # ports of each vertex in CSR order; contract edge by edge
    lab = {}
    for u in range(4):
        b = int(t.port_off[u])
        lab[u] = ["p%d" % u] + ["e%d" % int(t.port_edge[b + j]) for j in range(int(t.port_off[u + 1]) - b)]
    T, L = A[0], list(lab[0])
    for u in (1, 2, 3):
        sh = [x for x in L if x in lab[u]]
        T = np.tensordot(T, A[u], axes=([L.index(x) for x in sh], [lab[u].index(x) for x in sh]))
        L = [x for x in L if x not in sh] + [x for x in lab[u] if x not in sh]
    assert co.norm_exact(t, s) == pytest.approx(float(np.sum(np.abs(T) ** 2)), rel=1e-12)
    one = pkg.build_topology(pkg.named_path_graph(1), 1, 3)
    v = _csites(one, 4)
    assert co.norm_exact(one, v / np.linalg.norm(v)) == pytest.approx(1.0, rel=1e-14)


# ---------------------------------------------------------------- iterate_diff

def test_iterate_diff_complex_closed_forms(pkg):
    t = pkg.build_topology(pkg.named_path_graph(2), 2, plain=True)
    a = np.array([1.0, 1j, 1.0, 1j])                      # both directed messages (1, i)
    assert co.iterate_diff(t, a, a) == pytest.approx(0.0, abs=1e-15)   # dot without conj: 1
    assert co.iterate_diff(t, a, np.exp(0.7j) * a) == pytest.approx(0.0, abs=1e-15)
    b = np.array([1.0, -1j, 1.0, -1j])                    # orthogonal to (1, i)
    assert co.iterate_diff(t, a, b) == pytest.approx(1.0, abs=1e-15)
    c = np.array([1.0, 0.0, 1.0, 1.0 + 0j])               # (1, 0) vs (1, i): |<,>|^2 = 1/2
    d = np.array([1.0, 1j, 1.0, 1.0 + 0j])
    assert co.iterate_diff(t, c, d) == pytest.approx(0.5, abs=1e-15)
