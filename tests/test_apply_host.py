"""Host logic of gate application (no GPU): operator matrices, truncation parameters, the
gate descriptors, and the argument errors raised before any device call
(apply_operators.jl:208-228)."""
import numpy as np
import pytest


def _net(pkg, g, D=2, d=2, seed=0):
    rng = np.random.default_rng(seed)
    links = {frozenset(e): pkg.Index(D) for e in g.edges()}
    site = {v: pkg.Index(d) for v in g.vertices()}
    tn = pkg.tensornetwork(lambda v: pkg.random_itensor(
        [site[v]] + [links[frozenset(e)] for e in g.incident_edges(v)], rng), g.vertices())
    return tn, site


def test_operator_matrix_order(pkg):
    a, b = pkg.Index(2), pkg.Index(3)
    ap, bp = pkg.uniquename(a), pkg.uniquename(b)
    data = np.random.default_rng(1).standard_normal((2, 3, 2, 3))
    op = pkg.operator(data, [ap, bp], [a, b])
    np.testing.assert_array_equal(op.matrix([a, b]), data.reshape(6, 6))
    np.testing.assert_array_equal(op.matrix([b, a]),
                                  np.transpose(data, (1, 0, 3, 2)).reshape(6, 6))
    with pytest.raises(ValueError):
        pkg.operator(data, [ap], [a, b])


def test_truncation_parameters(pkg):
    tp = pkg.apply._trunc_params
    assert tp(None) == (0, -1.0)
    assert tp(pkg.truncrank(4)) == (4, -1.0)
    assert tp(pkg.trunctol(rtol=1e-3)) == (0, 1e-3)
    assert tp(pkg.apply.TruncBoth(pkg.truncrank(3), pkg.trunctol(rtol=0.1))) == (3, 0.1)
    with pytest.raises(pkg.ArgumentError):
        pkg.trunctol(atol=1e-3)


def test_gate_descriptor(pkg):
    g = pkg.named_grid((2, 3))
    tn, site = _net(pkg, g, D=3, d=2)
    env = pkg.message_environment(pkg.one, pkg.normnetwork(tn))
    op = pkg.operator(np.eye(4).reshape(2, 2, 2, 2),
                      [pkg.uniquename(site[(1, 2)]), pkg.uniquename(site[(2, 2)])],
                      [site[(1, 2)], site[(2, 2)]])
    gt = pkg.apply._Gate(op, tn, env, maxdim=5, normalize=True)
    d = gt.desc
    assert list(d[:3]) == [2, 2, 2]
    assert d[3] == 3 and d[4] == 3  # both vertices have three neighbours
    nb1, nb2 = gt.layouts[0][1], gt.layouts[1][1]
    assert nb1[d[5]] == (2, 2) and nb2[d[6]] == (1, 2)
    assert d[7] == 5 and d[8] == 1
    sizes = np.zeros(1, dtype=np.int64)
    # packed inputs: two tensors (2*27 each), two messages per site (9 each), the gate (16)
    assert gt.data.size == 2 * 54 + 4 * 9 + 16


def test_gate_argument_errors(pkg):
    g = pkg.named_path_graph(4)
    tn, site = _net(pkg, g)
    env = pkg.message_environment(pkg.one, pkg.normnetwork(tn))

    def op_on(inds):
        n = len(inds)
        dims = [i.dim for i in inds]
        return pkg.operator(np.zeros(dims + dims), [pkg.uniquename(i) for i in inds], inds)

    with pytest.raises(pkg.ArgumentError, match="3-site"):
        pkg.apply_operator(op_on([site[1], site[2], site[3]]), tn, env)
    with pytest.raises(pkg.ArgumentError, match="no indices"):
        pkg.apply_operator(op_on([pkg.Index(2)]), tn, env)
    with pytest.raises(pkg.ArgumentError, match="neighbouring"):
        pkg.apply_operator(op_on([site[1], site[3]]), tn, env)
    with pytest.raises(pkg.ArgumentError, match="not allowed"):
        pkg.apply_operator(op_on([site[1]]), tn, env, alg=pkg.BPApplyGate(), normalize=True)


def test_complex_element_types_are_refused(pkg):
    """The gate kernel is real: a complex gate or a complex site tensor raises before any
    device call (never a silent cast that drops the imaginary part)."""
    g = pkg.named_path_graph(3)
    tn, site = _net(pkg, g)
    env = pkg.message_environment(pkg.one, pkg.normnetwork(tn))
    d = site[2].dim
    cgate = pkg.operator(np.eye(d, dtype=np.complex128) * 1j, [pkg.uniquename(site[2])], [site[2]])
    with pytest.raises(pkg.ArgumentError, match="complex"):
        pkg.apply_operator(cgate, tn, env)
    tc = tn.copy()
    tc[2] = pkg.ITensor(tn[2].data.astype(np.complex128), tn[2].inds)
    rgate = pkg.operator(np.eye(d), [pkg.uniquename(site[2])], [site[2]])
    with pytest.raises(pkg.ArgumentError, match="complex"):
        pkg.apply_operator(rgate, tc, env)
