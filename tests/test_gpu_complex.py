"""Complex (ComplexF64) BP on the GPU against the complex restatement of the oracle
(oracle/complex_oracle.py), and the reference's complex tests on the reference-shaped API
(test/test_beliefpropagation.jl:84-99 and :153-202 with T = ComplexF64).

FP64 arithmetic throughout: stated tolerance 1e-12 relative for one sweep, 1e-10 for
converged messages and log-norms (the FP64 bar)."""
import cmath
import ctypes as C
import math

import numpy as np
import pytest

from oracle import complex_oracle as co

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _csites(t, seed):
    rng = np.random.default_rng(seed)
    n = int(t.site_off[-1])
    return (rng.standard_normal(n) + 1j * rng.standard_normal(n)) / math.sqrt(2 * 2 * 3 ** 4)


def _cmsgs(t, seed):
    rng = np.random.default_rng(seed)
    n = int(t.msg_off[-1])
    return rng.standard_normal(n) + 1j * rng.standard_normal(n)


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


def _norm_lattice(pkg, L=3, D=3, d=2):
    g = pkg.named_grid((L, L))
    return g, pkg.build_topology(g, D, d)


@pytest.mark.parametrize("plain", [False, True])
def test_one_sweep_matches_oracle(pkg, plain):
    _say(f"case complex one sweep plain={plain}")
    g = pkg.named_grid((3, 4))
    t = pkg.build_topology(g, 3, plain=True) if plain else pkg.build_topology(g, 3, 2)
    s, m0 = _csites(t, 1), _cmsgs(t, 2)
    with pkg.Engine(t, s, dtype="c128") as e:
        assert e.path() == "generic"
        e.set_messages(m0)
        e.sweep()
        got = e.messages()
        fid, mab = e.residuals()
    want = co.sweep_jacobi(t, s, m0)
    assert got.dtype == np.complex128
    assert _rel(got, want) < 1e-12
    for q in range(t.nports):
        a = m0[t.msg_off[q]:t.msg_off[q + 1]]
        b = want[t.msg_off[q]:t.msg_off[q + 1]]
        assert fid[q] == pytest.approx(co._fidelity(a, b), abs=1e-12)
        assert mab[q] == pytest.approx(np.max(np.abs(a - b)), rel=1e-11)


def test_unnormalised_sweep_matches_oracle(pkg):
    _say("case complex normalize=false")
    g, t = _norm_lattice(pkg)
    s, m0 = _csites(t, 3), _cmsgs(t, 4)
    with pkg.Engine(t, s, dtype="c128") as e:
        e.set_option("normalize", 0)
        e.set_messages(m0)
        e.sweep()
        got = e.messages()
    assert _rel(got, co.sweep_jacobi(t, s, m0, normalize=False)) < 1e-12


def test_converged_run_scalars_and_lognorm(pkg):
    _say("case complex converged run")
    g, t = _norm_lattice(pkg, L=4, D=2)
    s = _csites(t, 5)
    m0 = pkg.peps.identity_messages(t)
    mo, sw, dl, at = co.bp(t, s, m0, maxiter=200, tol=1e-13)
    with pkg.Engine(t, s, dtype="c128") as e:
        e.set_messages(m0)
        r = e.run(maxiter=200, tol=1e-13)
        assert int(r.sweeps[0]) == sw and int(r.at_iteration[0]) == at
        assert float(r.delta[0]) == pytest.approx(dl, rel=1e-3, abs=1e-15)
        got = e.messages()
        zv, ze = e.vertex_scalars(), e.edge_scalars()
        ln = e.lognorm()
    assert _rel(got, mo) < 1e-10
    assert _rel(zv, co.vertex_scalars(t, s, mo)) < 1e-10
    assert _rel(ze, co.edge_scalars(t, mo)) < 1e-10
    want = co.bethe_free_energy(t, s, mo)
    assert isinstance(ln, complex)
    assert abs(ln.real - want.real) < 1e-10 * abs(want.real)
    assert abs(cmath.exp(1j * ln.imag) - cmath.exp(1j * want.imag)) < 1e-10
    # the converged messages of a norm network are Hermitian
    for q in range(t.nports):
        M = got[t.msg_off[q]:t.msg_off[q + 1]].reshape(t.msg_shape(q))
        assert np.allclose(M, M.conj().T, atol=1e-12)


def test_sequential_run_matches_oracle(pkg):
    _say("case complex sequential")
    g, t = _norm_lattice(pkg, L=3, D=2)
    s = _csites(t, 6)
    order = [t.slot(a, b) for a, b in pkg.forest_cover_edge_sequence(g)]
    m0 = pkg.peps.identity_messages(t)
    mo, sw, dl, at = co.bp(t, s, m0, maxiter=100, tol=1e-12, schedule="sequential", order=order)
    with pkg.Engine(t, s, dtype="c128", path="generic") as e:
        e.set_messages(m0)
        r = e.run_sequential(order, maxiter=100, tol=1e-12)
        assert int(r.sweeps[0]) == sw and int(r.at_iteration[0]) == at
        assert _rel(e.messages(), mo) < 1e-10


def test_iterate_diff_and_noalloc(pkg):
    _say("case complex iterate_diff")
    g, t = _norm_lattice(pkg)
    s, a, b = _csites(t, 7), _cmsgs(t, 8), _cmsgs(t, 9)
    with pkg.Engine(t, s, dtype="c128") as e:
        e.set_messages(a)
        d = e.iterate_diff(b)
        d_same = e.iterate_diff(np.exp(0.3j) * a)  # a global phase leaves the fidelity at 0
        # the entry point without allocation, through the C ABI (interleaved doubles)
        bb = np.ascontiguousarray(b, dtype=np.complex128)
        out = C.c_double(0)
        pkg._lib.check(pkg._lib.load().bp_iterate_diff_noalloc(
            e.h, bb.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), C.byref(out)),
            "bp_iterate_diff_noalloc")
    assert d == pytest.approx(co.iterate_diff(t, a, b), abs=1e-12)
    assert abs(d_same) < 1e-13
    assert out.value == d


def test_batch_masks_and_graph_replay(pkg):
    """Two complex networks of different shapes in one context: per-network convergence
    and the cached sweep graph agree with eager sweeps and with the oracle per network."""
    _say("case complex batch")
    t1 = pkg.build_topology(pkg.named_grid((3, 3)), 2, 2)
    t2 = pkg.build_topology(pkg.named_grid((2, 4)), 3, 2)
    tb = pkg.concat_topologies([t1, t2])
    s1, s2 = _csites(t1, 10), _csites(t2, 11)
    sb = np.concatenate([s1, s2])
    res = {}
    for graph in (True, False):
        with pkg.Engine(tb, sb, dtype="c128") as e:
            e.init_messages("identity")
            r = e.run(maxiter=300, tol=1e-12, use_graph=graph)
            res[graph] = (r, e.messages(), e.lognorms())
    rg, mg, lg = res[True]
    re_, me, le = res[False]
    assert np.array_equal(rg.sweeps, re_.sweeps) and np.array_equal(mg, me)
    assert np.array_equal(lg, le)
    off = int(t1.msg_off[-1])
    for k, (t, s, m) in enumerate(((t1, s1, mg[:off]), (t2, s2, mg[off:]))):
        mo, sw, _, at = co.bp(t, s, pkg.peps.identity_messages(t), maxiter=300, tol=1e-12)
        assert int(rg.sweeps[k]) == sw and int(rg.at_iteration[k]) == at
        assert _rel(m, mo) < 1e-10
        want = co.bethe_free_energy(t, s, mo)
        assert abs(lg[k].real - want.real) < 1e-10 * abs(want.real)


def test_unsupported_entry_points_refuse(pkg):
    _say("case complex refusals")
    g, t = _norm_lattice(pkg, L=2, D=2)
    with pkg.Engine(t, _csites(t, 12), dtype="c128") as e:
        e.init_messages("identity")
        for call in (lambda: e.grad_lognorm(), lambda: e.polish_fp64(1),
                     lambda: e.tfim_energy(1.0, 3.0), lambda: e.set_path("vertex")):
            with pytest.raises(Exception, match="complex"):
                call()
        e.sweep()  # the context still works


# ------------------------------------------------------------------ the reference's tests

def _links(pkg, g, dim):
    return {frozenset(e): pkg.Index(dim) for e in g.edges()}


def _exact_plain(tn, g):
    letters = iter("abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ")
    name, subs, ops = {}, [], []
    for v in g.vertices():
        t = tn[v]
        subs.append("".join(name.setdefault(i, next(letters)) for i in t.inds))
        ops.append(t.data)
    return complex(np.einsum(",".join(subs) + "->", *ops))


def test_reference_algorithm_complexf64_chain_and_tree(pkg):
    """test_beliefpropagation.jl:153-202 with T = ComplexF64: ones(T) messages, maxiter = 1,
    z_bp ~ z_exact with rtol eps(Float64)^(1/3)."""
    _say("case complex api chain/tree")
    rng = np.random.default_rng(123)
    rtol = np.finfo(np.float64).eps ** (1 / 3)
    for g, dim in ((pkg.named_path_graph(2), 2), (pkg.named_comb_tree((4, 3)), 3)):
        l = _links(pkg, g, dim)

        def tensor(v):
            inds = [l[frozenset(e)] for e in g.incident_edges(v)]
            shape = tuple(i.dim for i in inds)
            return pkg.ITensor(rng.standard_normal(shape) + 1j * rng.standard_normal(shape), inds)

        tn = pkg.tensornetwork(tensor, g.vertices())
        msgs = {e: np.ones(int(np.prod([i.dim for i in tn.linkinds(e)])), dtype=np.complex128)
                for e in g.all_edges()}
        cache = pkg.beliefpropagation(tn, msgs, stopping_criterion={"maxiter": 1})
        z_bp = cmath.exp(pkg.bethe_free_energy(tn, cache))
        z_exact = _exact_plain(tn, g)
        assert abs(z_bp - z_exact) <= rtol * abs(z_exact)
        assert cache.state.iteration == 1
        assert all(np.iscomplexobj(m) for m in cache.values())


def test_reference_scalars_complexf32_tensors(pkg):
    """test_beliefpropagation.jl:84-99: ComplexF32 tensors with Float64 ones messages give a
    complex vertex scalar and a real edge scalar; region scalars are products of vertex
    scalars."""
    _say("case complex api scalars")
    g = pkg.named_path_graph(3)
    l = _links(pkg, g, 2)
    rng = np.random.default_rng(1)

    def tensor(v):
        inds = [l[frozenset(e)] for e in g.incident_edges(v)]
        shape = tuple(i.dim for i in inds)
        return pkg.ITensor((rng.standard_normal(shape) + 1j * rng.standard_normal(shape))
                           .astype(np.complex64), inds)

    tn = pkg.tensornetwork(tensor, g.vertices())
    bpc = pkg.messagecache(lambda e: np.ones(2), g.all_edges())
    z2 = pkg.vertex_scalar(tn, bpc, 2)
    assert isinstance(z2, complex)
    assert isinstance(pkg.edge_scalar(bpc, (1, 2)), float)
    assert pkg.region_scalar(tn, bpc, [1]) == pkg.vertex_scalar(tn, bpc, 1)
    zs = pkg.vertex_scalars(tn, bpc, [2, 3])
    assert pkg.region_scalar(tn, bpc, [2, 3]) == pytest.approx(complex(np.prod(zs)), rel=1e-15)
    # the vertex scalar of vertex 2 is its tensor summed over both (ones) links
    assert z2 == pytest.approx(complex(np.sum(tn[2].data.astype(np.complex128))), rel=1e-12)


def test_reference_norm_network_complex_bp_exact_on_tree(pkg):
    """A ComplexF64 PEPS on a comb tree: BP on its norm network gives <psi|psi> (real,
    positive) after one sequential sweep from identity messages."""
    _say("case complex api norm tree")
    g = pkg.named_comb_tree((3, 2))
    l = _links(pkg, g, 2)
    sidx = {v: pkg.Index(2) for v in g.vertices()}
    rng = np.random.default_rng(4)

    def tensor(v):
        inds = [sidx[v]] + [l[frozenset(e)] for e in g.incident_edges(v)]
        shape = tuple(i.dim for i in inds)
        return pkg.ITensor(rng.standard_normal(shape) + 1j * rng.standard_normal(shape), inds)

    tn = pkg.tensornetwork(tensor, g.vertices())
    nn = pkg.NormNetwork(tn)
    env = pkg.message_environment(pkg.one, nn)
    assert all(np.iscomplexobj(m) for m in env.values())
    cache = pkg.beliefpropagation(nn, env, stopping_criterion={"maxiter": 1})
    f = pkg.bethe_free_energy(nn, cache)
    # exact <psi|psi>
    letters = iter("abcdefghijklmnopqrstuvwxyzABCDEFGHIJKLMNOPQRSTUVWXYZ")
    name, subs, ops, outs = {}, [], [], ""
    for v in g.vertices():
        t = tn[v]
        subs.append("".join(name.setdefault(i, next(letters)) for i in t.inds))
        ops.append(t.data)
        outs += name[sidx[v]]
    psi = np.einsum(",".join(subs) + "->" + outs, *ops)
    exact = float(np.vdot(psi, psi).real)
    assert abs(math.exp(f.real) / exact - 1) < 1e-11
    assert abs(cmath.exp(1j * f.imag) - 1) < 1e-9
