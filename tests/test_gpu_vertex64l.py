"""The FP64 large-bond vertex sweep (vertex64l.h: largest bond dimension 9..16, padded to
16, three tiled passes on FP64 MFMA) against the FP64 oracle and the generic FP64 kernel.

Tolerances: the FP64 bar, 1e-12 relative for one sweep, 1e-10 for converged messages and
log-norms."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return float(np.max(np.abs(np.asarray(a) - np.asarray(b))) / np.max(np.abs(b)))


@pytest.mark.parametrize("shape,D,d", [((3, 3), 16, 2), ((3, 4), 10, 1), ((2, 3), 12, 3),
                                       ((3, 3), 9, 4), ((1, 3), 16, 2)])
def test_single_sweep(pkg, orc, shape, D, d):
    _say(f"case v64l sweep {shape} D={D} d={d}")
    g = pkg.named_grid(shape) if shape[0] > 1 else pkg.named_path_graph(shape[1])
    t = pkg.build_topology(g, D, d)
    s = pkg.peps.random_sites(t, 91, scale=pkg.peps.peps_scale(d, D))
    m = pkg.peps.random_messages(t, 91)
    want, fid_w, mab_w = orc.sweep_jacobi(t, s, m, nthreads=8)
    with pkg.Engine(t, s, dtype="f64") as e:
        assert e.path() == "vertex"
        e.set_messages(m)
        e.sweep()
        got = e.messages()
        fid, mab = e.residuals()
    assert _rel(got, want) < 1e-12
    assert np.allclose(mab, mab_w, rtol=1e-10, atol=1e-14)
    assert np.allclose(fid, fid_w, rtol=1e-8, atol=1e-13)


def test_mixed_bond_dims_and_low_degree(pkg, orc):
    """Bond dimensions 9..16 inside one comb tree (degree 1..3, padded ports)."""
    _say("case v64l mixed dims comb tree")
    g = pkg.named_comb_tree((3, 3))
    dims = {e: 9 + (i % 8) for i, e in enumerate(g.edges())}
    t = pkg.build_topology(g, dims, 2)
    s = pkg.peps.random_sites(t, 92, scale=0.05)
    m = pkg.peps.random_messages(t, 92)
    want, _, _ = orc.sweep_jacobi(t, s, m, nthreads=8)
    with pkg.Engine(t, s, dtype="f64") as e:
        assert e.path() == "vertex"
        e.set_messages(m)
        e.sweep()
        assert _rel(e.messages(), want) < 1e-12


def test_converged_matches_oracle_and_generic(pkg, orc):
    _say("case v64l converged 4x4 D=10")
    t = pkg.build_topology(pkg.named_grid((4, 4)), 10, 2)
    s = pkg.peps.random_sites(t, 93, scale=pkg.peps.peps_scale(2, 10))
    m0 = pkg.peps.identity_messages(t)
    want, sw, _, at = orc.bp(t, s, m0, maxiter=200, tol=1e-13, criterion="maxabs", nthreads=8)
    res = {}
    for path in ("auto", "generic"):
        with pkg.Engine(t, s, dtype="f64", path=path) as e:
            e.set_messages(m0)
            r = e.run(maxiter=200, tol=1e-13, criterion="maxabs")
            res[path] = (e.path(), r, e.messages(), e.lognorm())
    assert res["auto"][0] == "vertex" and res["generic"][0] == "generic"
    r = res["auto"][1]
    assert int(r.sweeps[0]) == sw and int(r.at_iteration[0]) == at
    assert _rel(res["auto"][2], want) < 1e-10
    lw = orc.bethe_free_energy(t, s, want)
    assert abs(res["auto"][3] - lw) < 1e-10 * abs(lw)
    assert _rel(res["auto"][2], res["generic"][2]) < 1e-12


def test_batch_masks_and_graph(pkg, orc):
    """Two networks (D = 16 and D = 11) in one FP64 context: per-network stopping, graph
    replay equal to eager sweeps, each against the oracle."""
    _say("case v64l batch")
    tops = [pkg.build_topology(pkg.named_grid((2, 3)), 16, 2),
            pkg.build_topology(pkg.named_grid((3, 3)), 11, 2)]
    sites = [pkg.peps.random_sites(t, 94 + i, scale=pkg.peps.peps_scale(2, t.edge_dim[0]))
             for i, t in enumerate(tops)]
    tb = pkg.concat_topologies(tops)
    out = {}
    for graph in (True, False):
        with pkg.Engine(tb, np.concatenate(sites), dtype="f64") as e:
            assert e.path() == "vertex"
            e.init_messages("identity")
            r = e.run(maxiter=300, tol=1e-12, criterion="maxabs", use_graph=graph)
            out[graph] = (r, e.messages(), e.lognorms())
    rg, mg, lg = out[True]
    re_, me, le = out[False]
    assert np.array_equal(rg.sweeps, re_.sweeps) and np.array_equal(mg, me)
    assert np.array_equal(lg, le)
    off = 0
    for i, (t, s) in enumerate(zip(tops, sites)):
        n = int(t.msg_off[-1])
        want, sw, _, at = orc.bp(t, s, pkg.peps.identity_messages(t), maxiter=300, tol=1e-12,
                                 criterion="maxabs", nthreads=8)
        assert int(rg.sweeps[i]) == sw and int(rg.at_iteration[i]) == at
        assert _rel(mg[off:off + n], want) < 1e-10
        off += n


def test_polish_matches_generic_fp64_sweeps(pkg):
    """The FP64 polish of a BF16 context at D = 16 (the large-bond kernel) equals three
    FP64 sweeps of the generic kernel from the same BF16 messages."""
    _say("case v64l polish")
    t = pkg.build_topology(pkg.named_grid((4, 4)), 16, 2)
    s = pkg.peps.random_sites(t, 95, scale=pkg.peps.peps_scale(2, 16))
    with pkg.Engine(t, s, dtype="bf16") as e:
        e.init_messages("identity")
        e.run(maxiter=50, tol=1e-5)
        m_bf = e.messages()
        e.polish_fp64(3)
        pol, ln = e.messages(), e.lognorm()
    with pkg.Engine(t, s, dtype="f64", path="generic") as e:
        e.set_messages(m_bf)
        e.run(maxiter=3)
        ref, lr = e.messages(), e.lognorm()
    assert _rel(pol, ref) < 1e-12
    assert ln == pytest.approx(lr, rel=1e-13)


@pytest.mark.parametrize("kind,D,d", [("grid", 16, 2), ("comb", 11, 3), ("path", 9, 1)])
def test_vertex_scalars_and_lognorm(pkg, orc, kind, D, d):
    """FP64 vertex scalars from the large-bond kernel's passes 1 and 2 (Z_u = <out_2, m_2>,
    e0 e0^T on a padded port 2: degrees 1..4) against the FP64 oracle and the generic FP64
    kernel (1e-12 relative), and the log-norm against the oracle's bethe_free_energy."""
    _say(f"case v64l vertex scalars {kind} D={D} d={d}")
    g = {"grid": lambda: pkg.named_grid((3, 4)), "comb": lambda: pkg.named_comb_tree((3, 3)),
         "path": lambda: pkg.named_path_graph(4)}[kind]()
    t = pkg.build_topology(g, D, d)
    s = pkg.peps.random_sites(t, 96, scale=pkg.peps.peps_scale(d, D))
    with pkg.Engine(t, s, dtype="f64", path="generic") as e:  # PSD messages: real logs
        e.init_messages("identity")
        e.run(maxiter=3)
        m = e.messages()
    res = {}
    for path in ("auto", "generic"):
        with pkg.Engine(t, s, dtype="f64", path=path) as e:
            e.set_messages(m)
            res[path] = (e.path(), e.vertex_scalars(), e.lognorm())
    assert res["auto"][0] == "vertex" and res["generic"][0] == "generic"
    want = orc.vertex_scalars(t, s, m)
    assert _rel(res["auto"][1], want) < 1e-12
    assert _rel(res["auto"][1], res["generic"][1]) < 1e-12
    lw = orc.bethe_free_energy(t, s, m)
    assert abs(res["auto"][2] - lw) < 1e-11 * max(1.0, abs(lw))


def test_lognorm_after_polish_uses_fp64_messages(pkg, orc):
    """After bp_polish_fp64 at D = 16 the log-norm comes from the FP64 messages on the
    large-bond kernel: equal (1e-12) to the log-norm of an FP64 context holding the same
    messages, and to the oracle's bethe_free_energy."""
    _say("case v64l lognorm after polish")
    t = pkg.build_topology(pkg.named_grid((4, 4)), 16, 2)
    s = pkg.peps.random_sites(t, 97, scale=pkg.peps.peps_scale(2, 16))
    with pkg.Engine(t, s, dtype="bf16") as e:
        e.init_messages("identity")
        e.run(maxiter=50, tol=1e-5)
        e.polish_fp64(3)
        pol, ln = e.messages(), e.lognorm()
    with pkg.Engine(t, s, dtype="f64", path="generic") as e:
        e.set_messages(pol)
        lg = e.lognorm()
    lw = orc.bethe_free_energy(t, s, pol)
    assert ln == pytest.approx(lg, rel=1e-12)
    assert ln == pytest.approx(lw, rel=1e-12)
