"""The FP64 vertex kernel (bond dimensions <= 8, k_vertex_d8_f64) against the FP64 oracle.

FP64 runs keep the 1e-10 bar: one Jacobi sweep from random non-symmetric messages agrees
entry by entry to 1e-12 relative; converged runs agree with the oracle's messages and
log-norm to 1e-10 and report the oracle's sweep counts.  Shapes cover corners, edges and
interior vertices, physical dimensions 1..4, bond dimensions below 8 (zero padding),
mixed bond dimensions, batches, and the FP64 polish of an FP32 solve.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("shape,d,D", [((6, 5), 2, 8), ((4, 4), 1, 8), ((3, 5), 3, 6),
                                       ((5, 4), 4, 7), ((7, 7), 2, 8)])
def test_fp64_vertex_single_sweep(pkg, orc, shape, d, D):
    _say(f"case fp64 vertex single sweep {shape} d={d} D={D}")
    t = pkg.build_topology(pkg.named_grid(shape), D, d)
    s = pkg.peps.random_sites(t, 21, scale=pkg.peps.peps_scale(d, D))
    m = pkg.peps.random_messages(t, 21)
    want, fid, mab = orc.sweep_jacobi(t, s, m)
    with pkg.Engine(t, s, dtype="f64") as e:
        assert e.path() == "vertex"
        e.set_messages(m)
        e.sweep()
        got = e.messages()
        gfid, gmab = e.residuals()
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15 * np.max(np.abs(want)))
    np.testing.assert_allclose(gmab, mab, rtol=1e-9, atol=1e-15)


def test_fp64_vertex_mixed_bond_dimensions(pkg, orc):
    _say("case fp64 vertex mixed bond dimensions")
    g = pkg.named_grid((5, 6))
    dims = {e: 6 + (i % 3) for i, e in enumerate(g.edges())}
    t = pkg.build_topology(g, dims, 2)
    s = pkg.peps.random_sites(t, 22, scale=pkg.peps.peps_scale(2, 7))
    m = pkg.peps.random_messages(t, 22)
    want, _, _ = orc.sweep_jacobi(t, s, m)
    with pkg.Engine(t, s, dtype="f64") as e:
        assert e.path() == "vertex"
        e.set_messages(m)
        e.sweep()
        got = e.messages()
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-15 * np.max(np.abs(want)))


def test_fp64_vertex_converged_and_counts(pkg, orc):
    _say("case fp64 vertex converged 10x10 D=8")
    t = pkg.build_topology(pkg.named_grid((10, 10)), 8, 2)
    s = pkg.peps.random_sites(t, 23, scale=pkg.peps.peps_scale(2, 8))
    m0 = pkg.peps.identity_messages(t)
    want, sw, dl, at = orc.bp(t, s, m0, maxiter=200, tol=1e-13, criterion="maxabs", nthreads=8)
    assert at > 0
    with pkg.Engine(t, s, dtype="f64") as e:
        assert e.path() == "vertex"
        e.set_messages(m0)
        r = e.run(maxiter=200, tol=1e-13, criterion="maxabs")
        got = e.messages()
        ln = e.lognorm()
    _say(f"sweeps gpu={r.sweeps[0]} oracle={sw} err={np.max(np.abs(got - want)):.2e}")
    assert (r.sweeps[0], r.at_iteration[0]) == (sw, at)
    assert np.max(np.abs(got - want)) < 1e-10
    assert abs(ln - orc.bethe_free_energy(t, s, want)) < 1e-10


def test_fp64_vertex_small_bond_when_enabled(pkg, orc):
    """D = 2 runs on the generic kernel by default; with the mask bit set it runs on the
    FP64 vertex kernel (zero padding to 8) with the same fixed point (4x4 d=2 D=2)."""
    _say("case fp64 vertex D=2 via mask")
    t = pkg.build_topology(pkg.named_grid((4, 4)), 2, 2)
    s = pkg.peps.random_sites(t, 1234)
    m0 = pkg.peps.identity_messages(t)
    want, sw, dl, at = orc.bp(t, s, m0, maxiter=100, tol=1e-10, criterion="maxabs")
    with pkg.Engine(t, s, dtype="f64") as e:
        assert e.path() == "generic"
        e.set_option("vertex_bond_mask", 0x1FE)
        assert e.path() == "vertex"
        e.set_messages(m0)
        r = e.run(maxiter=100, tol=1e-10, criterion="maxabs")
        assert (r.sweeps[0], r.at_iteration[0]) == (sw, at)
        assert np.max(np.abs(e.messages() - want)) < 1e-10
        assert abs(e.lognorm() - orc.bethe_free_energy(t, s, want)) < 1e-10


def test_fp64_vertex_batch(pkg, orc):
    _say("case fp64 vertex batch")
    tops = [pkg.build_topology(pkg.named_grid((L, L)), D, 2) for L, D in ((4, 8), (5, 6), (3, 7))]
    sites = [pkg.peps.random_sites(t, 70 + i, scale=pkg.peps.peps_scale(2, int(t.edge_dim[0])))
             for i, t in enumerate(tops)]
    b = pkg.concat_topologies(tops)
    with pkg.Engine(b, np.concatenate(sites), dtype="f64") as e:
        assert e.path() == "vertex"
        e.init_messages("identity")
        r = e.run(maxiter=300, tol=1e-12, criterion="maxabs")
        mb = e.messages()
        ln = e.lognorms()
    off = 0
    for i, (t, s) in enumerate(zip(tops, sites)):
        want, sw, dl, at = orc.bp(t, s, pkg.peps.identity_messages(t), maxiter=300, tol=1e-12,
                                  criterion="maxabs")
        n = int(t.msg_off[-1])
        assert (r.sweeps[i], r.at_iteration[i]) == (sw, at)
        np.testing.assert_allclose(mb[off:off + n], want, atol=1e-10)
        assert ln[i] == pytest.approx(orc.bethe_free_energy(t, s, want), abs=1e-10)
        off += n


def test_fp64_vertex_matches_generic_and_graphs(pkg):
    """The FP64 vertex path and the FP64 generic kernel reach the same messages (1e-13),
    eager and replayed from a captured graph."""
    _say("case fp64 vertex vs generic")
    t = pkg.build_topology(pkg.named_grid((8, 8)), 8, 2)
    s = pkg.peps.random_sites(t, 24, scale=pkg.peps.peps_scale(2, 8))
    out = {}
    for path, graph in (("auto", True), ("auto", False), ("generic", True)):
        with pkg.Engine(t, s, dtype="f64", path=path) as e:
            e.init_messages("identity")
            r = e.run(maxiter=40, use_graph=graph)
            out[(path, graph)] = (e.messages(), e.path())
    assert out[("auto", True)][1] == "vertex" and out[("generic", True)][1] == "generic"
    np.testing.assert_array_equal(out[("auto", True)][0], out[("auto", False)][0])
    assert _rel(out[("auto", True)][0], out[("generic", True)][0]) < 1e-13


def test_fp64_polish_on_vertex_kernel(pkg, orc):
    """FP32 solve, then FP64 polishing sweeps on the FP64 vertex kernel: messages converge
    to the oracle's fixed point."""
    _say("case fp64 polish via vertex kernel 12x12 D=8")
    t = pkg.build_topology(pkg.named_grid((12, 12)), 8, 2)
    s = pkg.peps.random_sites(t, 25, scale=pkg.peps.peps_scale(2, 8))
    m0 = pkg.peps.identity_messages(t)
    want, _, _, at = orc.bp(t, s, m0, maxiter=300, tol=1e-15, criterion="maxabs", nthreads=8)
    assert at > 0
    lw = orc.bethe_free_energy(t, s, want)
    errs = []
    for path in ("auto", "generic"):
        with pkg.Engine(t, s, dtype="f32", path=path) as e:
            e.set_messages(m0)
            assert e.run(maxiter=300, tol=1e-12).converged[0]
            e.polish_fp64(12)
            errs.append((_rel(e.messages(), want), abs(e.lognorm() - lw) / abs(lw)))
    _say(f"polished errors vertex={errs[0]} generic={errs[1]}")
    for em, el in errs:
        assert em < 1e-10 and el < 1e-12
