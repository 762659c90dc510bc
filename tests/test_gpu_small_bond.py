"""The small-bond sweep kernel (k_vertex_s<DP>, FP32 networks whose bond dimensions are all
<= 4) against the FP64 oracle and against the padded D <= 8 kernel it replaces.

Tolerances: one sweep from random non-symmetric messages within the FP32 error bound of
tests/helpers.py; from identity messages within 1e-5 relative; converged runs within 1e-4
(messages) and 1e-5 (log-norm) of the FP64 oracle, as the other FP32 kernels."""
import numpy as np
import pytest

from helpers import fp32_sweep_check

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.mark.parametrize("shape,D,d", [((5, 4), 4, 2), ((4, 4), 3, 2), ((3, 5), 2, 3),
                                       ((4, 3), 4, 1), ((3, 3), 1, 2), ((4, 5), 3, 4),
                                       ((2, 2), 4, 3), ((1, 3), 2, 2)])
def test_small_single_sweep(pkg, orc, shape, D, d):
    _say(f"case small-bond sweep {shape} D={D} d={d}")
    g = pkg.named_grid(shape) if shape[0] > 1 else pkg.named_path_graph(shape[1])
    t = pkg.build_topology(g, D, d)
    s = pkg.peps.random_sites(t, 71, scale=pkg.peps.peps_scale(d, D))
    m = pkg.peps.random_messages(t, 71)
    with pkg.Engine(t, s, dtype="f32") as e:
        assert e.path() == "vertex" and e.get_option("small_bond") == 1
        e.set_messages(m)
        e.sweep()
        got = e.messages()
        fid, mab = e.residuals()
    assert fp32_sweep_check(pkg, orc, t, s, m, got) < 1.0
    # residuals of the sweep: the oracle's distances between the input and the result
    for q in range(t.nports):
        a = m[t.msg_off[q]:t.msg_off[q + 1]]
        b = got[t.msg_off[q]:t.msg_off[q + 1]].astype(np.float64)
        assert mab[q] == pytest.approx(np.max(np.abs(a.astype(np.float32) - b)), rel=1e-6, abs=1e-7)
    mi = pkg.peps.identity_messages(t)
    want_i, _, _ = orc.sweep_jacobi(t, s, mi)
    with pkg.Engine(t, s, dtype="f32") as e:
        e.set_messages(mi)
        e.sweep()
        assert _rel(e.messages(), want_i) < 1e-5


@pytest.mark.parametrize("kind,D,d", [("comb", 3, 2), ("heavyhex", 4, 2), ("heavyhex", 2, 1)])
def test_small_low_degree(pkg, orc, kind, D, d):
    """Degree <= 3 networks (trees, heavy-hex) pad their missing ports."""
    _say(f"case small-bond low degree {kind} D={D} d={d}")
    g = pkg.named_comb_tree((4, 3)) if kind == "comb" else pkg.heavy_hex_lattice(2, 3)
    t = pkg.build_topology(g, D, d)
    s = pkg.peps.random_sites(t, 73, scale=pkg.peps.peps_scale(d, D))
    m = pkg.peps.random_messages(t, 73)
    with pkg.Engine(t, s, dtype="f32") as e:
        assert e.path() == "vertex"
        e.set_messages(m)
        e.sweep()
        got = e.messages()
    assert fp32_sweep_check(pkg, orc, t, s, m, got) < 1.0


def test_small_mixed_dims_converged(pkg, orc):
    """Bond dimensions 1..4 inside one network (one DP = 4 bucket, per-port padding):
    converged messages and log-norm against the FP64 oracle."""
    _say("case small-bond mixed dims converged")
    g = pkg.named_grid((6, 5))
    dims = {e: 1 + (i % 4) for i, e in enumerate(g.edges())}
    t = pkg.build_topology(g, dims, 2)
    s = pkg.peps.random_sites(t, 75)
    m0 = pkg.peps.identity_messages(t)
    want, sw, _, at = orc.bp(t, s, m0, maxiter=500, tol=1e-10, criterion="fidelity", nthreads=8)
    with pkg.Engine(t, s, dtype="f32") as e:
        assert e.path() == "vertex"
        e.set_messages(m0)
        r = e.run(maxiter=500, tol=1e-10, criterion="fidelity")
        got = e.messages()
        ln = e.lognorm()
    assert r.converged[0]
    assert _rel(got, want) < 1e-4
    lw = orc.bethe_free_energy(t, s, want)
    assert abs(ln - lw) / abs(lw) < 1e-5


@pytest.mark.parametrize("D,d", [(4, 2), (3, 3), (2, 2)])
def test_small_matches_padded_kernel(pkg, D, d):
    """small_bond = 1 (k_vertex_s) and 0 (zero-padded to 8 on the MFMA kernel) agree after
    30 sweeps from identity messages."""
    _say(f"case small-bond vs padded D={D} d={d}")
    t = pkg.build_topology(pkg.named_grid((8, 8)), D, d)
    s = pkg.peps.random_sites(t, 77, scale=pkg.peps.peps_scale(d, D))
    res = {}
    for sb in (1, 0):
        with pkg.Engine(t, s, dtype="f32") as e:
            e.set_option("small_bond", sb)
            assert e.path() == "vertex"
            e.init_messages("identity")
            e.run(maxiter=30)
            res[sb] = (e.messages(), e.lognorm())
    assert _rel(res[1][0], res[0][0]) < 1e-5
    assert res[1][1] == pytest.approx(res[0][1], rel=1e-6)


def test_small_batch_every_bucket_masks_and_graph(pkg, orc):
    """One context with networks on every bucket: DP = 2, 3, 4 small-bond, D = 6 (padded
    MFMA), D = 10 (bond-dimension-12 kernel) and a degree-5 network (generic kernel).
    Per-network stopping, graph replay equal to eager sweeps, each network against the
    FP64 oracle."""
    _say("case small-bond batch")
    specs = [(pkg.named_grid((4, 4)), 2), (pkg.named_grid((5, 3)), 3), (pkg.named_grid((3, 4)), 4),
             (pkg.named_grid((3, 3)), 6), (pkg.named_grid((2, 3)), 10)]
    tops = [pkg.build_topology(g, D, 2) for g, D in specs]
    star = pkg.NamedGraph(vertices=list(range(6)), edges=[(0, k) for k in range(1, 6)])
    tops.append(pkg.build_topology(star, 2, 2))
    sites = [pkg.peps.random_sites(t, 80 + i) for i, t in enumerate(tops)]
    tb = pkg.concat_topologies(tops)
    sb = np.concatenate(sites)
    out = {}
    for graph in (True, False):
        with pkg.Engine(tb, sb, dtype="f32") as e:
            assert e.path() == "mixed"
            e.init_messages("identity")
            r = e.run(maxiter=400, tol=1e-10, criterion="fidelity", use_graph=graph)
            out[graph] = (r, e.messages(), e.lognorms())
    rg, mg, lg = out[True]
    re_, me, le = out[False]
    assert np.array_equal(rg.sweeps, re_.sweeps) and np.array_equal(rg.at_iteration, re_.at_iteration)
    assert np.array_equal(mg, me) and np.array_equal(lg, le)
    assert np.all(rg.converged)
    off = 0
    for i, (t, s) in enumerate(zip(tops, sites)):
        n = int(t.msg_off[-1])
        want, _, _, _ = orc.bp(t, s, pkg.peps.identity_messages(t), maxiter=400, tol=1e-10,
                               criterion="fidelity")
        assert _rel(mg[off:off + n], want) < 1e-4, i
        lw = orc.bethe_free_energy(t, s, want)
        assert abs(lg[i] - lw) / abs(lw) < 1e-5, i
        off += n
