"""Batched uneven networks on the GPU: per-network convergence masks, mixed dispatch
(vertex kernels for the networks with every bond dimension <= 12 in FP32, generic kernel
for the rest, in the same sweep), against the FP64 oracle per network."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _check_batch(pkg, orc, tops, sites, dtype, mtol, ltol, maxiter=400, tol=1e-10):
    # FP32 runs stop on the reference's fidelity criterion (1e-10 ~ 1e-5 in angle): a
    # max-abs change below ~1e-6 is under FP32 resolution for some of these networks
    b = pkg.concat_topologies(tops)
    with pkg.Engine(b, np.concatenate(sites), dtype=dtype) as e:
        path = e.path()
        e.init_messages("identity")
        r = e.run(maxiter=maxiter, tol=tol, criterion="fidelity")
        mb = e.messages()
        ln = e.lognorms()
    off = 0
    for i, (t, s) in enumerate(zip(tops, sites)):
        n = int(t.msg_off[-1])
        want, sw, dl, at = orc.bp(t, s, pkg.peps.identity_messages(t), maxiter=maxiter * 4,
                                  tol=1e-13, criterion="maxabs", nthreads=8)
        assert r.converged[i], (i, r.delta[i])
        assert at > 0
        assert _rel(mb[off:off + n], want) < mtol, i
        lw = orc.bethe_free_energy(t, s, want)
        assert abs(ln[i] - lw) / abs(lw) < ltol, i
        off += n
    return path, r


def test_mixed_dispatch_batch(pkg, orc):
    _say("case mixed dispatch batch")
    # bond dimensions <= 8 run on the D <= 8 vertex kernels, D = 10 on the FP32 large-bond
    # kernel (9..12), D = 13 on the generic kernel
    specs = [(5, 8, False), (4, 3, False), (3, 10, False), (3, 13, False), (3, 5, True)]
    tops, sites = [], []
    for k, (L, D, hh) in enumerate(specs):
        g = pkg.heavy_hex_lattice(2, L) if hh else pkg.named_grid((L, L))
        t = pkg.build_topology(g, D, 2)
        tops.append(t)
        sites.append(pkg.peps.random_sites(t, 70 + k, scale=pkg.peps.peps_scale(2, D)))
    path, r = _check_batch(pkg, orc, tops, sites, "f32", 1e-4, 1e-5)
    assert path == "mixed"
    _say(f"mixed batch sweeps per network: {list(r.sweeps)}")


def test_uneven_workload_shape(pkg, orc):
    """The batched uneven workload's distribution (D in {4,6,8,10,12}, square or heavy-hex,
    scale 1/sqrt(d D^4)) at reduced L so the oracle checks every network in seconds."""
    _say("case uneven workload (reduced L)")
    tops, sites = pkg.peps.uneven_batch(seed=3, nnet=10, L_range=(3, 6))
    path, r = _check_batch(pkg, orc, tops, sites, "f32", 1e-4, 1e-5)
    _say(f"uneven batch path={path} sweeps={list(r.sweeps)}")


def test_uneven_workload_full_size(pkg, orc):
    """64 networks, L in 8..48, one launch per sweep: every network converges, per-network
    masks stop the converged ones, and the two smallest match the oracle."""
    _say("case uneven workload (full size)")
    tops, sites = pkg.peps.uneven_batch(seed=1, nnet=64)
    b = pkg.concat_topologies(tops)
    with pkg.Engine(b, np.concatenate(sites), dtype="f32") as e:
        e.init_messages("identity")
        r = e.run(maxiter=200, tol=1e-6, criterion="fidelity")
        mb = e.messages()
        ln = e.lognorms()
        _say(f"full batch: nv={b.nv} path={e.path()} ms={r.ms:.1f} converged="
             f"{int(np.sum(r.converged))}/64 sweeps min/max={r.sweeps.min()}/{r.sweeps.max()}")
    assert np.all(r.converged)
    assert np.all(np.isfinite(ln))
    assert len(set(r.sweeps.tolist())) > 1  # networks stop at different sweeps
    sizes = [int(t.site_off[-1]) for t in tops]
    offs = np.concatenate([[0], np.cumsum([int(t.msg_off[-1]) for t in tops])])
    for i in np.argsort(sizes)[:2]:
        t, s = tops[i], sites[i]
        want, _, _, _ = orc.bp(t, s, pkg.peps.identity_messages(t), maxiter=200, tol=1e-6,
                               criterion="fidelity", nthreads=8)
        assert _rel(mb[offs[i]:offs[i + 1]], want) < 1e-3


def test_folded_bookkeeping_matches_separate_kernel(pkg):
    """Networks of one batch that stop at different sweeps, in the middle of the 16-sweep
    graphs: the bookkeeping folded into the D <= 8 sweep kernels (each sweep applies the
    previous sweep's stopping rule) gives bit for bit the messages, sweep counts, deltas and
    convergence sweeps of a separate bookkeeping kernel after every sweep."""
    _say("case folded bookkeeping vs separate kernel")
    tops, sites = [], []
    for k, (L, D) in enumerate([(6, 8), (4, 8), (5, 6), (3, 4), (7, 8)]):
        t = pkg.build_topology(pkg.named_grid((L, L)), D, 2)
        tops.append(t)
        sites.append(pkg.peps.random_sites(t, 90 + k, scale=pkg.peps.peps_scale(2, D)))
    b = pkg.concat_topologies(tops)
    runs = [dict(maxiter=300, tol=1e-9, criterion="fidelity"),
            dict(maxiter=300, tol=1e-5, criterion="maxabs"),
            dict(maxiter=21, tol=1e-12, criterion="fidelity")]
    for kw in runs:
        out = {}
        for fold in (1, 0):
            for graph in (True, False):
                with pkg.Engine(b, np.concatenate(sites), dtype="f32") as e:
                    assert e.path() == "vertex"
                    e.set_option("fold_bookkeeping", fold)
                    e.init_messages("identity")
                    r = e.run(use_graph=graph, **kw)
                    out[(fold, graph)] = (r, e.messages())
        r1, m1 = out[(1, True)]
        _say(f"{kw}: sweeps {r1.sweeps.tolist()} at {r1.at_iteration.tolist()}")
        assert len(set(r1.sweeps.tolist())) > 1  # the networks stop at different sweeps
        for key, (r, m) in out.items():
            assert np.array_equal(r.sweeps, r1.sweeps), key
            assert np.array_equal(r.at_iteration, r1.at_iteration), key
            assert np.array_equal(r.delta, r1.delta), key
            assert np.array_equal(m, m1), key
