"""Cached sweep graphs: a context replays its captured 16-sweep graph across bp_run calls
with different stopping parameters, and drops it when the path or a kernel option changes.
Every run is compared bit for bit with the same run launched sweep by sweep (no graph) on a
fresh context: same messages, same per-network sweep counts."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _fresh_run(pkg, t, s, dtype, path, msgs, kw, options=()):
    e = pkg.Engine(t, s, dtype=dtype, path=path)
    for name, value in options:
        e.set_option(name, value)
    e.set_messages(msgs)
    r = e.run(use_graph=False, **kw)
    m = e.messages()
    e.close()
    return r, m


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_graph_reused_across_runs(pkg, dtype):
    _say(f"case graph reuse {dtype}")
    g = pkg.named_grid((6, 6))
    t = pkg.build_topology(g, 4, 2)
    s = pkg.peps.random_sites(t, 51, scale=pkg.peps.peps_scale(2, 4))
    e = pkg.Engine(t, s, dtype=dtype)
    e.init_messages("identity")
    runs = [dict(maxiter=40, tol=-1.0, criterion="fidelity"),
            dict(maxiter=200, tol=1e-6, criterion="maxabs"),
            dict(maxiter=17, tol=1e-30, criterion="fidelity"),
            dict(maxiter=300, tol=1e-9, criterion="fidelity")]
    options = []
    for k, kw in enumerate(runs):
        if k == 2:
            e.set_path("generic")
        if k == 3:
            e.set_path("auto")
            if dtype == "f32":
                e.set_option("vertex_variant", 1)
                options = [("vertex_variant", 1)]
        path = "generic" if k == 2 else "auto"
        start = e.messages()
        kw = {a: b for a, b in kw.items() if not (a == "tol" and b < 0)}
        r = e.run(**kw)
        got = e.messages()
        rw, want = _fresh_run(pkg, t, s, dtype, path, start, kw, options)
        _say(f"run {k} {kw} sweeps {r.sweeps.tolist()} fresh {rw.sweeps.tolist()}")
        assert np.array_equal(r.sweeps, rw.sweeps)
        assert np.array_equal(got, want)
    e.close()


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_prepared_run_tail_graphs(pkg, dtype):
    """bp_prepare_run captures the 16-sweep graphs and the tail chunk (maxiter % 16) from
    both message buffers without running a sweep (messages unchanged); runs afterwards
    (from either buffer, with and without a tolerance) equal eager runs bit for bit."""
    _say(f"case prepared run {dtype}")
    t = pkg.build_topology(pkg.named_grid((6, 5)), 6, 2)
    s = pkg.peps.random_sites(t, 52, scale=pkg.peps.peps_scale(2, 6))
    e = pkg.Engine(t, s, dtype=dtype)
    e.init_messages("identity")
    m0 = e.messages()
    e.prepare_run(20)
    assert np.array_equal(e.messages(), m0)
    runs = [dict(maxiter=20), dict(maxiter=5), dict(maxiter=20),
            dict(maxiter=36, tol=1e-30, criterion="maxabs"), dict(maxiter=20, tol=1e-6)]
    for k, kw in enumerate(runs):
        start = e.messages()
        r = e.run(use_graph=True, **kw)
        got = e.messages()
        rw, want = _fresh_run(pkg, t, s, dtype, "auto", start, kw)
        _say(f"run {k} {kw} sweeps {r.sweeps.tolist()} fresh {rw.sweeps.tolist()}")
        assert np.array_equal(r.sweeps, rw.sweeps) and np.array_equal(r.at_iteration, rw.at_iteration)
        assert np.array_equal(got, want)
    e.close()
