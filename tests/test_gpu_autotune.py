"""Autotuner: timed choices on the engine's own data (vertex kernels or generic kernel per
largest bond dimension <= 12 in FP32; workgroup size and storage per generic bucket, here
the D = 13 networks), persisted to
a JSON cache and applied again from it.  Tuning changes how a sweep is scheduled, never what
it computes: results after tuning match the untuned engine within the FP32 bound."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _batch(pkg):
    tops, sites = pkg.peps.uneven_batch(seed=5, nnet=8, L_range=(3, 6),
                                        D_choices=(3, 6, 8, 12, 13))
    return pkg.concat_topologies(tops), np.concatenate(sites), tops


def test_autotune_cache_roundtrip(pkg, tmp_path):
    _say("case autotune cache round trip")
    b, s, tops = _batch(pkg)
    cache = str(tmp_path / "tune.json")
    with pkg.Engine(b, s, dtype="f32") as e:
        e.init_messages("identity")
        m0 = e.messages()
        n = e.autotune(cache=cache, mode="missing")
        assert n > 0
        assert np.array_equal(e.messages(), m0)  # messages restored
        r = e.run(maxiter=300, tol=1e-10, criterion="fidelity")
        tuned = e.messages()
        path_tuned = e.path()
    table = json.load(open(cache))
    assert any(k.startswith("gen|f32|") for k in table)
    for k, v in table.items():
        if k.startswith("gen|"):
            assert v["block"] in (128, 256, 512) and v["lds"] in (0, 1, 2)
            assert v["us"] <= v["us_alt"] * 1.0001
        elif k.startswith("vxk|"):
            # FP32 weighs six sweep kernels (four with MFMA open contractions)
            assert k.startswith("vxk|f32|") and v["variant"] in (0, 1, 2, 3, 4, 5)
        else:
            assert k.startswith("vx|f32|") and v["vertex"] in (0, 1)
    _say(f"tuned {n} decisions, path={path_tuned}: {table}")
    with pkg.Engine(b, s, dtype="f32") as e:
        e.init_messages("identity")
        r0 = e.run(maxiter=300, tol=1e-10, criterion="fidelity")
        untuned = e.messages()
    assert np.all(r.converged) and np.all(r0.converged)
    assert _rel(tuned, untuned) < 1e-4
    # a fresh engine applies the cached choices without timing anything
    with pkg.Engine(b, s, dtype="f32") as e:
        e.init_messages("identity")
        assert e.autotune(cache=cache, mode="cached") == 0
        assert e.autotune(cache=cache, mode="missing") == 0
        assert e.path() == path_tuned
        e.run(maxiter=300, tol=1e-10, criterion="fidelity")
        assert _rel(e.messages(), tuned) < 1e-6


def test_autotune_fp64_generic_buckets(pkg, orc):
    """FP64 (generic kernel only): tuned buckets keep 1e-10 parity with the oracle."""
    _say("case autotune fp64")
    t = pkg.build_topology(pkg.heavy_hex_lattice(2, 4), 5, 2)
    s = pkg.peps.random_sites(t, 9, scale=pkg.peps.peps_scale(2, 5))
    m0 = pkg.peps.identity_messages(t)
    with pkg.Engine(t, s, dtype="f64") as e:
        e.set_messages(m0)
        assert e.autotune(mode="all") > 0
        e.run(maxiter=400, tol=1e-13, criterion="maxabs")
        got = e.messages()
    want, _, _, at = orc.bp(t, s, m0, maxiter=400, tol=1e-13, criterion="maxabs")
    assert at > 0
    assert np.max(np.abs(got - want)) < 1e-10
