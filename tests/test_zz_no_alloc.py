"""bp_iterate_diff_noalloc (the entry point that allocates nothing) against bp_iterate_diff,
through the C ABI: two calls in a row on one GPU, the second on a smaller network (the
static scratch array still holds the first call's contents), and the size limit."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _both(pkg, e, other):
    lib = pkg._lib.load()
    o = np.ascontiguousarray(other, dtype=np.float64)
    p = o.ctypes.data_as(C.POINTER(C.c_double))
    a, b = C.c_double(0), C.c_double(0)
    pkg._lib.check(lib.bp_iterate_diff(e.h, p, C.byref(a)), "bp_iterate_diff")
    pkg._lib.check(lib.bp_iterate_diff_noalloc(e.h, p, C.byref(b)), "bp_iterate_diff_noalloc")
    return a.value, b.value


def test_noalloc_matches_iterate_diff_twice(pkg):
    # first call: 32x32 D=8 FP32 after two sweeps, against the messages after one
    _say("case noalloc 32x32 D=8 f32")
    t = pkg.build_topology(pkg.named_grid((32, 32)), 8, 2)
    s = pkg.peps.random_sites(t, 1, scale=pkg.peps.peps_scale(2, 8))
    with pkg.Engine(t, s, dtype="f32") as e:
        e.init_messages("identity")
        e.sweep()
        m1 = e.messages()
        e.sweep()
        a, b = _both(pkg, e, m1)
        _say(f"iterate_diff {a:.6e} noalloc {b:.6e}")
        assert a > 0.0 and b == a
        a, b = _both(pkg, e, e.messages())
        assert b == a and abs(a) < 1e-12
    # second call: a smaller FP64 network on the same GPU
    _say("case noalloc 6x6 D=3 f64")
    t = pkg.build_topology(pkg.named_grid((6, 6)), 3, 2)
    s = pkg.peps.random_sites(t, 2)
    with pkg.Engine(t, s, dtype="f64") as e:
        m = pkg.peps.random_messages(t, 3)
        e.set_messages(m)
        other = pkg.peps.random_messages(t, 4)
        a, b = _both(pkg, e, other)
        _say(f"iterate_diff {a:.6e} noalloc {b:.6e}")
        assert a > 0.0 and b == a
        other[7] = np.nan  # a NaN distance is kept, as maximum() keeps it
        a, b = _both(pkg, e, other)
        assert np.isnan(a) and np.isnan(b)


def test_noalloc_rejects_oversized_networks(pkg):
    _say("case noalloc size limit")
    t = pkg.build_topology(pkg.named_path_graph(200), 64, 2)
    assert 2 * t.nports + int(t.msg_off[-1]) + 1 > 1 << 20
    with pkg.Engine(t, np.zeros(int(t.site_off[-1])), dtype="f64", path="generic") as e:
        e.init_messages("identity")
        lib = pkg._lib.load()
        other = np.zeros(int(t.msg_off[-1]))
        out = C.c_double(0)
        rc = lib.bp_iterate_diff_noalloc(e.h, other.ctypes.data_as(C.POINTER(C.c_double)),
                                         C.byref(out))
        assert rc == pkg._lib.BP_ERR_ARG
        assert b"scratch" in lib.bp_last_error()
