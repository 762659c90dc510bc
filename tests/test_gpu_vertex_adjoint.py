"""Fixed-point adjoint on the bond-dimension-8 vertex kernels (FP32 MFMA) against the
generic-kernel adjoint on the same converged messages and against the FP64 oracle.

Energy convention: H = -J sum_<uv> Z_u Z_v - h sum_v X_v, J = 1, h = 3.
Stated tolerances: vertex FP32 vs the FP64 oracle: 1e-6 relative on the energy, 1e-3 of the
largest gradient entry; vertex vs generic FP32 (same messages, different summation order):
2e-6 relative on the energy (each path within 1e-6 of the FP64 value, on opposite sides at
worst: measured 5.4e-7 and 5.6e-7 on the 6x6 lattice), 2e-4 of the largest gradient entry.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


def _lattice(pkg, kind, L, seed, D=8):
    g = pkg.heavy_hex_lattice(2, L) if kind == "heavyhex" else pkg.named_grid((L, L))
    t = pkg.build_topology(g, D, 2)
    s = pkg.peps.random_sites(t, seed, scale=pkg.peps.peps_scale(2, D))
    return t, s


def _pair(pkg, t, s):
    """A vertex-path engine converged in FP32 and a generic-path engine holding the same
    messages."""
    ev = pkg.Engine(t, s, dtype="f32")
    assert ev.path() == "vertex"
    ev.init_messages("identity")
    r = ev.run(maxiter=500, tol=1e-12, criterion="fidelity")
    assert r.converged[0], r
    eg = pkg.Engine(t, s, dtype="f32", path="generic")
    eg.set_messages(ev.messages())
    return ev, eg


@pytest.mark.parametrize("kind,L", [("square", 6), ("heavyhex", 5)])
def test_vertex_tfim_matches_generic(pkg, kind, L):
    _say(f"case vertex adjoint tfim vs generic {kind} L={L} D=8")
    t, s = _lattice(pkg, kind, L, 41)
    ev, eg = _pair(pkg, t, s)
    Ev, gv, itv, resv = ev.tfim_energy(1.0, 3.0, grad=True, adj_tol=1e-6, adj_maxiter=500)
    Eg, gg, itg, resg = eg.tfim_energy(1.0, 3.0, grad=True, adj_tol=1e-6, adj_maxiter=500)
    ev.close()
    eg.close()
    _say(f"E vertex={Ev:.9f} generic={Eg:.9f} sweeps {itv}/{itg} res {resv:.1e}/{resg:.1e} "
         f"grad rel diff={_rel(gv, gg):.2e}")
    assert Ev == pytest.approx(Eg, rel=2e-6)
    assert resv < 1e-6
    assert _rel(gv, gg) < 2e-4


def test_vertex_tfim_energy_only(pkg):
    _say("case vertex tfim energy without gradient")
    t, s = _lattice(pkg, "square", 4, 42)
    ev, eg = _pair(pkg, t, s)
    Ev = ev.tfim_energy(1.0, 3.0)
    Eg = eg.tfim_energy(1.0, 3.0)
    ev.close()
    eg.close()
    assert Ev == pytest.approx(Eg, rel=2e-6)


def test_vertex_grad_lognorm_matches_generic(pkg):
    _say("case vertex log-norm gradient vs generic 6x6 D=8")
    t, s = _lattice(pkg, "square", 6, 43)
    ev, eg = _pair(pkg, t, s)
    gv = ev.grad_lognorm()
    gg = eg.grad_lognorm()
    ev.close()
    eg.close()
    _say(f"log-norm grad rel diff={_rel(gv, gg):.2e}")
    assert _rel(gv, gg) < 2e-4


def test_vertex_tfim_vs_fp64_oracle(pkg, orc):
    _say("case vertex adjoint tfim vs FP64 oracle 5x5 D=8")
    t, s = _lattice(pkg, "square", 5, 44)
    ev = pkg.Engine(t, s, dtype="f32")
    assert ev.path() == "vertex"
    ev.init_messages("identity")
    r = ev.run(maxiter=500, tol=1e-12, criterion="fidelity")
    assert r.converged[0]
    E, g, it, res = ev.tfim_energy(1.0, 3.0, grad=True, adj_tol=1e-6, adj_maxiter=500)
    gl = ev.grad_lognorm()
    ev.close()
    m64, _, _, at = orc.bp(t, s, pkg.peps.identity_messages(t), maxiter=3000, tol=1e-15,
                           criterion="maxabs", nthreads=8)
    assert at > 0
    Eo, dm, da = orc.tfim_energy(t, s, m64, 1.0, 3.0, partials=True)
    want, _, _ = orc.fixed_point_gradient(t, s, m64, dm, da, maxiter=5000, tol=1e-12)
    wl = orc.grad_lognorm_explicit(t, s, m64)
    _say(f"E={E:.9f} oracle={Eo:.9f} sweeps={it} grad err={_rel(g, want):.2e} "
         f"lognorm grad err={_rel(gl, wl):.2e}")
    assert E == pytest.approx(Eo, rel=1e-6)
    assert _rel(g, want) < 1e-3
    assert _rel(gl, wl) < 1e-3


@pytest.mark.parametrize("dtype,path", [("f32", "auto"), ("f32", "generic"), ("f64", "auto")])
def test_device_gradient_matches_host(pkg, dtype, path):
    """The device-resident gradient entry points give the host entry points' values."""
    _say(f"case device gradient {dtype} {path}")
    t, s = _lattice(pkg, "square", 4, 45)
    e = pkg.Engine(t, s, dtype=dtype, path=path)
    e.init_messages("identity")
    e.run(maxiter=500, tol=1e-12, criterion="fidelity")
    E, gh, _, _ = e.tfim_energy(1.0, 3.0, grad=True, adj_tol=1e-7, adj_maxiter=500)
    E2, gd, _, _ = e.tfim_energy(1.0, 3.0, grad=True, adj_tol=1e-7, adj_maxiter=500, device=True)
    lh = e.grad_lognorm()
    ld = e.grad_lognorm(device=True)
    e.close()
    assert gd.dtype == (np.float64 if dtype == "f64" else np.float32)
    assert E2 == E
    tol = 1e-12 if dtype == "f64" else 1e-6
    assert _rel(gd.numpy().astype(np.float64), gh) < tol
    assert _rel(ld.numpy().astype(np.float64), lh) < tol
    gd.free()
    ld.free()


@pytest.mark.parametrize("path", ["auto", "generic"])
def test_fp32_adjoint_stops_at_its_rounding_floor(pkg, path):
    """An FP32 adjoint asked for a tolerance below its rounding floor (relative change about
    1e-7) stops once the change no longer falls (within 128 sweeps here, not at adj_maxiter),
    reports the residual it reached, and gives the gradient of a 1e-6 solve to 1e-5."""
    _say(f"case fp32 adjoint floor {path}")
    t = pkg.build_topology(pkg.named_grid((8, 8)), 6, 2)
    s = pkg.peps.random_sites(t, 46, scale=pkg.peps.peps_scale(2, 6))
    with pkg.Engine(t, s, dtype="f32", path=path) as e:
        e.init_messages("identity")
        assert e.run(maxiter=500, tol=1e-12).converged[0]
        _, g, it, res = e.tfim_energy(1.0, 3.0, grad=True, adj_tol=1e-10, adj_maxiter=1000)
        _, g6, it6, res6 = e.tfim_energy(1.0, 3.0, grad=True, adj_tol=1e-6, adj_maxiter=1000)
    _say(f"floor stop after {it} sweeps at {res:.2e}; 1e-6 solve {it6} sweeps at {res6:.2e}")
    assert it <= 128 and 1e-10 <= res < 1e-6
    assert res6 < 1e-6 and it6 < it
    assert _rel(g, g6) < 1e-5
