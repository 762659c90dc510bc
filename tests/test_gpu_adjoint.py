"""GPU fixed-point adjoint: log-norm and TFIM-energy gradients against the FP64 oracle
(implicit differentiation of the same fixed point) and against central finite
differences of the GPU's own converged BP.  No reference counterpart exists (the
reference has no AD); the oracle's adjoint is itself pinned by finite differences in
test_oracle.py.

Energy convention: H = -J sum_<uv> Z_u Z_v - h sum_v X_v, J = 1, h = 3.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _setup(pkg, L, D, seed):
    t = pkg.build_topology(pkg.named_grid((L, L)), D, 2)
    s = pkg.peps.random_sites(t, seed, scale=pkg.peps.peps_scale(2, D))
    return t, s


def _converged_engine(pkg, t, s, dtype, tol, path="auto"):
    e = pkg.Engine(t, s, dtype=dtype, path=path)
    e.init_messages("identity")
    r = e.run(maxiter=2000, tol=tol, criterion="maxabs" if dtype == "f64" else "fidelity")
    assert r.converged[0], r
    return e


def test_grad_lognorm_fp64_matches_oracle(pkg, orc):
    _say("case adjoint lognorm fp64 6x6 D=3")
    t, s = _setup(pkg, 6, 3, 31)
    e = _converged_engine(pkg, t, s, "f64", 1e-14)
    m = e.messages()
    g = e.grad_lognorm()
    e.close()
    want = orc.grad_lognorm_explicit(t, s, m)
    np.testing.assert_allclose(g, want, rtol=1e-10, atol=1e-12 * np.max(np.abs(want)))


def test_tfim_fp64_matches_oracle(pkg, orc):
    _say("case adjoint tfim fp64 5x5 D=3")
    t, s = _setup(pkg, 5, 3, 32)
    e = _converged_engine(pkg, t, s, "f64", 1e-14)
    m = e.messages()
    E, g, it, res = e.tfim_energy(1.0, 3.0, grad=True, adj_tol=1e-13, adj_maxiter=5000)
    e.close()
    Eo, dm, da = orc.tfim_energy(t, s, m, 1.0, 3.0, partials=True)
    assert E == pytest.approx(Eo, rel=1e-12)
    want, ito, reso = orc.fixed_point_gradient(t, s, m, dm, da, maxiter=5000, tol=1e-13)
    _say(f"adjoint sweeps gpu={it} oracle={ito}")
    assert res < 1e-13
    np.testing.assert_allclose(g, want, rtol=1e-8, atol=1e-9 * np.max(np.abs(want)))


def test_tfim_fd_16x16_d6_fp64(pkg):
    """Differentiation workload, FP64 path: adjoint gradient vs central differences
    (step 1e-3 on 20 seeded entries) of the GPU's own converged BP energy."""
    _say("case adjoint tfim fd 16x16 D=6")
    t, s = _setup(pkg, 16, 6, 33)
    e = _converged_engine(pkg, t, s, "f64", 1e-13, path="generic")
    E, g, it, res = e.tfim_energy(1.0, 3.0, grad=True, adj_tol=1e-12, adj_maxiter=5000)
    m0 = e.messages()
    idx = np.random.default_rng(7).choice(s.size, 20, replace=False)
    step = 1e-3
    fd = []
    for i in idx:
        vals = []
        for sg in (1.0, -1.0):
            s2 = s.copy()
            s2[i] += sg * step
            e.set_sites(s2)
            e.set_messages(m0)
            r = e.run(maxiter=2000, tol=1e-13, criterion="maxabs")
            assert r.converged[0]
            vals.append(e.tfim_energy(1.0, 3.0))
        fd.append((vals[0] - vals[1]) / (2 * step))
    e.close()
    fd = np.array(fd)
    scale = np.max(np.abs(fd))
    err = np.max(np.abs(g[idx] - fd)) / scale
    _say(f"fd check: E={E:.10f} adjoint sweeps={it} max rel err={err:.3e}")
    assert err < 2e-3  # O(step^2) truncation of the central difference


def test_tfim_fp32_vs_fp64_oracle(pkg, orc):
    """FP32 forward + FP32 adjoint against the FP64 oracle gradient (stated tolerance:
    1e-3 relative to the largest gradient entry, 1e-6 relative on the energy)."""
    _say("case adjoint tfim fp32 10x10 D=6")
    t, s = _setup(pkg, 10, 6, 34)
    e = _converged_engine(pkg, t, s, "f32", 1e-12)
    E, g, it, res = e.tfim_energy(1.0, 3.0, grad=True, adj_tol=1e-6, adj_maxiter=2000)
    e.close()
    m64, _, _, at = orc.bp(t, s, pkg.peps.identity_messages(t), maxiter=3000, tol=1e-15,
                           criterion="maxabs", nthreads=8)
    assert at > 0
    Eo, dm, da = orc.tfim_energy(t, s, m64, 1.0, 3.0, partials=True)
    want, _, _ = orc.fixed_point_gradient(t, s, m64, dm, da, maxiter=5000, tol=1e-12)
    _say(f"fp32 E={E:.8f} oracle={Eo:.8f} adjoint sweeps={it} res={res:.2e}")
    assert E == pytest.approx(Eo, rel=1e-6)
    assert np.max(np.abs(g - want)) / np.max(np.abs(want)) < 1e-3


def test_autograd_function(pkg):
    """torch.autograd over the device adjoint: d/dA of the TFIM energy and of the log-norm
    equal the engine's gradients, and a first-order Taylor check holds."""
    import importlib

    import torch

    _say("case autograd")
    ag = importlib.import_module(pkg.__name__ + ".autograd")
    t, s = _setup(pkg, 4, 2, 35)
    x = torch.tensor(s, dtype=torch.float64, requires_grad=True)
    E = ag.bp_tfim_energy(x, t, 1.0, 3.0)
    E.backward()
    g = x.grad.numpy().copy()
    e = _converged_engine(pkg, t, s, "f64", 1e-12)
    E2, g2, _, _ = e.tfim_energy(1.0, 3.0, grad=True, adj_tol=1e-10)
    e.close()
    assert float(E.detach()) == pytest.approx(E2, rel=1e-10)
    np.testing.assert_allclose(g, g2, rtol=1e-8, atol=1e-10 * np.max(np.abs(g2)))
    d = np.random.default_rng(1).standard_normal(s.size) * 1e-6
    E3 = ag.bp_tfim_energy(torch.tensor(s + d), t, 1.0, 3.0)
    assert float(E3) - float(E) == pytest.approx(float(g @ d), rel=1e-3)
    y = torch.tensor(s, dtype=torch.float64, requires_grad=True)
    ag.bp_lognorm(y, t).backward()
    assert np.all(np.isfinite(y.grad.numpy()))
