"""Full-size parity against the FP64 oracle at the headline shapes.

32x32 open-lattice PEPS norm network, d = 2, D = 8, sites iid N(0,1)/sqrt(d D^4) (seed 1),
identity messages, Jacobi; the oracle converges the same network in FP64 to a max-abs
change of 1e-15 (29 sweeps).  24x24, D = 16: BF16 MFMA sweeps, then FP64 polishing sweeps.

Stated tolerances (relative to the largest entry of the oracle's quantity; measured values
on MI355X in brackets):
  FP32, stopped at the configured fidelity 1e-7:     messages 1e-4 [3.4e-5], log-norm 1e-9
                                                      [3.1e-11]
  FP32, converged to fidelity 1e-12:                  messages 1e-6 [1.4e-7], log-norm 1e-9
                                                      [4.1e-11]; log-norm gradient 1e-6
                                                      [1.6e-7]; TFIM energy 1e-6 [2.6e-8]
                                                      and gradient 1e-6 [2.1e-7] (adjoint
                                                      solved to relative change 1e-6)
  then FP64 polishing sweeps:                         messages and log-norm 1e-10 (FP64 bar)
  BF16 at D = 16, fidelity 1e-5:                      messages 1e-2 [2.8e-3], log-norm 1e-8
                                                      [2.1e-10]
  + 3 FP64 polishing sweeps:                          messages 1e-4 [1.7e-5], log-norm 1e-12
                                                      [1.5e-14]
FP32 sweeps cannot drive the adjoint's relative change below about 8e-8 (rounding of each
FP32 VJP sweep), so FP32 gradients use an adjoint tolerance of 1e-6; their error against the
oracle is set by FP32 arithmetic (1.9e-7 even from the oracle's exact messages).
"""
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

NT = min(os.cpu_count() or 1, 16)


def _say(msg):
    print(msg, flush=True)


def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


@pytest.fixture(scope="module")
def c2(pkg, orc):
    t = pkg.build_topology(pkg.named_grid((32, 32)), 8, 2)
    s = pkg.peps.random_sites(t, 1, scale=pkg.peps.peps_scale(2, 8))
    m0 = pkg.peps.identity_messages(t)
    m64, sw, dl, at = orc.bp(t, s, m0, maxiter=400, tol=1e-15, criterion="maxabs", nthreads=NT)
    assert at > 0
    return dict(t=t, s=s, m0=m0, m64=m64, ln64=orc.bethe_free_energy(t, s, m64))


def test_c2_fp32_forward_fullsize(pkg, orc, c2):
    _say("case c2 32x32 D=8 fp32 forward vs fp64 oracle")
    t, s, m0, m64, ln64 = c2["t"], c2["s"], c2["m0"], c2["m64"], c2["ln64"]
    with pkg.Engine(t, s, dtype="f32") as e:
        assert e.path() == "vertex"
        e.set_messages(m0)
        r = e.run(maxiter=200, tol=1e-7)
        em, el = _rel(e.messages(), m64), abs(e.lognorm() - ln64) / abs(ln64)
        _say(f"fidelity 1e-7: sweeps={r.sweeps[0]} msg={em:.2e} lognorm={el:.2e}")
        assert r.converged[0] and em < 1e-4 and el < 1e-9
        e.set_messages(m0)
        r = e.run(maxiter=200, tol=1e-12)
        em, el = _rel(e.messages(), m64), abs(e.lognorm() - ln64) / abs(ln64)
        _say(f"fidelity 1e-12: sweeps={r.sweeps[0]} msg={em:.2e} lognorm={el:.2e}")
        assert r.converged[0] and em < 1e-6 and el < 1e-9
        e.polish_fp64(10)
        em, el = _rel(e.messages(), m64), abs(e.lognorm() - ln64) / abs(ln64)
        _say(f"+10 FP64 polishing sweeps: msg={em:.2e} lognorm={el:.2e}")
        assert em < 1e-10 and el < 1e-10


def test_c2_fp32_gradients_fullsize(pkg, orc, c2):
    _say("case c2 32x32 D=8 fp32 forward + adjoint gradients vs fp64 oracle")
    t, s, m0, m64 = c2["t"], c2["s"], c2["m0"], c2["m64"]
    gl64 = orc.grad_lognorm_explicit(t, s, m64)
    E64, dm, da = orc.tfim_energy(t, s, m64, 1.0, 3.0, partials=True)
    g64, it64, res64 = orc.fixed_point_gradient(t, s, m64, dm, da, maxiter=500, tol=1e-13,
                                                nthreads=NT)
    assert res64 < 1e-13
    with pkg.Engine(t, s, dtype="f32") as e:
        e.set_messages(m0)
        r = e.run(maxiter=200, tol=1e-12)
        assert r.converged[0]
        gl = e.grad_lognorm()
        E, g, it, res = e.tfim_energy(1.0, 3.0, grad=True, adj_tol=1e-6, adj_maxiter=200)
    egl, eE, eg = _rel(gl, gl64), abs(E - E64) / abs(E64), _rel(g, g64)
    _say(f"log-norm grad={egl:.2e} energy={eE:.2e} tfim grad={eg:.2e} adjoint sweeps={it} "
         f"res={res:.1e} (oracle {it64} sweeps)")
    assert egl < 1e-6 and eE < 1e-6 and eg < 1e-6
    assert res < 1e-6


def test_c5_bf16_then_polish_fullsize(pkg, orc):
    _say("case c5 24x24 D=16 bf16 + fp64 polish vs fp64 oracle")
    t = pkg.build_topology(pkg.named_grid((24, 24)), 16, 2)
    s = pkg.peps.random_sites(t, 1, scale=pkg.peps.peps_scale(2, 16))
    m0 = pkg.peps.identity_messages(t)
    m64, sw, dl, at = orc.bp(t, s, m0, maxiter=60, tol=1e-13, criterion="maxabs", nthreads=NT)
    assert at > 0
    ln64 = orc.bethe_free_energy(t, s, m64)
    with pkg.Engine(t, s, dtype="bf16") as e:
        assert e.path() == "vertex"
        e.set_messages(m0)
        r = e.run(maxiter=200, tol=1e-5)
        assert r.converged[0]
        stages = [("bf16", _rel(e.messages(), m64), abs(e.lognorm() - ln64) / abs(ln64))]
        e.polish_fp64(3)
        stages.append(("+3 fp64", _rel(e.messages(), m64), abs(e.lognorm() - ln64) / abs(ln64)))
        e.polish_fp64(3)
        stages.append(("+6 fp64", _rel(e.messages(), m64), abs(e.lognorm() - ln64) / abs(ln64)))
    for name, em, el in stages:
        _say(f"{name}: msg={em:.2e} lognorm={el:.2e}")
    assert stages[0][1] < 1e-2 and stages[0][2] < 1e-8
    assert stages[1][1] < 1e-4 and stages[1][2] < 1e-12
    assert stages[2][1] < 1e-6 and stages[2][2] < 1e-12
    assert stages[2][1] < stages[1][1] < stages[0][1]


def test_c4_fp32_gradient_fullsize(pkg, orc):
    """Differentiation workload at full size: 16x16 PEPS, d = 2, D = 6, FP32 forward (the
    D <= 8 vertex kernels) and FP32 adjoint sweeps, against the FP64 oracle's fixed-point
    gradient (stated tolerance 1e-5 relative to the largest entry; measured 2.3e-7) and
    against central differences (step 1e-3, 20 seeded entries) of the FP64 engine's
    converged energy (2e-3: the O(step^2) truncation of the difference quotient; measured
    7.6e-6).  The energy is a sum of terms of both signs that nearly cancel here (about
    -0.8 from terms of order 3 per site), so its stated tolerance is 5e-6 relative
    (measured 1.3e-6)."""
    _say("case c4 16x16 D=6 fp32 gradient vs oracle and central differences")
    t = pkg.build_topology(pkg.named_grid((16, 16)), 6, 2)
    s = pkg.peps.random_sites(t, 33, scale=pkg.peps.peps_scale(2, 6))
    m0 = pkg.peps.identity_messages(t)
    m64, _, _, at = orc.bp(t, s, m0, maxiter=400, tol=1e-15, criterion="maxabs", nthreads=NT)
    assert at > 0
    E64, dm, da = orc.tfim_energy(t, s, m64, 1.0, 3.0, partials=True)
    g64, _, res64 = orc.fixed_point_gradient(t, s, m64, dm, da, maxiter=500, tol=1e-13,
                                             nthreads=NT)
    with pkg.Engine(t, s, dtype="f32") as e:
        assert e.path() == "vertex"
        e.set_messages(m0)
        assert e.run(maxiter=400, tol=1e-12).converged[0]
        E, g, it, res = e.tfim_energy(1.0, 3.0, grad=True, adj_tol=1e-6, adj_maxiter=400)
    eg = _rel(g, g64)
    idx = np.random.default_rng(7).choice(s.size, 20, replace=False)
    step = 1e-3
    fd = []
    with pkg.Engine(t, s, dtype="f64") as e:
        for i in idx:
            vals = []
            for sg in (1.0, -1.0):
                s2 = s.copy()
                s2[i] += sg * step
                e.set_sites(s2)
                e.set_messages(m64)
                assert e.run(maxiter=2000, tol=1e-13, criterion="maxabs").converged[0]
                vals.append(e.tfim_energy(1.0, 3.0))
            fd.append((vals[0] - vals[1]) / (2 * step))
    fd = np.array(fd)
    efd = float(np.max(np.abs(g[idx] - fd)) / np.max(np.abs(fd)))
    _say(f"E rel={abs(E - E64) / abs(E64):.2e} grad vs oracle={eg:.2e} vs FD={efd:.2e} "
         f"adjoint sweeps={it}")
    assert abs(E - E64) / abs(E64) < 5e-6
    assert eg < 1e-5
    assert efd < 2e-3
