"""The FP32 error bound used by the GPU parity tests is tight enough to catch plausible
mistakes.  The mistakes are planted in the oracle, on the CPU."""
import numpy as np

from helpers import fp32_sweep_check


def _setup(pkg):
    t = pkg.build_topology(pkg.named_grid((4, 3)), 8, 2)
    s = pkg.peps.random_sites(t, 17, scale=pkg.peps.peps_scale(2, 8))
    m = pkg.peps.random_messages(t, 17)
    return t, s, m


def test_bound_accepts_fp32_rounding(pkg, orc):
    t, s, m = _setup(pkg)
    want, _, _ = orc.sweep_jacobi(t, s.astype(np.float32).astype(np.float64),
                                  m.astype(np.float32).astype(np.float64))
    got = want.astype(np.float32).astype(np.float64)
    assert fp32_sweep_check(pkg, orc, t, s, m, got) < 0.05


def test_bound_rejects_transposed_messages(pkg, orc):
    t, s, m = _setup(pkg)
    mt = m.copy()
    for q in range(t.nports):
        sl = slice(int(t.msg_off[q]), int(t.msg_off[q + 1]))
        mt[sl] = m[sl].reshape(8, 8).T.ravel()
    wrong, _, _ = orc.sweep_jacobi(t, s, mt)
    assert fp32_sweep_check(pkg, orc, t, s, m, wrong) > 100


def test_bound_rejects_dropped_normalisation_term(pkg, orc):
    t, s, m = _setup(pkg)
    want, _, _ = orc.sweep_jacobi(t, s, m)
    # the best-conditioned message (smallest sum|g| / |sum g|, about 33 here) carries a 1%
    # scale error
    cond = []
    for q in range(t.nports):
        g = orc.message(t, s, m, q)
        G = orc.message(t, np.abs(s), np.abs(m), q)
        cond.append(G.sum() / abs(g.sum()))
    q = int(np.argmin(cond))
    wrong = want.copy()
    sl = slice(int(t.msg_off[q]), int(t.msg_off[q + 1]))
    wrong[sl] *= 1.0 + 1e-2
    assert fp32_sweep_check(pkg, orc, t, s, m, wrong) > 1.0
