"""Shared test helpers (error models)."""
import numpy as np

# FP32 error model for one normalised sweep: with inputs rounded to FP32, the computed
# unnormalised entry g_i carries an error bounded by gamma * G_i, where G is the same
# contraction of |A| (x) |A| (x) |m| (a standard dot-product error bound; gamma covers the
# accumulation depth), and the division by S = sum(g) propagates it as
#   |dx_i| <= gamma * (G_i + |x_i| * sum(G)) / |S|.
GAMMA32 = 600 * np.finfo(np.float32).eps


def fp32_sweep_check(pkg, orc, t, s, m, got, gamma=GAMMA32):
    s32 = s.astype(np.float32).astype(np.float64)
    m32 = m.astype(np.float32).astype(np.float64)
    want, _, _ = orc.sweep_jacobi(t, s32, m32)
    worst = 0.0
    for q in range(t.nports):
        sl = slice(int(t.msg_off[q]), int(t.msg_off[q + 1]))
        g = orc.message(t, s32, m32, q)
        G = orc.message(t, np.abs(s32), np.abs(m32), q)
        S = g.sum()
        bound = gamma * (G + np.abs(want[sl]) * G.sum()) / abs(S)
        err = np.abs(got[sl] - want[sl])
        worst = max(worst, float(np.max(err / bound)))
    return worst
