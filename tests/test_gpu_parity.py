"""GPU parity: the HIP engine against the FP64 CPU oracle, through the C ABI.

Tolerances: FP64 runs agree with the oracle within 1e-10 (messages and BP log-norm);
FP32 runs within the bounds stated per test.
"""
import math

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)


def _grid(pkg, L, D, d=2, periodic=False):
    return pkg.build_topology(pkg.named_grid((L, L), periodic=periodic), D, d)


@pytest.mark.parametrize("path", ["generic", "auto"])
def test_c1_fp64_converged_parity(pkg, orc, path):
    """Config 1: 4x4, d=2, D=2, FP64, identity init, Jacobi, tol 1e-10 on max change."""
    _say(f"case c1 fp64 path={path}")
    t = _grid(pkg, 4, 2)
    s = pkg.peps.random_sites(t, 1234)
    m0 = pkg.peps.identity_messages(t)
    want, sw_o, dl_o, at_o = orc.bp(t, s, m0, maxiter=100, tol=1e-10, criterion="maxabs")
    with pkg.Engine(t, s, dtype="f64", path=path if path == "generic" else "auto") as e:
        e.set_messages(m0)
        r = e.run(maxiter=100, tol=1e-10, criterion="maxabs")
        got = e.messages()
        assert r.sweeps[0] == sw_o and r.at_iteration[0] == at_o
        assert np.max(np.abs(got - want)) < 1e-10
        ln = e.lognorm()
        assert abs(ln - orc.bethe_free_energy(t, s, want)) < 1e-10
        # BP vs exact <psi|psi>: the loopy-BP gap is reported, not asserted to vanish
        exact = orc.norm_exact(t, s)
        _say(f"c1 lognorm bp={ln:.12f} exact={math.log(exact):.12f}")


def test_single_sweep_parity_fp64(pkg, orc):
    """One Jacobi sweep from random non-symmetric messages on a ragged lattice
    (degrees 2/3/4, mixed bond dims), before any convergence."""
    _say("case single sweep fp64 ragged")
    g = pkg.named_grid((5, 4))
    dims = {e: 2 + (i % 3) for i, e in enumerate(g.edges())}
    t = pkg.build_topology(g, dims, 2)
    s = pkg.peps.random_sites(t, 7)
    m = pkg.peps.random_messages(t, 7)
    want, fid, mab = orc.sweep_jacobi(t, s, m)
    with pkg.Engine(t, s, dtype="f64") as e:
        e.set_messages(m)
        e.sweep()
        got = e.messages()
        gfid, gmab = e.residuals()
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14)
    np.testing.assert_allclose(gmab, mab, rtol=1e-9, atol=1e-14)
    np.testing.assert_allclose(gfid, fid, rtol=1e-6, atol=1e-13)


@pytest.mark.parametrize("dims,D", [((2, 1), 2), ((4, 3), 3)])
def test_norm_tree_exact_on_gpu(pkg, orc, dims, D):
    _say(f"case norm tree {dims} D={D}")
    g = pkg.named_comb_tree(dims) if dims[1] > 1 else pkg.named_path_graph(dims[0])
    t = pkg.build_topology(g, D, 2)
    s = pkg.peps.random_sites(t, 11)
    with pkg.Engine(t, s, dtype="f64") as e:
        e.init_messages("ones")
        r = e.run(maxiter=60, tol=1e-15, criterion="maxabs")
        assert r.converged[0]
        z = math.exp(e.lognorm())
    assert z / orc.norm_exact(t, s) == pytest.approx(1.0, abs=1e-12)


def test_plain_tree_and_spin_ice_sequential(pkg, orc):
    """Plain networks: tree exactness (Jacobi) and spin ice via per-slot sequential
    updates in the forest-cover order (the reference's default schedule)."""
    _say("case plain tree + spin ice")
    g = pkg.named_comb_tree((4, 3))
    t = pkg.build_topology(g, 3, plain=True)
    s = pkg.peps.random_sites(t, 3)
    with pkg.Engine(t, s, dtype="f64") as e:
        e.init_messages("ones")
        e.run(maxiter=30, tol=1e-15, criterion="maxabs")
        z = e.lognorm()
    assert np.exp(complex(z)).real / orc.plain_exact(t, s) == pytest.approx(1.0, abs=1e-11)
    for n in (3, 4, 5):
        g = pkg.named_grid((n, n), periodic=True)
        t = pkg.build_topology(g, 2, plain=True)
        s = pkg.peps.spin_ice_sites(t)
        order = [t.slot(a, b) for a, b in pkg.forest_cover_edge_sequence(g)]
        with pkg.Engine(t, s, dtype="f64") as e:
            e.set_messages(pkg.peps.random_messages(t, 123))
            for _ in range(10):
                for q in order:
                    e.update_slot(q)
            z = math.exp(e.lognorm())
        assert z == pytest.approx(1.5 ** (n * n), rel=1e-8)


def test_scalars_parity(pkg, orc):
    _say("case scalars")
    t = _grid(pkg, 5, 3)
    s = pkg.peps.random_sites(t, 5)
    m = pkg.peps.random_messages(t, 5)
    with pkg.Engine(t, s, dtype="f64") as e:
        e.set_messages(m)
        np.testing.assert_allclose(e.vertex_scalars(), orc.vertex_scalars(t, s, m), rtol=1e-12)
        np.testing.assert_allclose(e.edge_scalars(), orc.edge_scalars(t, m), rtol=1e-12)
        assert e.lognorm() == pytest.approx(orc.bethe_free_energy(t, s, m), abs=1e-10)
        assert e.iterate_diff(m) == pytest.approx(0.0, abs=1e-14)


def test_graph_and_eager_agree(pkg, orc):
    _say("case graph vs eager")
    t = _grid(pkg, 6, 3)
    s = pkg.peps.random_sites(t, 3, scale=pkg.peps.peps_scale(2, 3))
    out = []
    for use_graph in (False, True, True):
        with pkg.Engine(t, s, dtype="f64") as e:
            e.init_messages("identity")
            r = e.run(maxiter=200, tol=1e-12, use_graph=use_graph)
            out.append((e.messages(), r.sweeps[0], r.at_iteration[0]))
    for m, sw, at in out[1:]:
        assert sw == out[0][1] and at == out[0][2]
        np.testing.assert_array_equal(m, out[0][0])


def test_batch_matches_individual_runs(pkg, orc):
    _say("case batch")
    tops = [_grid(pkg, L, D) for L, D in ((3, 2), (5, 3), (4, 2))]
    sites = [pkg.peps.random_sites(t, 40 + i, scale=pkg.peps.peps_scale(2, int(t.edge_dim[0])))
             for i, t in enumerate(tops)]
    b = pkg.concat_topologies(tops)
    with pkg.Engine(b, np.concatenate(sites), dtype="f64") as e:
        e.init_messages("identity")
        r = e.run(maxiter=300, tol=1e-12, criterion="maxabs")
        mb = e.messages()
        ln = e.lognorms()
    off = 0
    for i, (t, s) in enumerate(zip(tops, sites)):
        want, sw, dl, at = orc.bp(t, s, pkg.peps.identity_messages(t), maxiter=300, tol=1e-12,
                                  criterion="maxabs")
        n = int(t.msg_off[-1])
        assert r.sweeps[i] == sw and r.at_iteration[i] == at
        np.testing.assert_allclose(mb[off:off + n], want, atol=1e-10)
        assert ln[i] == pytest.approx(orc.bethe_free_energy(t, s, want), abs=1e-10)
        off += n


def test_fp32_parity_16x16_d8(pkg, orc):
    """FP32 contractions against the FP64 oracle on a 16x16 D=8 lattice."""
    _say("case fp32 16x16 D=8")
    t = _grid(pkg, 16, 8)
    s = pkg.peps.random_sites(t, 1, scale=pkg.peps.peps_scale(2, 8))
    m0 = pkg.peps.identity_messages(t)
    want, sw, dl, at = orc.bp(t, s, m0, maxiter=200, tol=1e-7, nthreads=8)
    with pkg.Engine(t, s, dtype="f32") as e:
        e.set_messages(m0)
        r = e.run(maxiter=200, tol=1e-7)
        got = e.messages()
        ln = e.lognorm()
        _say(f"fp32 path={e.path()} sweeps={r.sweeps[0]} oracle sweeps={sw}")
    rel = np.max(np.abs(got - want)) / np.max(np.abs(want))
    assert rel < 1e-4
    lw = orc.bethe_free_energy(t, s, want)
    assert abs(ln - lw) / abs(lw) < 1e-5


def test_errors_are_reported(pkg):
    _say("case errors")
    t = _grid(pkg, 3, 2)
    s = pkg.peps.random_sites(t, 0)
    e = pkg.Engine(t, s, dtype="f64")
    with pytest.raises(pkg._lib.BPError):
        e.run()  # neither maxiter nor tol
    with pytest.raises(pkg._lib.BPError):
        e.update_slot(10 ** 6)
    e.close()


# ---------------------------------------------------------------- vertex (MFMA) path

def _rel(a, b):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))


from helpers import fp32_sweep_check as _fp32_sweep_check  # noqa: E402


@pytest.mark.parametrize("shape,d", [((6, 5), 2), ((4, 4), 1), ((3, 5), 3), ((5, 4), 4)])
def test_vertex_single_sweep_parity(pkg, orc, shape, d):
    """One Jacobi sweep on the D=8 vertex path (FP32 MFMA) from random non-symmetric
    messages against the FP64 oracle; corners, edges and interior vertices all present."""
    _say(f"case vertex single sweep {shape} d={d}")
    t = pkg.build_topology(pkg.named_grid(shape), 8, d)
    s = pkg.peps.random_sites(t, 17, scale=pkg.peps.peps_scale(d, 8))
    m = pkg.peps.random_messages(t, 17)
    want, fid, mab = orc.sweep_jacobi(t, s, m)
    with pkg.Engine(t, s, dtype="f32", path="vertex") as e:
        assert e.path() == "vertex"
        e.set_messages(m)
        e.sweep()
        got = e.messages()
    # every entry within the FP32 error bound (a wrong index gives O(1) errors)
    assert _fp32_sweep_check(pkg, orc, t, s, m, got) < 1.0
    # and on well-conditioned (identity) incoming messages a plain relative check
    mi = pkg.peps.identity_messages(t)
    want_i, _, _ = orc.sweep_jacobi(t, s, mi)
    with pkg.Engine(t, s, dtype="f32", path="vertex") as e:
        e.set_messages(mi)
        e.sweep()
        assert _rel(e.messages(), want_i) < 1e-5


def test_vertex_path_graph_degree2(pkg, orc):
    _say("case vertex path graph")
    t = pkg.build_topology(pkg.named_path_graph(5), 8, 2)
    s = pkg.peps.random_sites(t, 3)
    m = pkg.peps.random_messages(t, 3)
    with pkg.Engine(t, s, dtype="f32", path="vertex") as e:
        e.set_messages(m)
        e.sweep()
        assert _fp32_sweep_check(pkg, orc, t, s, m, e.messages()) < 1.0


def test_vertex_matches_generic_fp32(pkg):
    _say("case vertex vs generic fp32")
    t = _grid(pkg, 8, 8)
    s = pkg.peps.random_sites(t, 5, scale=pkg.peps.peps_scale(2, 8))
    res = {}
    for path in ("generic", "vertex"):
        with pkg.Engine(t, s, dtype="f32", path=path) as e:
            e.init_messages("identity")
            r = e.run(maxiter=30)
            res[path] = (e.messages(), e.lognorm())
    assert _rel(res["vertex"][0], res["generic"][0]) < 1e-5
    assert res["vertex"][1] == pytest.approx(res["generic"][1], rel=1e-6)


def test_vertex_batch_with_masks(pkg, orc):
    _say("case vertex batch")
    tops = [_grid(pkg, L, 8) for L in (3, 6, 4)]
    sites = [pkg.peps.random_sites(t, 60 + i, scale=pkg.peps.peps_scale(2, 8))
             for i, t in enumerate(tops)]
    b = pkg.concat_topologies(tops)
    with pkg.Engine(b, np.concatenate(sites), dtype="f32") as e:
        assert e.path() == "vertex"
        e.init_messages("identity")
        r = e.run(maxiter=300, tol=1e-6, criterion="maxabs")
        mb = e.messages()
    off = 0
    for i, (t, s) in enumerate(zip(tops, sites)):
        want, sw, dl, at = orc.bp(t, s, pkg.peps.identity_messages(t), maxiter=300, tol=1e-6,
                                  criterion="maxabs")
        n = int(t.msg_off[-1])
        assert r.converged[i]
        assert abs(int(r.sweeps[i]) - sw) <= 2
        assert _rel(mb[off:off + n], want) < 1e-5
        off += n


# ---------------------------------------------------------------- BF16 + FP64 polish

@pytest.mark.parametrize("variant", [0, 1, 5])
def test_bf16_single_sweep_within_bound(pkg, orc, variant):
    """BF16 operands / FP32 accumulation: one sweep within the error bound with the BF16
    unit roundoff (gamma = 24 u, u = 2^-8, covering the ~10 roundings along a product
    chain), from well-conditioned (identity) messages; the D <= 8 sweep kernels with two LDS
    buffers per wave, one buffer and two waves per SIMD, and the open contractions on MFMA
    (BF16 mode products, FP32 opens)."""
    _say(f"case bf16 single sweep, kernel variant {variant}")
    t = pkg.build_topology(pkg.named_grid((6, 5)), 8, 2)
    s = pkg.peps.random_sites(t, 19, scale=pkg.peps.peps_scale(2, 8))
    m = pkg.peps.identity_messages(t)
    with pkg.Engine(t, s, dtype="bf16") as e:
        assert e.path() == "vertex"
        e.set_option("vertex_variant", variant)
        e.set_messages(m)
        e.sweep()
        got = e.messages()
    assert _fp32_sweep_check(pkg, orc, t, s, m, got, gamma=24 * 2.0 ** -8) < 1.0


@pytest.mark.parametrize("variant", [0, 1, 5])
def test_bf16_converged_then_fp64_polish(pkg, orc, variant):
    """Mixed-precision workload at D = 8: converge in BF16, report the error against the
    FP64 oracle, then 3 FP64 polishing sweeps shrink it (stated bounds: BF16 messages
    within 5e-2 relative, log-norm within 1e-2 relative; after polishing, messages at
    least 4x closer and log-norm within 1e-4 relative); the three BF16 D <= 8 sweep
    kernels."""
    _say(f"case bf16 converged + polish, kernel variant {variant}")
    t = _grid(pkg, 12, 8)
    s = pkg.peps.random_sites(t, 23, scale=pkg.peps.peps_scale(2, 8))
    m0 = pkg.peps.identity_messages(t)
    want, _, _, at = orc.bp(t, s, m0, maxiter=500, tol=1e-14, criterion="maxabs", nthreads=8)
    assert at > 0
    lw = orc.bethe_free_energy(t, s, want)
    with pkg.Engine(t, s, dtype="bf16") as e:
        e.set_option("vertex_variant", variant)
        e.set_messages(m0)
        r = e.run(maxiter=200, tol=1e-5)
        m_bf = e.messages()
        ln_bf = e.lognorm()
        e.polish_fp64(3)
        m_pol = e.messages()
        ln_pol = e.lognorm()
    err_bf = _rel(m_bf, want)
    err_pol = _rel(m_pol, want)
    _say(f"bf16 sweeps={r.sweeps[0]} msg rel err bf16={err_bf:.3e} polished={err_pol:.3e} "
         f"lognorm rel err bf16={abs(ln_bf - lw) / abs(lw):.3e} "
         f"polished={abs(ln_pol - lw) / abs(lw):.3e}")
    assert err_bf < 5e-2
    assert abs(ln_bf - lw) / abs(lw) < 1e-2
    assert err_pol < err_bf / 4
    assert abs(ln_pol - lw) / abs(lw) < 1e-4


def test_bf16_generic_path_single_sweep(pkg, orc):
    """BF16 operands on the generic kernel (bond dimension 10): one sweep from identity
    messages within the BF16 error bound (gamma = 24 u, u = 2^-8)."""
    _say("case bf16 generic single sweep D=10")
    t = pkg.build_topology(pkg.named_grid((4, 3)), 10, 2)
    s = pkg.peps.random_sites(t, 37, scale=pkg.peps.peps_scale(2, 10))
    m = pkg.peps.identity_messages(t)
    with pkg.Engine(t, s, dtype="bf16", path="generic") as e:
        assert e.path() == "generic"
        e.set_messages(m)
        e.sweep()
        got = e.messages()
    assert _fp32_sweep_check(pkg, orc, t, s, m, got, gamma=24 * 2.0 ** -8) < 1.0
    # the same sweep in FP32 is more accurate than the BF16 one: the rounding is real
    with pkg.Engine(t, s, dtype="f32") as e:
        e.set_messages(m)
        e.sweep()
        got32 = e.messages()
    want, _, _ = orc.sweep_jacobi(t, s, m)
    assert _rel(got32, want) < _rel(got, want)


def test_bf16_d16_converged_then_fp64_polish(pkg, orc):
    """The large-bond mixed-precision workload at reduced lattice size (5x5, D = 16): BF16
    operands with FP32 sums to convergence, then 3 FP64 polishing sweeps; errors against
    the FP64 oracle at each stage (stated bounds: BF16 messages within 5e-2 and log-norm
    within 1e-2 relative; polished messages at least 4x closer, log-norm within 1e-4)."""
    _say("case bf16 D=16 converged + polish")
    t = _grid(pkg, 5, 16)
    s = pkg.peps.random_sites(t, 41, scale=pkg.peps.peps_scale(2, 16))
    m0 = pkg.peps.identity_messages(t)
    want, _, _, at = orc.bp(t, s, m0, maxiter=500, tol=1e-13, criterion="maxabs", nthreads=8)
    assert at > 0
    lw = orc.bethe_free_energy(t, s, want)
    with pkg.Engine(t, s, dtype="bf16") as e:
        assert e.path() == "vertex"  # the bond-dimension-16 MFMA kernel
        e.set_messages(m0)
        r = e.run(maxiter=200, tol=1e-5)
        m_bf = e.messages()
        ln_bf = e.lognorm()
        e.polish_fp64(3)
        m_pol = e.messages()
        ln_pol = e.lognorm()
    err_bf, err_pol = _rel(m_bf, want), _rel(m_pol, want)
    _say(f"D=16 bf16 sweeps={r.sweeps[0]} msg err bf16={err_bf:.3e} polished={err_pol:.3e} "
         f"lognorm err bf16={abs(ln_bf - lw) / abs(lw):.3e} "
         f"polished={abs(ln_pol - lw) / abs(lw):.3e}")
    assert r.converged[0]
    assert err_bf < 5e-2
    assert abs(ln_bf - lw) / abs(lw) < 1e-2
    assert err_pol < err_bf / 4
    assert abs(ln_pol - lw) / abs(lw) < 1e-4


# ---------------------------------------------------------------- bond dimensions below 8

@pytest.mark.parametrize("shape,D,d", [((5, 4), 6, 2), ((4, 4), 3, 2), ((3, 5), 5, 3),
                                       ((4, 3), 7, 1), ((3, 3), 1, 2)])
def test_vertex_padded_bond_single_sweep(pkg, orc, shape, D, d):
    """Bond dimensions below 8 run on the vertex kernels zero-padded to 8: one sweep from
    random messages within the FP32 error bound, and identity messages to 1e-5."""
    _say(f"case vertex padded D={D} {shape} d={d}")
    t = pkg.build_topology(pkg.named_grid(shape), D, d)
    s = pkg.peps.random_sites(t, 29, scale=pkg.peps.peps_scale(d, D))
    m = pkg.peps.random_messages(t, 29)
    with pkg.Engine(t, s, dtype="f32") as e:
        assert e.path() == "vertex"
        e.set_messages(m)
        e.sweep()
        got = e.messages()
    assert _fp32_sweep_check(pkg, orc, t, s, m, got) < 1.0
    mi = pkg.peps.identity_messages(t)
    want_i, _, _ = orc.sweep_jacobi(t, s, mi)
    with pkg.Engine(t, s, dtype="f32") as e:
        e.set_messages(mi)
        e.sweep()
        assert _rel(e.messages(), want_i) < 1e-5


def test_vertex_mixed_bond_dims_in_one_network(pkg, orc):
    """Edges of one network with different bond dimensions (2..8)."""
    _say("case vertex mixed bond dims")
    g = pkg.named_grid((4, 5))
    dims = {e: 2 + (i % 7) for i, e in enumerate(g.edges())}
    t = pkg.build_topology(g, dims, 2)
    s = pkg.peps.random_sites(t, 31)
    m0 = pkg.peps.identity_messages(t)
    want, sw, _, at = orc.bp(t, s, m0, maxiter=500, tol=1e-9, criterion="fidelity", nthreads=8)
    with pkg.Engine(t, s, dtype="f32") as e:
        assert e.path() == "vertex"
        e.set_messages(m0)
        r = e.run(maxiter=500, tol=1e-9, criterion="fidelity")
        got = e.messages()
        ln = e.lognorm()
    assert r.converged[0]
    assert _rel(got, want) < 1e-4
    lw = orc.bethe_free_energy(t, s, want)
    assert abs(ln - lw) / abs(lw) < 1e-5


def test_vertex_padded_matches_generic_d6(pkg):
    """The 16x16 D=6 differentiation lattice: vertex (padded) and generic FP32 paths agree
    on the converged messages, the log-norm and the TFIM gradient."""
    _say("case vertex vs generic 16x16 D=6")
    t = _grid(pkg, 16, 6)
    s = pkg.peps.random_sites(t, 33, scale=pkg.peps.peps_scale(2, 6))
    out = {}
    for path in ("generic", "auto"):
        with pkg.Engine(t, s, dtype="f32", path=path) as e:
            e.init_messages("identity")
            r = e.run(maxiter=300, tol=1e-12, criterion="fidelity")
            assert r.converged[0]
            E, g, it, res = e.tfim_energy(1.0, 3.0, grad=True, adj_tol=1e-6, adj_maxiter=500)
            out[path] = (e.path(), e.messages(), e.lognorm(), E, g)
    assert out["auto"][0] == "vertex"
    assert _rel(out["auto"][1], out["generic"][1]) < 1e-4
    assert out["auto"][2] == pytest.approx(out["generic"][2], rel=1e-6)
    assert out["auto"][3] == pytest.approx(out["generic"][3], rel=1e-6)
    assert _rel(out["auto"][4], out["generic"][4]) < 2e-4


@pytest.mark.parametrize("shape,D,d", [((6, 5), 8, 2), ((4, 4), 8, 1), ((3, 5), 6, 3),
                                       ((5, 4), 8, 4)])
def test_vertex_low_lds_variant(pkg, orc, shape, D, d):
    """The one-buffer-per-wave FP32 vertex kernel (two waves per SIMD), the kernels with MFMA
    open contractions, and the one with half of them on MFMA in place compute what the
    default kernel computes: one sweep within the FP32 bound and 20 sweeps agreeing."""
    _say(f"case vertex kernel variants {shape} D={D} d={d}")
    t = pkg.build_topology(pkg.named_grid(shape), D, d)
    s = pkg.peps.random_sites(t, 43, scale=pkg.peps.peps_scale(d, D))
    m = pkg.peps.random_messages(t, 43)
    res = {}
    for v in (0, 1, 2, 3, 4, 5):
        with pkg.Engine(t, s, dtype="f32") as e:
            e.set_option("vertex_variant", v)
            e.set_messages(m)
            e.sweep()
            one = e.messages()
            assert _fp32_sweep_check(pkg, orc, t, s, m, one) < 1.0
            e.init_messages("identity")
            e.run(maxiter=20)
            res[v] = (one, e.messages())
    for v in (1, 2, 3, 4, 5):
        # one sweep from random non-symmetric messages: each variant is within the FP32
        # bound above (ill-conditioned messages make a flat relative check loose); after
        # 20 sweeps from the identity the messages are well conditioned
        assert _rel(res[v][0], res[0][0]) < 1e-4
        assert _rel(res[v][1], res[0][1]) < 1e-5


def test_kernel_options_round_trip(pkg):
    """bp_get_option reads back what bp_set_option set, and rejects unknown names."""
    _say("case kernel options round trip")
    t = pkg.build_topology(pkg.named_grid((3, 3)), 8, 2)
    s = pkg.peps.random_sites(t, 44, scale=pkg.peps.peps_scale(2, 8))
    with pkg.Engine(t, s, dtype="f32") as e:
        for name, value in (("vertex_variant", 1), ("vertex_variant", 2), ("vertex_variant", 3),
                            ("vertex_variant", 4), ("vertex_variant", 5), ("vertex_variant", 0),
                            ("vertex16_waves", 8), ("vertex_bond_mask", 0x1FE)):
            e.set_option(name, value)
            assert e.get_option(name) == value
        with pytest.raises(pkg._lib.BPError, match="unknown option"):
            e.get_option("no_such_option")


# ---------------------------------------------------------------- bond dimensions 9..12 (FP32)

@pytest.mark.parametrize("shape,D,d", [((4, 3), 12, 2), ((3, 3), 10, 2), ((3, 4), 9, 1),
                                       ((3, 3), 12, 3), ((2, 5), 11, 4), ((3, 4), 10, 4)])
def test_vertex12_single_sweep_fp32(pkg, orc, shape, D, d):
    """The FP32 MFMA kernel for bond dimensions 9..12 (padded to 12, or to 10 when every bond
    dimension of the vertex is <= 10): one sweep from random
    non-symmetric messages within the FP32 error bound, and from identity messages within
    1e-5 relative of the FP64 oracle."""
    _say(f"case vertex12 single sweep {shape} D={D} d={d}")
    t = pkg.build_topology(pkg.named_grid(shape), D, d)
    s = pkg.peps.random_sites(t, 61, scale=pkg.peps.peps_scale(d, D))
    m = pkg.peps.random_messages(t, 61)
    with pkg.Engine(t, s, dtype="f32") as e:
        assert e.path() == "vertex"
        e.set_messages(m)
        e.sweep()
        got = e.messages()
    assert _fp32_sweep_check(pkg, orc, t, s, m, got) < 1.0
    mi = pkg.peps.identity_messages(t)
    want_i, _, _ = orc.sweep_jacobi(t, s, mi)
    with pkg.Engine(t, s, dtype="f32") as e:
        e.set_messages(mi)
        e.sweep()
        err = _rel(e.messages(), want_i)
    _say(f"D={D} d={d}: identity-message sweep rel err {err:.2e}")
    assert err < 1e-5


@pytest.mark.parametrize("kind,D", [("heavyhex", 12), ("heavyhex", 10), ("path", 11)])
def test_vertex12_low_degree_single_sweep(pkg, orc, kind, D):
    """Degree <= 3 vertices of the FP32 large-bond bucket (their own 12^3-slice kernel):
    heavy-hex lattices (degrees 2 and 3) and a path graph (degrees 1 and 2), one sweep from
    random messages within the FP32 bound and from identity messages within 1e-5."""
    _say(f"case vertex12 low degree {kind} D={D}")
    g = pkg.heavy_hex_lattice(3, 3) if kind == "heavyhex" else pkg.named_path_graph(5)
    t = pkg.build_topology(g, D, 2)
    s = pkg.peps.random_sites(t, 64, scale=pkg.peps.peps_scale(2, D))
    m = pkg.peps.random_messages(t, 64)
    with pkg.Engine(t, s, dtype="f32") as e:
        assert e.path() == "vertex"
        e.set_messages(m)
        e.sweep()
        got = e.messages()
    assert _fp32_sweep_check(pkg, orc, t, s, m, got) < 1.0
    mi = pkg.peps.identity_messages(t)
    want_i, _, _ = orc.sweep_jacobi(t, s, mi)
    with pkg.Engine(t, s, dtype="f32") as e:
        e.set_messages(mi)
        e.sweep()
        assert _rel(e.messages(), want_i) < 1e-5


@pytest.mark.parametrize("kind,D,d", [("heavyhex", 8, 2), ("heavyhex", 5, 2), ("path", 8, 3),
                                       ("heavyhex", 7, 1)])
def test_vertex_k3_d8_single_sweep(pkg, orc, kind, D, d):
    """FP32 networks of bond dimension <= 8 whose vertices all have degree <= 3 (heavy-hex
    lattices, paths) sweep on the 8^3-slice kernel: one sweep from random messages within
    the FP32 bound, from identity messages within 1e-5, and 30 sweeps agree with the
    generic FP32 kernel within 1e-5 (log-norm 1e-6)."""
    _say(f"case vertex k3 D<=8 {kind} D={D} d={d}")
    g = pkg.heavy_hex_lattice(3, 4) if kind == "heavyhex" else pkg.named_path_graph(6)
    t = pkg.build_topology(g, D, d)
    s = pkg.peps.random_sites(t, 65, scale=pkg.peps.peps_scale(d, D))
    m = pkg.peps.random_messages(t, 65)
    with pkg.Engine(t, s, dtype="f32") as e:
        assert e.path() == "vertex"
        e.set_messages(m)
        e.sweep()
        got = e.messages()
    assert _fp32_sweep_check(pkg, orc, t, s, m, got) < 1.0
    mi = pkg.peps.identity_messages(t)
    want_i, _, _ = orc.sweep_jacobi(t, s, mi)
    res = {}
    for path in ("auto", "generic"):
        with pkg.Engine(t, s, dtype="f32", path=path) as e:
            e.set_messages(mi)
            e.sweep()
            if path == "auto":
                assert _rel(e.messages(), want_i) < 1e-5
            e.run(maxiter=29)
            res[path] = (e.messages(), e.lognorm())
    assert _rel(res["auto"][0], res["generic"][0]) < 1e-5
    assert res["auto"][1] == pytest.approx(res["generic"][1], rel=1e-6)


def test_vertex12_mixed_dims_and_generic_agree(pkg, orc):
    """Bond dimensions 9..12 within one network: within the FP32 bound, and the MFMA kernel
    and the generic FP32 kernel agree after 20 sweeps."""
    _say("case vertex12 mixed bond dims")
    g = pkg.named_grid((3, 4))
    dims = {e: 9 + (i % 4) for i, e in enumerate(g.edges())}
    t = pkg.build_topology(g, dims, 2)
    s = pkg.peps.random_sites(t, 62)
    m = pkg.peps.random_messages(t, 62)
    with pkg.Engine(t, s, dtype="f32") as e:
        assert e.path() == "vertex"
        e.set_messages(m)
        e.sweep()
        assert _fp32_sweep_check(pkg, orc, t, s, m, e.messages()) < 1.0
    res = {}
    for path in ("generic", "auto"):
        with pkg.Engine(t, s, dtype="f32", path=path) as e:
            e.init_messages("identity")
            e.run(maxiter=20)
            res[path] = (e.messages(), e.lognorm())
    assert _rel(res["auto"][0], res["generic"][0]) < 1e-5
    assert res["auto"][1] == pytest.approx(res["generic"][1], rel=1e-6)


def test_vertex12_side10_and_side12_in_one_network(pkg, orc):
    """Degree-4 vertices whose bond dimensions are all <= 10 run at side 10, the rest at side
    12, in the same sweep (a 5x5 lattice with D = 10 on every edge with an end in the two
    left columns, D = 12 elsewhere, d = 3): one sweep within the FP32 bound, then converged
    within 1e-4 (messages) and 1e-5 (log-norm) of the FP64 oracle."""
    _say("case vertex12 side 10 and 12 in one network")
    g = pkg.named_grid((5, 5))
    dims = {e: 10 if min(e[0][1], e[1][1]) <= 2 else 12 for e in g.edges()}  # columns 1..5
    t = pkg.build_topology(g, dims, 3)
    s = pkg.peps.random_sites(t, 66, scale=pkg.peps.peps_scale(3, 11))
    m = pkg.peps.random_messages(t, 66)
    with pkg.Engine(t, s, dtype="f32") as e:
        assert e.path() == "vertex"
        e.set_messages(m)
        e.sweep()
        assert _fp32_sweep_check(pkg, orc, t, s, m, e.messages()) < 1.0
    m0 = pkg.peps.identity_messages(t)
    want, _, _, _ = orc.bp(t, s, m0, maxiter=400, tol=1e-12, criterion="maxabs", nthreads=8)
    with pkg.Engine(t, s, dtype="f32") as e:
        e.set_messages(m0)
        r = e.run(maxiter=400, tol=1e-10, criterion="fidelity")
        got, ln = e.messages(), e.lognorm()
    assert r.converged[0]
    lw = orc.bethe_free_energy(t, s, want)
    _say(f"msg rel err {_rel(got, want):.2e}, log-norm rel err {abs(ln - lw) / abs(lw):.2e}")
    assert _rel(got, want) < 1e-4
    assert abs(ln - lw) / abs(lw) < 1e-5


def test_vertex12_converged_fp32_vs_oracle(pkg, orc):
    """A 5x5 D=12 lattice converged in FP32 on the large-bond kernel against the FP64
    oracle: messages within 1e-4 relative, log-norm within 1e-5."""
    _say("case vertex12 converged 5x5 D=12")
    t = _grid(pkg, 5, 12)
    s = pkg.peps.random_sites(t, 63, scale=pkg.peps.peps_scale(2, 12))
    m0 = pkg.peps.identity_messages(t)
    want, sw, _, _ = orc.bp(t, s, m0, maxiter=400, tol=1e-12, criterion="maxabs", nthreads=8)
    with pkg.Engine(t, s, dtype="f32") as e:
        assert e.path() == "vertex"
        e.set_messages(m0)
        r = e.run(maxiter=400, tol=1e-10, criterion="fidelity")
        got = e.messages()
        ln = e.lognorm()
    assert r.converged[0]
    err = _rel(got, want)
    lw = orc.bethe_free_energy(t, s, want)
    _say(f"sweeps {r.sweeps[0]} (oracle {sw}), msg rel err {err:.2e}, "
         f"log-norm rel err {abs(ln - lw) / abs(lw):.2e}")
    assert err < 1e-4
    assert abs(ln - lw) / abs(lw) < 1e-5


# ---------------------------------------------------------------- bond dimensions 9..16 (BF16)

@pytest.mark.parametrize("shape,D,d", [((4, 3), 16, 2), ((3, 3), 12, 2), ((3, 4), 10, 1),
                                       ((3, 3), 16, 3)])
def test_vertex16_single_sweep_within_bf16_bound(pkg, orc, shape, D, d):
    """The BF16 MFMA kernel for bond dimensions 9..16 (padded to 16): one sweep from identity
    messages within the BF16 error bound (gamma = 24 u, u = 2^-8)."""
    _say(f"case vertex16 single sweep {shape} D={D} d={d}")
    t = pkg.build_topology(pkg.named_grid(shape), D, d)
    s = pkg.peps.random_sites(t, 47, scale=pkg.peps.peps_scale(d, D))
    m = pkg.peps.identity_messages(t)
    with pkg.Engine(t, s, dtype="bf16") as e:
        assert e.path() == "vertex"
        e.set_messages(m)
        e.sweep()
        got = e.messages()
    assert _fp32_sweep_check(pkg, orc, t, s, m, got, gamma=24 * 2.0 ** -8) < 1.0
    want, _, _ = orc.sweep_jacobi(t, s, m)
    _say(f"D={D}: rel err {_rel(got, want):.2e}")


def test_vertex16_random_messages_and_mixed_dims(pkg, orc):
    """Non-symmetric random incoming messages and bond dimensions 9..16 within one network."""
    _say("case vertex16 random messages, mixed bond dims")
    g = pkg.named_grid((3, 4))
    dims = {e: 9 + (i % 8) for i, e in enumerate(g.edges())}
    t = pkg.build_topology(g, dims, 2)
    s = pkg.peps.random_sites(t, 53)
    m = pkg.peps.random_messages(t, 53)
    with pkg.Engine(t, s, dtype="bf16") as e:
        assert e.path() == "vertex"
        e.set_messages(m)
        e.sweep()
        got = e.messages()
    assert _fp32_sweep_check(pkg, orc, t, s, m, got, gamma=24 * 2.0 ** -8) < 1.0


def test_bf16_batch_all_kernels(pkg, orc):
    """One BF16 batch on the three kernels at once: D = 8 (vertex), D = 16 (vertex16) and
    D = 20 (generic), converged with per-network masks against the FP64 oracle."""
    _say("case bf16 batch across kernels")
    specs = [((4, 4), 8), ((3, 3), 16), ((2, 3), 20)]
    tops, sites = [], []
    for k, (shape, D) in enumerate(specs):
        t = pkg.build_topology(pkg.named_grid(shape), D, 2)
        tops.append(t)
        sites.append(pkg.peps.random_sites(t, 90 + k, scale=pkg.peps.peps_scale(2, D)))
    b = pkg.concat_topologies(tops)
    with pkg.Engine(b, np.concatenate(sites), dtype="bf16") as e:
        assert e.path() == "mixed"
        e.init_messages("identity")
        r = e.run(maxiter=300, tol=1e-5, criterion="fidelity")
        mb = e.messages()
    off = 0
    for i, (t, s) in enumerate(zip(tops, sites)):
        want, _, _, at = orc.bp(t, s, pkg.peps.identity_messages(t), maxiter=500, tol=1e-13,
                                criterion="maxabs", nthreads=8)
        n = int(t.msg_off[-1])
        assert r.converged[i]
        err = _rel(mb[off:off + n], want)
        _say(f"network {i} D={specs[i][1]}: sweeps {r.sweeps[i]} rel err {err:.2e}")
        assert err < 5e-2
        off += n


@pytest.mark.parametrize("shape,D,d", [((4, 5), 8, 2), ((3, 3), 6, 1), ((3, 4), 7, 3)])
def test_vertex_scalars_fp32_vertex_path(pkg, orc, shape, D, d):
    """FP32 vertex scalars on the D <= 8 vertex path (the message-form kernel's
    Z_u = <G_{u, port 0}, m_in,0>, degrees 2..4) against the FP64 oracle and the generic FP32
    kernel (1e-5 relative to the largest), and the log-norm within 1e-6 relative."""
    _say(f"case vertex scalars fp32 {shape} D={D} d={d}")
    t = pkg.build_topology(pkg.named_grid(shape), D, d)
    s = pkg.peps.random_sites(t, 67, scale=pkg.peps.peps_scale(d, D))
    with pkg.Engine(t, s, dtype="f64", path="generic") as e:  # PSD messages: real logs
        e.init_messages("identity")
        e.run(maxiter=4)
        m = e.messages()
    res = {}
    for path in ("auto", "generic"):
        with pkg.Engine(t, s, dtype="f32", path=path) as e:
            e.set_messages(m)
            res[path] = (e.path(), e.vertex_scalars(), e.lognorm())
    assert res["auto"][0] == "vertex" and res["generic"][0] == "generic"
    want = orc.vertex_scalars(t, s, m.astype(np.float32).astype(np.float64))
    assert _rel(res["auto"][1], want) < 1e-5
    assert _rel(res["auto"][1], res["generic"][1]) < 1e-5
    lw = orc.bethe_free_energy(t, s, m.astype(np.float32).astype(np.float64))
    assert abs(res["auto"][2] - lw) < 1e-6 * abs(lw)
