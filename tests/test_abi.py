"""The C-ABI library loads and exports every symbol include/bp_mi355x.h declares; the
host-side logic (topologies, schedules, argument checks) behaves.  CPU only: no compute
call is made without a GPU."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bp_mi355x.h")


def declared_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(bp_\w+)\s*\(", text, flags=re.M)))


def test_header_declares_api():
    syms = declared_symbols()
    for must in ("bp_create", "bp_destroy", "bp_set_topology", "bp_upload_sites", "bp_run",
                 "bp_sweep", "bp_update_slot", "bp_lognorm", "bp_vertex_scalars",
                 "bp_edge_scalars", "bp_grad_lognorm", "bp_tfim_energy", "bp_polish_fp64"):
        assert must in syms


def test_library_exports_every_declared_symbol(pkg):
    lib = pkg._lib.load()
    raw = ctypes.CDLL(pkg._lib.LIB_PATH)
    for s in declared_symbols():
        assert hasattr(raw, s), f"missing export {s}"
        assert s in pkg._lib.SIGNATURES, f"no ctypes signature for {s}"
    assert lib.bp_version() >= 10000


def test_create_without_gpu_fails_cleanly(pkg):
    """With no HIP device the library reports an error instead of crashing."""
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    lib = pkg._lib.load()
    h = ctypes.c_void_p()
    rc = lib.bp_create(ctypes.byref(h), 0, 0)
    assert rc != 0
    assert lib.bp_last_error()


def test_topology_arrays_are_consistent(pkg):
    for g in (pkg.named_grid((5, 4)), pkg.named_grid((4, 4), periodic=True),
              pkg.named_comb_tree((4, 3)), pkg.heavy_hex_lattice(3, 4), pkg.named_path_graph(6)):
        t = pkg.build_topology(g, 3, 2)
        own = t.port_owner()
        for q in range(t.nports):
            r = t.port_rev[q]
            assert t.port_rev[r] == q
            assert t.port_nbr[r] == own[q]
            assert t.port_edge[r] == t.port_edge[q]
        assert t.nports == 2 * t.ne
        for u in range(t.nv):
            assert t.site_off[u + 1] - t.site_off[u] == np.prod(t.site_shape(u))


def test_grid_sizes(pkg):
    for L in (4, 32):
        g = pkg.named_grid((L, L))
        assert g.nv() == L * L and g.ne() == 2 * L * (L - 1)
    hh = pkg.heavy_hex_lattice(4, 5)
    assert max(hh.degree(v) for v in hh.vertices()) == 3


def test_forest_cover_sequence_covers_each_directed_edge_once(pkg):
    for g in (pkg.named_grid((4, 4)), pkg.named_grid((3, 3), periodic=True),
              pkg.named_comb_tree((4, 3))):
        seq = pkg.forest_cover_edge_sequence(g)
        assert len(seq) == 2 * g.ne()
        assert len(set(seq)) == len(seq)
        for a, b in seq:
            assert g.has_edge(a, b)


def test_batch_concat(pkg):
    ts = [pkg.build_topology(pkg.named_grid((L, L)), D, 2) for L, D in ((3, 2), (4, 3))]
    b = pkg.concat_topologies(ts)
    assert b.nnet == 2 and b.nv == 9 + 16
    assert int(b.site_off[-1]) == sum(int(t.site_off[-1]) for t in ts)
    assert list(np.unique(b.vertex_net)) == [0, 1]
