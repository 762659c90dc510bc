"""Seeded synthetic inputs: random PEPS site tensors and message initialisations.

Random numbers come from numpy's counter-based Philox4x32-10 generator (Salmon et al.,
SC'11) keyed by (seed, network index), so every process that builds the same workload
gets bit-identical inputs.  Entries are iid N(0, 1), optionally scaled by
1/sqrt(d * D^4) as the 32x32 and batched workloads specify.
"""
from __future__ import annotations

import math

import numpy as np

from .graphs import Topology


def philox(seed: int, stream: int = 0) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(key=[int(seed) & 0xFFFFFFFFFFFFFFFF,
                                                     int(stream) & 0xFFFFFFFFFFFFFFFF]))


def random_sites(topo: Topology, seed: int = 0, scale: float | None = None,
                 stream: int = 0) -> np.ndarray:
    """Flat FP64 site-tensor buffer (layout of `topo.site_off`), iid N(0,1) * scale.
    `scale=None` uses 1; pass `peps_scale(d, D)` for the 1/sqrt(d D^4) convention."""
    rng = philox(seed, stream)
    x = rng.standard_normal(int(topo.site_off[-1]))
    if scale is not None:
        x *= scale
    return x


def peps_scale(d: int, D: int) -> float:
    return 1.0 / math.sqrt(d * D ** 4)


def identity_messages(topo: Topology) -> np.ndarray:
    """message_environment(one, nn): every D x D message is the identity
    (src/beliefpropagation/messagecache.jl:227-229 with f = one)."""
    out = np.zeros(int(topo.msg_off[-1]))
    for q in range(topo.nports):
        D = int(topo.edge_dim[topo.port_edge[q]])
        if topo.plain:
            out[topo.msg_off[q]:topo.msg_off[q + 1]] = 1.0
        else:
            out[topo.msg_off[q]:topo.msg_off[q + 1]] = np.eye(D).ravel()
    return out


def ones_messages(topo: Topology) -> np.ndarray:
    """`fill!(msg, true)` initialisation used by the reference tests."""
    return np.ones(int(topo.msg_off[-1]))


def random_messages(topo: Topology, seed: int = 0, stream: int = 1) -> np.ndarray:
    """Uniform [0,1) entries (`rand(rng, T, ...)` as in the spin-ice test)."""
    return philox(seed, stream).random(int(topo.msg_off[-1]))


def uneven_batch(seed: int = 0, nnet: int = 64, L_range=(8, 48), D_choices=(4, 6, 8, 10, 12),
                 d: int = 2):
    """The batched uneven workload: `nnet` independent PEPS norm networks, L uniform on
    L_range (inclusive), D uniform on D_choices, square or heavy-hex lattice chosen by the
    seed stream (heavy-hex: max(2, L // 2) rows x L columns of degree-3 sites plus the
    degree-2 bridge sites), sites iid N(0,1) / sqrt(d D^4).  Returns (topologies, sites)."""
    from .graphs import build_topology, heavy_hex_lattice, named_grid

    rng = philox(seed, 7)
    tops, sites = [], []
    for n in range(nnet):
        L = int(rng.integers(L_range[0], L_range[1] + 1))
        D = int(D_choices[int(rng.integers(0, len(D_choices)))])
        hh = bool(rng.integers(0, 2))
        g = heavy_hex_lattice(max(2, L // 2), L) if hh else named_grid((L, L))
        t = build_topology(g, D, d)
        tops.append(t)
        sites.append(random_sites(t, seed * 1000003 + n, scale=peps_scale(d, D)))
    return tops, sites


def spin_ice_sites(topo: Topology) -> np.ndarray:
    """Spin-ice vertex tensors (test/test_beliefpropagation.jl:16-36): a degree-4 vertex
    carries t[i,j,k,l] = 1 when i+j+k+l == 2 (two in, two out), else 0."""
    out = np.zeros(int(topo.site_off[-1]))
    for u in range(topo.nv):
        shp = topo.site_shape(u)[1:]
        if len(shp) != 4 or any(s != 2 for s in shp):
            raise ValueError("spin ice needs degree-4 vertices with dimension-2 links")
        t = np.zeros((2, 2, 2, 2))
        for idx in np.ndindex(2, 2, 2, 2):
            if sum(idx) == 2:
                t[idx] = 1.0
        out[topo.site_off[u]:topo.site_off[u + 1]] = t.ravel()
    return out
