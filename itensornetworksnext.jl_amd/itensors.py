"""Named tensors and networks, the containers the reference's BP path consumes.

Mirrors src/tensornetwork.jl (ITensorNetwork: vertex -> tensor, edges inferred from
index names shared by exactly two tensors, tensornetwork.jl:138-173), src/normnetwork.jl
(NormNetwork: lazy <psi|psi> whose bra layer is the conjugated ket with link names
renamed, normnetwork.jl:13-93) and the link/site queries of src/abstracttensornetwork.jl
(linkinds, siteinds: abstracttensornetwork.jl:52-95).  Flattening to the engine's port
topology (`to_topology`) permutes each tensor to [site indices][port 0]...[port k-1].
"""
from __future__ import annotations

import itertools
from typing import Dict, Hashable, Iterable, List, Sequence, Tuple

import numpy as np

from .graphs import NamedGraph, Topology, build_topology

_uid = itertools.count(1)


class Index:
    """A named dimension (ITensorBase `Index`): equal only to itself (or a copy sharing its
    name)."""

    __slots__ = ("dim", "name", "tags")

    def __init__(self, dim: int, tags: str = "", name=None):
        self.dim = int(dim)
        self.name = name if name is not None else next(_uid)
        self.tags = tags

    def __eq__(self, other):
        return isinstance(other, Index) and other.name == self.name

    def __hash__(self):
        return hash(("Index", self.name))

    def __repr__(self):
        return f"Index(dim={self.dim}, name={self.name}{', ' + self.tags if self.tags else ''})"


def uniquename(ind: Index) -> Index:
    """A fresh index of the same dimension (ITensorBase `uniquename`)."""
    return Index(ind.dim, ind.tags)


class ITensor:
    """Dense array with one Index per axis."""

    def __init__(self, data, inds: Sequence[Index]):
        data = np.asarray(data)
        inds = tuple(inds)
        if data.ndim != len(inds) or any(d != i.dim for d, i in zip(data.shape, inds)):
            raise ValueError("tensor shape does not match its indices")
        self.data = data
        self.inds = inds

    def dimnames(self):
        return list(self.inds)

    def __getitem__(self, key):
        """Entries by position; `t[()]` is the value of a 0-dimensional tensor."""
        return self.data[key]

    def permute(self, inds: Sequence[Index]) -> np.ndarray:
        perm = [self.inds.index(i) for i in inds]
        return np.transpose(self.data, perm)

    def __repr__(self):
        return f"ITensor({self.data.shape}, {self.inds})"


def random_itensor(inds: Sequence[Index], rng=None, dtype=np.float64) -> ITensor:
    rng = rng if rng is not None else np.random.default_rng()
    return ITensor(rng.standard_normal(tuple(i.dim for i in inds)).astype(dtype), inds)


class ITensorNetwork:
    """vertex -> ITensor with edges inferred from shared index names.  Edges are added in
    the order the tensors and their indices arrive (never in hash order), so the same
    construction gives the same edge order, and with it the same default BP schedule."""

    def __init__(self, tensors: Dict[Hashable, ITensor] | None = None):
        self.tensors: Dict[Hashable, ITensor] = {}
        # index -> vertices holding it, in insertion order (a dict used as an ordered set)
        self.dimname_vertices: Dict[Index, dict] = {}
        self.graph = NamedGraph()
        for v, t in (tensors or {}).items():
            self[v] = t

    # -- upsert with edge inference (tensornetwork.jl:138-173)
    def __setitem__(self, v, t: ITensor):
        if not isinstance(t, ITensor):
            raise TypeError("ITensorNetwork holds ITensors")
        old = list(self.tensors[v].inds) if v in self.tensors else []
        new = list(t.inds)
        self.graph.add_vertex(v)
        for ind in old:
            if ind in new:
                continue
            vs = self.dimname_vertices.get(ind, {})
            vs.pop(v, None)
            if not vs:
                self.dimname_vertices.pop(ind, None)
            self._rebuild_edges()
        for ind in new:
            if ind in old:
                continue
            vs = self.dimname_vertices.setdefault(ind, {})
            if len(vs) > 1:
                raise ValueError(f"index {ind} can appear in at most one existing tensor, "
                                 f"got {len(vs)}.")
            vs[v] = None
            if len(vs) == 2:
                a, b = tuple(vs)
                if not self.graph.has_edge(a, b):
                    self.graph.add_edge(a, b)
        self.tensors[v] = t

    def _rebuild_edges(self):
        g = NamedGraph(vertices=list(self.tensors.keys()) or self.graph.vertices())
        for ind, vs in self.dimname_vertices.items():
            if len(vs) == 2:
                a, b = tuple(vs)
                if not g.has_edge(a, b):
                    g.add_edge(a, b)
        self.graph = g

    def __getitem__(self, v) -> ITensor:
        return self.tensors[v]

    def __contains__(self, v):
        return v in self.tensors

    def vertices(self) -> List[Hashable]:
        return self.graph.vertices()

    def edges(self):
        return self.graph.edges()

    def linkinds(self, edge) -> List[Index]:
        u, v = edge
        return [i for i in self.tensors[u].inds if i in set(self.tensors[v].inds)]

    def siteinds(self, v) -> List[Index]:
        return [i for i in self.tensors[v].inds if len(self.dimname_vertices.get(i, ())) == 1]

    def copy(self) -> "ITensorNetwork":
        return ITensorNetwork({v: ITensor(t.data.copy(), t.inds) for v, t in self.tensors.items()})


def tensornetwork(f, vertices: Iterable[Hashable]) -> ITensorNetwork:
    """tensornetwork(f, vs) (tensornetwork.jl:115): vertex v holds f(v)."""
    return ITensorNetwork({v: f(v) for v in vertices})


def linkinds(tn, edge):
    return (tn.ket if isinstance(tn, NormNetwork) else tn).linkinds(edge)


def siteinds(tn, v):
    return (tn.ket if isinstance(tn, NormNetwork) else tn).siteinds(v)


class NormNetwork:
    """Lazy <tn|tn>: the ket network plus a ket -> bra link-name map; site names map to
    themselves (normnetwork.jl:16-27, 66-73).  Only the ket tensors are stored."""

    def __init__(self, tn: ITensorNetwork, braname: Dict[Index, Index] | None = None):
        self.ket = tn
        if braname is None:
            braname = {i: uniquename(i) for i, vs in tn.dimname_vertices.items() if len(vs) == 2}
        self.braname = {i: braname[i] for i, vs in tn.dimname_vertices.items() if len(vs) == 2}

    def vertices(self):
        return self.ket.vertices()

    def edges(self):
        return self.ket.edges()

    @property
    def graph(self):
        return self.ket.graph

    def bra_index(self, ind: Index) -> Index:
        if ind not in self.ket.dimname_vertices:
            raise KeyError(f"index name {ind} not found underlying tensor network.")
        return self.braname.get(ind, ind)

    def kettensor(self, v) -> ITensor:
        return self.ket[v]

    def conj_bratensor(self, v) -> ITensor:
        t = self.ket[v]
        return ITensor(t.data, [self.bra_index(i) for i in t.inds])

    def bratensor(self, v) -> ITensor:
        t = self.conj_bratensor(v)
        return ITensor(np.conj(t.data), t.inds)


def normnetwork(tn: ITensorNetwork, braname=None) -> NormNetwork:
    return NormNetwork(tn, braname)


def braname(nn: NormNetwork, name: Index) -> Index:
    """The bra-layer name of `name`: its link partner, or itself for a site index
    (normnetwork.jl:74-81)."""
    return nn.bra_index(name)


def kettensor(nn: NormNetwork, v) -> ITensor:
    return nn.kettensor(v)


def bratensor(nn: NormNetwork, v) -> ITensor:
    return nn.bratensor(v)


def conj_bratensor(nn: NormNetwork, v) -> ITensor:
    return nn.conj_bratensor(v)


class _LayerView:
    """One layer of a NormNetwork seen as a network of its own (normnetworkview.jl:6-58):
    the parent's graph, vertices and edges; vertex data is the layer's tensor; settable
    and insertable only if the parent is (a NormNetwork is neither)."""

    def __init__(self, parent: NormNetwork):
        self.parent = parent

    def vertices(self):
        return self.parent.vertices()

    def edges(self):
        return self.parent.edges()

    @property
    def graph(self):
        return self.parent.graph

    def __getitem__(self, v) -> ITensor:
        return self.get_vertex_data(v)

    def get_vertex_data(self, v) -> ITensor:  # pragma: no cover - interface
        raise NotImplementedError

    def is_vertex_assigned(self, v) -> bool:
        return v in self.parent.ket

    def issettable(self) -> bool:
        return False

    def isinsertable(self) -> bool:
        return False

    def linkinds(self, edge) -> List[Index]:
        """The layer's link indices shared by the edge's two tensors."""
        u, w = edge
        mine = self[u].inds
        other = set(self[w].inds)
        return [i for i in mine if i in other]

    def linknames(self, edge) -> List[Index]:
        return self.linkinds(edge)


class KetView(_LayerView):
    """The ket layer: the bare ket tensors (`kettensor`)."""

    def get_vertex_data(self, v) -> ITensor:
        return self.parent.kettensor(v)


class BraView(_LayerView):
    """The bra layer: conj of the ket tensors with the bra link names (`bratensor`)."""

    def get_vertex_data(self, v) -> ITensor:
        return self.parent.bratensor(v)


def is_vertex_assigned(tn, v) -> bool:
    if isinstance(tn, (KetView, BraView)):
        return tn.is_vertex_assigned(v)
    return v in (tn.ket if isinstance(tn, NormNetwork) else tn)


def issettable(tn) -> bool:
    return tn.issettable() if isinstance(tn, (KetView, BraView)) else not isinstance(tn, NormNetwork)


def isinsertable(tn) -> bool:
    return tn.isinsertable() if isinstance(tn, (KetView, BraView)) else not isinstance(tn, NormNetwork)


def linknames(tn, edge) -> List[Index]:
    if isinstance(tn, (KetView, BraView)):
        return tn.linknames(edge)
    return linkinds(tn, edge)


# ------------------------------------------------------------------ flattening

class Flattened:
    """A network flattened to the engine's topology, with the maps back to names."""

    def __init__(self, topo: Topology, sites: np.ndarray, port_inds: Dict[Tuple, List[Index]],
                 plain: bool):
        self.topo = topo
        self.sites = sites
        self.port_inds = port_inds  # (u, v) -> shared link indices, in the order they are fused
        self.plain = plain

    def slot(self, edge) -> int:
        return self.topo.slot(edge[0], edge[1])


def to_topology(factors) -> Flattened:
    """Flatten an ITensorNetwork (plain BP: vector messages) or a NormNetwork (double layer:
    matrix messages) to the engine's port topology.  Several indices shared by the same
    pair of tensors are fused into one link; several site indices into one site index."""
    plain = not isinstance(factors, NormNetwork)
    tn = factors if plain else factors.ket
    g = tn.graph
    vs = g.vertices()
    port_inds = {}
    dims = {}
    for (u, v) in g.edges():
        li = tn.linkinds((u, v))
        port_inds[(u, v)] = li
        port_inds[(v, u)] = li
        dims[(u, v)] = int(np.prod([i.dim for i in li]))
    site_dims = {}
    for v in vs:
        si = tn.siteinds(v)
        site_dims[v] = int(np.prod([i.dim for i in si])) if si else 1
        if plain and si:
            raise ValueError("plain BP needs every index shared by two tensors (vertex "
                             f"{v} has unshared indices {si})")
    topo = build_topology(g, dims, site_dims, plain=plain)
    arrs = []
    for ui, u in enumerate(topo.vertices):
        t = tn[u]
        order = list(tn.siteinds(u))
        for q in range(topo.port_off[ui], topo.port_off[ui + 1]):
            w = topo.vertices[topo.port_nbr[q]]
            order += port_inds[(u, w)]
        arrs.append(t.permute(order))
    # complex tensors (ComplexF32 promotes, as Julia's scalars do) give a complex128 site
    # buffer for a complex context; real ones float64
    ct = np.complex128 if any(np.iscomplexobj(a) for a in arrs) else np.float64
    chunks = [np.ascontiguousarray(a, dtype=ct).reshape(-1) for a in arrs]
    sites = np.concatenate(chunks) if chunks else np.zeros(0, dtype=ct)
    return Flattened(topo, sites, port_inds, plain)
