"""Graphs and the flat port/CSR topology the engine runs on.

Vertex names follow NamedGraphs' conventions (tuples of 1-based coordinates for grids),
so host code reads like the reference's tests (`named_grid`, `named_comb_tree`,
`named_path_graph`).  The device sees only the flat topology built by `Topology`:

* each vertex ``u`` owns ports ``port_off[u] .. port_off[u+1]-1`` (its incident edges, in
  the order of ``incident_edges``), port ``q`` points to ``port_nbr[q]`` over undirected
  edge ``port_edge[q]`` and ``port_rev[q]`` is the neighbour's port pointing back;
* the directed message ``u -> port_nbr[q]`` lives in message slot ``q`` (the reference
  stores one tensor per directed edge, src/beliefpropagation/messagecache.jl:15-23);
* the site tensor of ``u`` is ``[phys][D_port0]...[D_port(k-1)]`` (site index first, links
  in incident-edge order, as test/test_normnetwork.jl:22-31 builds PEPS tensors).
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field
from typing import Dict, Hashable, List, Sequence, Tuple

import numpy as np

Vertex = Hashable
Edge = Tuple[Vertex, Vertex]


class NamedGraph:
    """A simple undirected graph over hashable vertex names, with stable orders."""

    def __init__(self, vertices: Sequence[Vertex] = (), edges: Sequence[Edge] = ()):
        self._vertices: List[Vertex] = []
        self._index: Dict[Vertex, int] = {}
        self._adj: Dict[Vertex, List[Vertex]] = {}
        self._edges: List[Edge] = []
        self._edge_set = set()
        for v in vertices:
            self.add_vertex(v)
        for u, v in edges:
            self.add_edge(u, v)

    # -- construction
    def add_vertex(self, v: Vertex) -> None:
        if v in self._index:
            return
        self._index[v] = len(self._vertices)
        self._vertices.append(v)
        self._adj[v] = []

    def add_edge(self, u: Vertex, v: Vertex) -> None:
        if u == v:
            raise ValueError("self-loops are not supported")
        self.add_vertex(u)
        self.add_vertex(v)
        key = frozenset((u, v))
        if key in self._edge_set:
            raise ValueError(f"multi-edge {u} -- {v} is not supported")
        self._edge_set.add(key)
        self._edges.append((u, v))
        self._adj[u].append(v)
        self._adj[v].append(u)

    # -- queries
    def vertices(self) -> List[Vertex]:
        return list(self._vertices)

    def edges(self) -> List[Edge]:
        return list(self._edges)

    def has_edge(self, u: Vertex, v: Vertex) -> bool:
        return frozenset((u, v)) in self._edge_set

    def neighbors(self, v: Vertex) -> List[Vertex]:
        return sorted(self._adj[v], key=self._index.__getitem__)

    def incident_edges(self, v: Vertex) -> List[Edge]:
        return [(v, w) for w in self.neighbors(v)]

    def degree(self, v: Vertex) -> int:
        return len(self._adj[v])

    def nv(self) -> int:
        return len(self._vertices)

    def ne(self) -> int:
        return len(self._edges)

    def vertex_index(self, v: Vertex) -> int:
        return self._index[v]

    def all_edges(self) -> List[Edge]:
        """Both orientations of every edge (NamedGraphs `all_edges`)."""
        out = []
        for u, v in self._edges:
            out.append((u, v))
            out.append((v, u))
        return out

    def is_tree(self) -> bool:
        return self.ne() == self.nv() - 1 and _connected(self)

    def __repr__(self) -> str:
        return f"NamedGraph(nv={self.nv()}, ne={self.ne()})"


def _connected(g: NamedGraph) -> bool:
    vs = g.vertices()
    if not vs:
        return True
    seen = {vs[0]}
    stack = [vs[0]]
    while stack:
        x = stack.pop()
        for y in g.neighbors(x):
            if y not in seen:
                seen.add(y)
                stack.append(y)
    return len(seen) == len(vs)


# ------------------------------------------------------------------ generators

def named_grid(dims: Sequence[int], periodic: bool = False) -> NamedGraph:
    """Open (or periodic) grid with vertices named by 1-based coordinate tuples.

    Vertex order is column-major like Julia's `CartesianIndices` (first coordinate
    fastest), as NamedGraphs' `named_grid` produces.
    """
    dims = tuple(int(d) for d in dims)
    g = NamedGraph()
    ranges = [range(1, d + 1) for d in dims]
    coords = [tuple(reversed(c)) for c in itertools.product(*reversed(ranges))]
    for c in coords:
        g.add_vertex(c)
    for c in coords:
        for ax, d in enumerate(dims):
            nxt = list(c)
            if c[ax] < d:
                nxt[ax] = c[ax] + 1
                g.add_edge(c, tuple(nxt))
            elif periodic and d > 2:
                nxt[ax] = 1
                g.add_edge(c, tuple(nxt))
    return g


def named_path_graph(n: int) -> NamedGraph:
    g = NamedGraph(vertices=list(range(1, n + 1)))
    for i in range(1, n):
        g.add_edge(i, i + 1)
    return g


def named_comb_tree(dims: Tuple[int, int]) -> NamedGraph:
    """Comb tree: a backbone along the first coordinate at x2 = 1, with teeth along x2."""
    n1, n2 = dims
    g = NamedGraph()
    for j in range(1, n2 + 1):
        for i in range(1, n1 + 1):
            g.add_vertex((i, j))
    for i in range(1, n1):
        g.add_edge((i, 1), (i + 1, 1))
    for i in range(1, n1 + 1):
        for j in range(1, n2):
            g.add_edge((i, j), (i, j + 1))
    return g


def heavy_hex_lattice(nrows: int, ncols: int) -> NamedGraph:
    """Heavy-hexagon lattice: a hexagonal (brick-wall) lattice of nrows x ncols qubit
    sites with an extra degree-2 vertex on every edge.  Degrees are 2 and 3."""
    base = NamedGraph()
    for r in range(1, nrows + 1):
        for c in range(1, ncols + 1):
            base.add_vertex(("q", r, c))
    for r in range(1, nrows + 1):
        for c in range(1, ncols):
            base.add_edge(("q", r, c), ("q", r, c + 1))
    for r in range(1, nrows):
        for c in range(1, ncols + 1):
            # brick-wall vertical links alternate by row parity
            if (r + c) % 2 == 0:
                base.add_edge(("q", r, c), ("q", r + 1, c))
    g = NamedGraph()
    for v in base.vertices():
        g.add_vertex(v)
    for k, (u, v) in enumerate(base.edges()):
        mid = ("b",) + u[1:] + v[1:]
        g.add_edge(u, mid)
        g.add_edge(mid, v)
    return g


# ------------------------------------------------------------------ topology


@dataclass
class Topology:
    """Flat CSR topology plus tensor/message offsets for one network (or a batch)."""

    vertices: List[Vertex]
    edges: List[Edge]
    port_off: np.ndarray
    port_nbr: np.ndarray
    port_edge: np.ndarray
    port_rev: np.ndarray
    edge_dim: np.ndarray
    phys_dim: np.ndarray
    site_off: np.ndarray
    msg_off: np.ndarray
    plain: bool = False
    vertex_net: np.ndarray = field(default=None)  # network id per vertex (batches)
    nnet: int = 1

    @property
    def nv(self) -> int:
        return len(self.vertices)

    @property
    def ne(self) -> int:
        return len(self.edges)

    @property
    def nports(self) -> int:
        return int(self.port_off[-1])

    def degree(self, u: int) -> int:
        return int(self.port_off[u + 1] - self.port_off[u])

    def port_owner(self) -> np.ndarray:
        own = np.empty(self.nports, dtype=np.int32)
        for u in range(self.nv):
            own[self.port_off[u]:self.port_off[u + 1]] = u
        return own

    def site_shape(self, u: int) -> Tuple[int, ...]:
        ports = range(self.port_off[u], self.port_off[u + 1])
        return (int(self.phys_dim[u]),) + tuple(int(self.edge_dim[self.port_edge[q]]) for q in ports)

    def msg_shape(self, q: int) -> Tuple[int, ...]:
        D = int(self.edge_dim[self.port_edge[q]])
        return (D,) if self.plain else (D, D)

    def slot(self, u: Vertex, v: Vertex) -> int:
        """Message slot of the directed edge u -> v."""
        ui = self._vidx[u]
        vi = self._vidx[v]
        for q in range(self.port_off[ui], self.port_off[ui + 1]):
            if self.port_nbr[q] == vi:
                return q
        raise KeyError(f"no edge {u} -> {v}")

    def directed_edge(self, q: int) -> Edge:
        own = self.port_owner()
        return (self.vertices[own[q]], self.vertices[self.port_nbr[q]])

    def __post_init__(self):
        self._vidx = {v: i for i, v in enumerate(self.vertices)}
        if self.vertex_net is None:
            self.vertex_net = np.zeros(len(self.vertices), dtype=np.int32)

    def uniform_bond_dim(self) -> int:
        """The common bond dimension, or 0 when edges differ."""
        if self.ne == 0:
            return 0
        d = np.unique(self.edge_dim)
        return int(d[0]) if len(d) == 1 else 0


def build_topology(g: NamedGraph, bond_dim, phys_dim=2, plain: bool = False,
                   port_order=None) -> Topology:
    """Flatten a graph.  `bond_dim` is an int or a dict edge -> dim; `phys_dim` an int or a
    dict vertex -> dim.  `port_order(v)` may give the neighbour order of each vertex."""
    vs = g.vertices()
    vidx = {v: i for i, v in enumerate(vs)}
    es = g.edges()
    eidx = {}
    for k, (u, v) in enumerate(es):
        eidx[frozenset((u, v))] = k
    edge_dim = np.empty(len(es), dtype=np.int32)
    for k, (u, v) in enumerate(es):
        if isinstance(bond_dim, dict):
            D = bond_dim.get((u, v), bond_dim.get((v, u)))
        else:
            D = bond_dim
        edge_dim[k] = int(D)
    pdim = np.empty(len(vs), dtype=np.int32)
    for i, v in enumerate(vs):
        pdim[i] = 1 if plain else int(phys_dim[v] if isinstance(phys_dim, dict) else phys_dim)
    port_off = np.zeros(len(vs) + 1, dtype=np.int32)
    nbrs = []
    for i, v in enumerate(vs):
        nb = list(port_order(v)) if port_order is not None else g.neighbors(v)
        nbrs.append(nb)
        port_off[i + 1] = port_off[i] + len(nb)
    nports = int(port_off[-1])
    port_nbr = np.empty(nports, dtype=np.int32)
    port_edge = np.empty(nports, dtype=np.int32)
    for i, v in enumerate(vs):
        for j, w in enumerate(nbrs[i]):
            q = port_off[i] + j
            port_nbr[q] = vidx[w]
            port_edge[q] = eidx[frozenset((v, w))]
    port_rev = np.empty(nports, dtype=np.int32)
    for i in range(len(vs)):
        for q in range(port_off[i], port_off[i + 1]):
            w = port_nbr[q]
            for r in range(port_off[w], port_off[w + 1]):
                if port_nbr[r] == i:
                    port_rev[q] = r
                    break
    site_off = np.zeros(len(vs) + 1, dtype=np.int64)
    for i in range(len(vs)):
        n = int(pdim[i])
        for q in range(port_off[i], port_off[i + 1]):
            n *= int(edge_dim[port_edge[q]])
        site_off[i + 1] = site_off[i] + n
    msg_off = np.zeros(nports + 1, dtype=np.int64)
    for q in range(nports):
        D = int(edge_dim[port_edge[q]])
        msg_off[q + 1] = msg_off[q] + (D if plain else D * D)
    return Topology(vs, es, port_off, port_nbr, port_edge, port_rev, edge_dim, pdim, site_off,
                    msg_off, plain=plain)


def concat_topologies(tops: Sequence[Topology]) -> Topology:
    """Disjoint union of several networks (a batch); vertex_net records each vertex's
    network.  Vertex names become (network index, name)."""
    vs, es = [], []
    po, pn, pe, pr, ed, pd, so, mo, vn = [np.zeros(1, np.int32)], [], [], [], [], [], \
        [np.zeros(1, np.int64)], [np.zeros(1, np.int64)], []
    voff = eoff = poff = 0
    soff = moff = 0
    plain = tops[0].plain if tops else False
    for n, t in enumerate(tops):
        if t.plain != plain:
            raise ValueError("cannot batch plain and norm networks")
        vs += [(n, v) for v in t.vertices]
        es += [((n, a), (n, b)) for a, b in t.edges]
        po.append(t.port_off[1:] + poff)
        pn.append(t.port_nbr + voff)
        pe.append(t.port_edge + eoff)
        pr.append(t.port_rev + poff)
        ed.append(t.edge_dim)
        pd.append(t.phys_dim)
        so.append(t.site_off[1:] + soff)
        mo.append(t.msg_off[1:] + moff)
        vn.append(np.full(t.nv, n, dtype=np.int32))
        voff += t.nv
        eoff += t.ne
        poff += t.nports
        soff += int(t.site_off[-1])
        moff += int(t.msg_off[-1])
    cat = lambda xs, dt: np.concatenate(xs).astype(dt) if xs else np.zeros(0, dt)
    return Topology(vs, es, cat(po, np.int32), cat(pn, np.int32), cat(pe, np.int32),
                    cat(pr, np.int32), cat(ed, np.int32), cat(pd, np.int32),
                    cat(so, np.int64), cat(mo, np.int64), plain=plain,
                    vertex_net=cat(vn, np.int32), nnet=len(tops))


# ------------------------------------------------------------------ schedules

def forest_cover_edge_sequence(g: NamedGraph) -> List[Edge]:
    """Sequential BP schedule in the manner of NamedGraphs' `forest_cover_edge_sequence`
    (used as the default by src/beliefpropagation/beliefpropagation.jl:14): the edges are
    split into spanning forests; for each tree of each forest, edges are listed in
    post-order towards a root, followed by their reverses (root back to leaves).  Every
    directed edge appears exactly once; on a tree one pass is exact."""
    remaining = list(g.edges())
    seq: List[Edge] = []
    while remaining:
        # greedy spanning forest over the remaining edges (union-find)
        parent = {}

        def find(x):
            while parent.get(x, x) != x:
                parent[x] = parent.get(parent[x], parent[x])
                x = parent[x]
            return x

        forest, rest = [], []
        for u, v in remaining:
            ru, rv = find(u), find(v)
            if ru != rv:
                parent[ru] = rv
                forest.append((u, v))
            else:
                rest.append((u, v))
        remaining = rest
        fg = NamedGraph(edges=forest)
        seen = set()
        for root in fg.vertices():
            if root in seen:
                continue
            # iterative DFS post-order of tree edges toward root
            order = []
            stack = [(root, None, iter(fg.neighbors(root)))]
            seen.add(root)
            while stack:
                node, par, it = stack[-1]
                nxt = next(it, None)
                if nxt is None:
                    stack.pop()
                    if par is not None:
                        order.append((node, par))
                elif nxt != par and nxt not in seen:
                    seen.add(nxt)
                    stack.append((nxt, node, iter(fg.neighbors(nxt))))
            seq += order
            seq += [(b, a) for a, b in reversed(order)]
    return seq
