"""MessageCache and the BP scalars, evaluated on the GPU.

Mirrors src/beliefpropagation/messagecache.jl: a directed-edge -> message store whose
graph is built from its keys (:15-49), incoming_messages / incoming_edge_data (:124-137),
vertex_scalar(s) (:139-151), edge_scalar(s) (:153-178), region_scalar (:180-182),
bethe_free_energy (:185-201) and message_environment (:205-229).

Message layout: a norm-network message on edge {u, v} is a (D, D) array indexed
[ket link][bra link] (D the fused dimension of the indices u and v share); a plain-network
message is a length-D vector.  Every contraction below runs in the HIP engine.
"""
from __future__ import annotations

import math
from typing import Dict, Hashable, Iterable, List, Tuple

import numpy as np

from .graphs import NamedGraph, build_topology
from .itensors import NormNetwork, to_topology

Edge = Tuple[Hashable, Hashable]


def _edge(e) -> Edge:
    if isinstance(e, tuple) and len(e) == 2:
        return e
    raise TypeError(f"expected a directed edge (src, dst), got {e!r}")


class MessageCache:
    """Directed edge -> message, over a directed graph of its own (messagecache.jl:15-49).

    The graph starts from the vertices given and the keys of the messages (every key is an
    edge); setting a message on a missing edge adds the edge (`set_edge_data!`, :92-96),
    `rem_edge_` drops both (:54-58).  Numeric messages are stored as float64 arrays
    (complex128 when complex); any other value (the reference's tests store strings) is
    kept as it is."""

    def __init__(self, messages=None, vertices=None):
        self.messages: Dict[Edge, object] = {}
        self._vertices: Dict[Hashable, None] = {}
        self._edges: Dict[Edge, None] = {}
        self.state = None  # stopping-criterion state of the run that produced this cache
        for v in (vertices or ()):
            self._vertices.setdefault(v, None)
        if messages is not None:
            items = messages.items() if hasattr(messages, "items") else messages
            for e, m in items:
                self[e] = m

    # -- graph interface
    def _add_edge(self, e: Edge) -> None:
        self._vertices.setdefault(e[0], None)
        self._vertices.setdefault(e[1], None)
        self._edges.setdefault(e, None)

    def add_vertex(self, v) -> "MessageCache":
        self._vertices.setdefault(v, None)
        return self

    def add_edge(self, e) -> "MessageCache":
        self._add_edge(_edge(e))
        return self

    def rem_edge(self, e) -> "MessageCache":
        e = _edge(e)
        self.messages.pop(e, None)
        self._edges.pop(e, None)
        return self

    def edges(self) -> List[Edge]:
        return list(self._edges)

    def vertices(self) -> List[Hashable]:
        return list(self._vertices)

    def has_edge(self, e) -> bool:
        return _edge(e) in self._edges

    def has_vertex(self, v) -> bool:
        return v in self._vertices

    # -- dictionary interface
    def __getitem__(self, e):
        return self.messages[_edge(e)]

    def __setitem__(self, e, m) -> None:
        e = _edge(e)
        self._add_edge(e)
        self.messages[e] = _value(m)

    def __contains__(self, e) -> bool:
        return _edge(e) in self.messages

    def __len__(self) -> int:
        return len(self.messages)

    def is_edge_assigned(self, e) -> bool:
        return _edge(e) in self.messages

    def keys(self):
        return self.messages.keys()

    def values(self):
        return self.messages.values()

    def items(self):
        return self.messages.items()

    def copy(self) -> "MessageCache":
        c = MessageCache(vertices=self.vertices())
        for e in self._edges:
            c._add_edge(e)
        for e, m in self.messages.items():
            c.messages[e] = m.copy() if isinstance(m, np.ndarray) else m
        c.state = self.state
        return c

    def map(self, f) -> "MessageCache":
        """map(f, cache): same graph, f applied to every message (messagecache.jl:109-114)."""
        c = MessageCache(vertices=self.vertices())
        for e in self._edges:
            c._add_edge(e)
        for e, m in self.messages.items():
            c[e] = f(m)
        return c

    def __repr__(self):
        return f"MessageCache({len(self)} directed edges)"

    def is_complex(self) -> bool:
        """Whether any message is complex (the cache's element type is then complex)."""
        return any(np.iscomplexobj(m) for m in self.messages.values())

    # -- flat buffers in a topology's slot order
    def to_flat(self, topo) -> np.ndarray:
        out = np.zeros(int(topo.msg_off[-1]),
                       dtype=np.complex128 if self.is_complex() else np.float64)
        own = topo.port_owner()
        for q in range(topo.nports):
            e = (topo.vertices[own[q]], topo.vertices[topo.port_nbr[q]])
            if e not in self.messages:
                raise KeyError(f"no message for directed edge {e}")
            m = self.messages[e]
            if m.size != topo.msg_off[q + 1] - topo.msg_off[q]:
                raise ValueError(f"message {e} has {m.size} entries, expected "
                                 f"{topo.msg_off[q + 1] - topo.msg_off[q]}")
            out[topo.msg_off[q]:topo.msg_off[q + 1]] = m.reshape(-1)
        return out

    @classmethod
    def from_flat(cls, topo, flat) -> "MessageCache":
        own = topo.port_owner()
        c = cls(vertices=list(topo.vertices))
        for q in range(topo.nports):
            e = (topo.vertices[own[q]], topo.vertices[topo.port_nbr[q]])
            c[e] = np.array(flat[topo.msg_off[q]:topo.msg_off[q + 1]]).reshape(topo.msg_shape(q))
        return c


def _value(m):
    if isinstance(m, (np.ndarray, list, tuple, float, int, complex, np.number)):
        return np.asarray(m, dtype=np.complex128 if np.iscomplexobj(m) else np.float64)
    return m


def _scalar(x):
    """A Python scalar of the element type: complex for complex networks."""
    return complex(x) if np.iscomplexobj(x) else float(x)


def context_dtype(dtype: str, sites, *caches) -> str:
    """The engine dtype for `dtype` requested on these sites and messages: complex data
    (complex site tensors or messages) runs in complex FP64 ("c128"), as Julia promotes the
    arithmetic to the complex element type."""
    if np.iscomplexobj(sites) or any(c is not None and c.is_complex() for c in caches):
        return "c128"
    return dtype


def messagecache(f_or_pairs, edges=None) -> MessageCache:
    """messagecache(pairs) and messagecache(f, edges) = messagecache(e => f(e) for e in
    edges) (messagecache.jl:51-52)."""
    if edges is None:
        pairs = f_or_pairs
        return MessageCache(dict(pairs.items() if hasattr(pairs, "items") else pairs))
    return MessageCache({_edge(e): f_or_pairs(_edge(e)) for e in edges})


def has_edge(cache: MessageCache, e) -> bool:
    return cache.has_edge(e)


def rem_edge_(cache: MessageCache, e) -> MessageCache:
    """Graphs.rem_edge!(cache, edge): drops the message and the edge (messagecache.jl:54-58)."""
    return cache.rem_edge(e)


def add_vertex_(cache: MessageCache, v) -> MessageCache:
    return cache.add_vertex(v)


def set_edge_data_(cache: MessageCache, val, e) -> MessageCache:
    """set_edge_data!: adds the edge when missing (messagecache.jl:92-96)."""
    cache[e] = val
    return cache


def insert_edge_data_(cache: MessageCache, e, val) -> MessageCache:
    """insert_edge_data!: a new key only (Dictionaries' insert! fails on an existing one)."""
    if cache.is_edge_assigned(e):
        raise KeyError(f"edge {e!r} already holds a message")
    cache[e] = val
    return cache


def edge_data(cache: MessageCache) -> Dict[Edge, object]:
    """The edge -> message dictionary."""
    return cache.messages


def edge_data_type(cache: MessageCache) -> type:
    """The common type of the stored messages (valtype / edge_data_type)."""
    ts = {type(m) for m in cache.values()}
    return ts.pop() if len(ts) == 1 else object


valtype = edge_data_type


def copyto_(dst: MessageCache, src, edges=None) -> MessageCache:
    """copyto!(dst, src) and copyto!(dst, src, edges): the messages of `src` (any edge ->
    message mapping), or only those on `edges`, written into `dst` (messagecache.jl:42-49;
    test_beliefpropagation.jl:61-79)."""
    items = src.items() if hasattr(src, "items") else src
    if edges is None:
        for e, m in items:
            dst[e] = m
        return dst
    for e in edges:
        dst[e] = src[_edge(e)]
    return dst


def map_(f, dst: MessageCache, src) -> MessageCache:
    """map!(f, dst, src): dst[e] = f(src[e]) for every key of src (messagecache.jl:116-120)."""
    for e in list(src.keys()):
        dst[e] = f(src[e])
    return dst


def subgraph(cache: MessageCache, vertices) -> MessageCache:
    """The cache induced on `vertices`: those vertices, the edges between them and their
    messages (DataGraphs' subgraph of an edge-data graph; test_beliefpropagation.jl:117-132)."""
    keep = list(dict.fromkeys(vertices))
    ks = set(keep)
    sub = MessageCache(vertices=keep)
    for e in cache.edges():
        if e[0] in ks and e[1] in ks:
            sub._add_edge(e)
            if cache.is_edge_assigned(e):
                m = cache[e]
                sub.messages[e] = m.copy() if isinstance(m, np.ndarray) else m
    return sub


def incoming_messages(cache: MessageCache, edge) -> List[np.ndarray]:
    """In-edges of src(edge) except reverse(edge) (messagecache.jl:124-131)."""
    u, v = _edge(edge)
    return [m for (a, b), m in cache.items() if b == u and a != v]


def incoming_edge_data(cache: MessageCache, vertices: Iterable[Hashable]) -> List[np.ndarray]:
    """Messages on the in-boundary of a vertex set (messagecache.jl:134-137)."""
    vs = set(vertices)
    return [m for (a, b), m in cache.items() if b in vs and a not in vs]


# ------------------------------------------------------------------ device evaluation

def _flat(factors):
    cached = getattr(factors, "_engine_flat", None)
    if cached is None:
        cached = to_topology(factors)
        try:
            factors._engine_flat = cached
        except AttributeError:
            pass
    return cached


def _engine(factors, messages, dtype="f64"):
    from .engine import Engine

    fl = _flat(factors)
    mc = messages if isinstance(messages, MessageCache) else MessageCache(messages)
    dtype = context_dtype(dtype, fl.sites, mc)
    e = Engine(fl.topo, fl.sites, dtype=dtype, path="auto" if dtype in ("f32", "bf16") else "generic")
    e.set_messages(mc.to_flat(fl.topo))
    return fl, e


def _cache_engine(cache: MessageCache, *others: MessageCache):
    """An engine over the graph of a cache alone (no factors): zero site tensors, used for
    the message-only quantities (iterate_diff, edge scalars); complex when any of the
    caches is."""
    from .engine import Engine

    vs = cache.vertices()
    g = NamedGraph(vertices=vs)
    dims = {}
    plain = None
    for (u, v), m in cache.items():
        if not g.has_edge(u, v):
            g.add_edge(u, v)
        is_plain = m.ndim == 1
        plain = is_plain if plain is None else plain
        if plain != is_plain:
            raise ValueError("mixed vector and matrix messages")
        dims[(u, v)] = m.shape[0]
    topo = build_topology(g, dims, 1, plain=bool(plain))
    dt = context_dtype("f64", None, cache, *others)
    e = Engine(topo, np.zeros(int(topo.site_off[-1])), dtype=dt, path="generic")
    return topo, e


def vertex_scalars(factors, messages, vertices=None) -> np.ndarray:
    """Z_v for each vertex: the full contraction of factor v with its incoming messages."""
    fl, e = _engine(factors, messages)
    try:
        z = e.vertex_scalars()
    finally:
        e.close()
    if vertices is None:
        return z
    idx = {v: i for i, v in enumerate(fl.topo.vertices)}
    return np.array([z[idx[v]] for v in vertices])


def vertex_scalar(factors, messages, vertex):
    """Z_v (messagecache.jl:139-143): a float, or a complex for complex networks."""
    return _scalar(vertex_scalars(factors, messages, [vertex])[0])


def edge_scalars(cache: MessageCache, edges=None) -> np.ndarray:
    """One Z_e = <m_uv, m_vu> per undirected edge, skipping repeats and reverses
    (messagecache.jl:159-178), in first-appearance order of `edges` (default: the keys)."""
    topo, e = _cache_engine(cache)
    try:
        e.set_messages(cache.to_flat(topo))
        ze = e.edge_scalars()
    finally:
        e.close()
    eidx = {frozenset(ed): k for k, ed in enumerate(topo.edges)}
    out, seen = [], set()
    for ed in (edges if edges is not None else cache.edges()):
        key = frozenset(ed)
        if key in seen:
            continue
        seen.add(key)
        out.append(ze[eidx[key]])
    return np.array(out)


def edge_scalar(cache: MessageCache, edge):
    return _scalar(edge_scalars(cache, [edge])[0])


def region_scalar(factors, messages, region):
    """The product of the vertex scalars of the region (messagecache.jl:180-182)."""
    return _scalar(np.prod(vertex_scalars(factors, messages, list(region))))


def bethe_free_energy(factors, messages):
    """sum log Z_v - sum log Z_e, complex when a term is negative, -Inf when an edge term
    vanishes (messagecache.jl:185-201); compensated FP64 sums on the device."""
    fl, e = _engine(factors, messages)
    try:
        v = e.lognorm()
    finally:
        e.close()
    return v


def expect(factors, messages, ops) -> dict:
    """BP expectation values <O_v>_BP = Z_v[O_v] / Z_v of one-site operators.

    `ops` maps vertices to real d x d matrices acting on the physical index of the ket
    layer; returns {vertex: value} for the vertices in `ops`.  The region scalar of
    messagecache.jl:139-143 with an operator inserted (the reference itself has no
    expectation API); contractions on the device in FP64."""
    fl, e = _engine(factors, messages)
    try:
        t = fl.topo
        ld = max(int(np.max(t.phys_dim)), max((np.shape(o)[0] for o in ops.values()), default=1))
        blk = np.zeros((t.nv, ld, ld))
        for u in range(t.nv):
            blk[u] = np.eye(ld)
        index = {v: i for i, v in enumerate(t.vertices)}
        for v, o in ops.items():
            o = np.asarray(o, dtype=np.float64)
            d = int(t.phys_dim[index[v]])
            if o.shape != (d, d):
                raise ValueError(f"operator on {v!r} must be {d} x {d}")
            blk[index[v]] = 0.0
            blk[index[v], :d, :d] = o
        vals = e.expect_vertex(blk)
    finally:
        e.close()
    return {v: float(vals[index[v]]) for v in ops}


def expect2(factors, messages, ops) -> dict:
    """BP expectation values <O_u O_v>_BP of nearest-neighbour two-site operators: the
    two-site region {u, v} with its boundary messages (messagecache.jl:134-137),
    <g^{O_u}_{u->v}, g^{O_v}_{v->u}> / <g_{u->v}, g_{v->u}> with g^O the unnormalised
    message of the sender with O inserted.  `ops` maps edges (u, v) to pairs (O_u, O_v);
    returns {edge: value}.  FP64 on the device."""
    fl, e = _engine(factors, messages)
    try:
        t = fl.topo
        ld = int(np.max(t.phys_dim))
        a = np.broadcast_to(np.eye(ld), (t.ne, ld, ld)).copy()
        b = a.copy()
        index = {v: i for i, v in enumerate(t.vertices)}
        own = t.port_owner()
        first = {}
        for q in range(t.nports):
            if q < int(t.port_rev[q]):
                first[int(t.port_edge[q])] = int(own[q])
        eid = {}
        for q in range(t.nports):
            eid[(int(own[q]), int(t.port_nbr[q]))] = int(t.port_edge[q])
        keys = []
        for (u, v), (ou, ov) in ops.items():
            iu, iv = index[u], index[v]
            if (iu, iv) not in eid:
                raise ValueError(f"({u!r}, {v!r}) is not an edge")
            k = eid[(iu, iv)]
            ou, ov = np.asarray(ou, dtype=np.float64), np.asarray(ov, dtype=np.float64)
            if first[k] != iu:
                ou, ov = ov, ou
                iu, iv = iv, iu
            a[k] = 0.0
            b[k] = 0.0
            a[k, :ou.shape[0], :ou.shape[1]] = ou
            b[k, :ov.shape[0], :ov.shape[1]] = ov
            keys.append(((u, v), k))
        vals = e.expect_edge(a, b)
    finally:
        e.close()
    return {uv: float(vals[k]) for uv, k in keys}


def iterate_diff(cache1: MessageCache, cache2: MessageCache) -> float:
    """max over directed edges of 1 - |<m1^, m2^>|^2 (beliefpropagation.jl:261-267), the
    first argument conjugated for complex messages."""
    topo, e = _cache_engine(cache1, cache2)
    try:
        e.set_messages(cache1.to_flat(topo))
        return e.iterate_diff(cache2.to_flat(topo))
    finally:
        e.close()


def one(template: np.ndarray) -> np.ndarray:
    """`f = one` of message_environment: the identity message."""
    if template.ndim == 1:
        return np.ones_like(template)
    return np.eye(template.shape[0], dtype=template.dtype)


def similar_message_environment(nn) -> MessageCache:
    """A zero message for every in-edge of every vertex, of the site tensors' element type
    (messagecache.jl:205-225)."""
    fl = _flat(nn)
    topo = fl.topo
    own = topo.port_owner()
    c = MessageCache(vertices=list(topo.vertices))
    dt = np.complex128 if np.iscomplexobj(fl.sites) else np.float64
    for q in range(topo.nports):
        e = (topo.vertices[own[q]], topo.vertices[topo.port_nbr[q]])
        c[e] = np.zeros(topo.msg_shape(q), dtype=dt)
    return c


def message_environment(f, nn) -> MessageCache:
    """map(f, similar_message_environment(nn)) (messagecache.jl:227-229)."""
    return similar_message_environment(nn).map(f)
