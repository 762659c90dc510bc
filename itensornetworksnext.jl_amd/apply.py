"""BP-gauged gate application on the device: apply_operator / apply_operators with the
BPApplyGate strategy (src/apply/apply_operators.jl:1-283).

    state, env = apply_operator(gate, state, env; alg=None, trunc=None, normalize=False)
    state, env = apply_operators(gates, state, env; ...)

`state` is an ITensorNetwork (the ket), `env` a MessageCache of converged norm-network
messages ([ket][bra] per directed edge, as `beliefpropagation` returns).  A gate is an
Operator (`operator(data, codomain, domain)`): its domain indices are site indices of one
vertex or of two neighbouring vertices.

  * one site: the operator acts on the vertex tensor; with `normalize`, the tensor is divided
    by its norm with the square roots of its incoming messages absorbed (:214-227);
  * two sites (:229-283): the tensors absorb the square roots of their other incoming
    messages, are factored Q R, the gate acts on R1 R2, an SVD splits it back (truncated by
    `trunc`: truncrank(k) and/or trunctol(rtol=...); normalised when `normalize`), the
    inverse square roots are absorbed back, the bond becomes one new index of the kept rank,
    and the messages on that bond become diag(S) both ways.

Every contraction and factorisation runs in libbpmi355x.so (bp_apply_gates, one workgroup
per gate); `apply_operators` sends each run of gates on disjoint vertex pairs as one batch
(such gates commute exactly, so the result is the sequential one).  Real FP64 tensors;
every vertex's site indices are acted on together (a gate's domain holds all of them).
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Hashable, List, Sequence, Tuple

import numpy as np

from . import _lib as L
from .algorithms import AbstractAlgorithm, ArgumentError, register_default, select_algorithm
from .contract import _context
from .itensors import Index, ITensor, ITensorNetwork
from .messagecache import MessageCache

KMAX = 8


class Operator(ITensor):
    """An operator tensor (ITensorBase `operator(data, codomain, domain)`): `data` is indexed
    [codomain..., domain...]; applying it contracts the domain with a state's indices of the
    same names and gives the codomain indices the domain's names."""

    def __init__(self, data, codomain: Sequence[Index], domain: Sequence[Index]):
        codomain, domain = tuple(codomain), tuple(domain)
        if len(codomain) != len(domain) or any(a.dim != b.dim for a, b in zip(codomain, domain)):
            raise ValueError("codomain and domain must pair indices of equal dimensions")
        super().__init__(data, list(codomain) + list(domain))
        self.codomain = codomain
        self.domain = domain

    def matrix(self, domain: Sequence[Index]) -> np.ndarray:
        """The operator as a matrix [out][in] over `domain` in the given order."""
        pos = [self.domain.index(i) for i in domain]
        n = len(self.domain)
        perm = pos + [n + p for p in pos]
        dim = int(np.prod([i.dim for i in domain]))
        return np.transpose(self.data, perm).reshape(dim, dim)


def operator(data, codomain, domain) -> Operator:
    return Operator(data, codomain, domain)


class TruncRank:
    """Keep at most `k` singular values (MatrixAlgebraKit `truncrank`)."""

    def __init__(self, k: int):
        if int(k) < 1:
            raise ArgumentError("truncrank needs k >= 1")
        self.k = int(k)


class TruncTol:
    """Drop singular values <= rtol * max (MatrixAlgebraKit `trunctol` with rtol)."""

    def __init__(self, rtol: float = 0.0, atol: float = 0.0):
        if atol:
            raise ArgumentError("trunctol: only a relative tolerance is supported")
        self.rtol = float(rtol)


class TruncBoth:
    """Both a rank cap and a relative tolerance."""

    def __init__(self, *parts):
        self.parts = parts


def truncrank(k: int) -> TruncRank:
    return TruncRank(k)


def trunctol(rtol: float = 0.0, atol: float = 0.0) -> TruncTol:
    return TruncTol(rtol, atol)


def _trunc_params(trunc) -> Tuple[int, float]:
    maxdim, rtol = 0, -1.0
    parts = trunc.parts if isinstance(trunc, TruncBoth) else (trunc,)
    for t in parts:
        if t is None:
            continue
        if isinstance(t, TruncRank):
            maxdim = t.k if maxdim == 0 else min(maxdim, t.k)
        elif isinstance(t, TruncTol):
            rtol = max(rtol, t.rtol)
        else:
            raise ArgumentError(f"unsupported truncation {t!r}")
    return maxdim, rtol


class BPApplyGate(AbstractAlgorithm):
    """The BP simple-update strategy (apply_operators.jl:181-193)."""

    def __init__(self, trunc=None, normalize: bool = False):
        self.trunc = trunc
        self.normalize = bool(normalize)


# -------------------------------------------------------------------- marshalling

def _edge_links(tn: ITensorNetwork, a, b) -> List[Index]:
    """The links of edge {a, b} in the order the message layout fuses them: the indices of
    the edge's first vertex (in the graph's edge orientation) shared with the other."""
    if (a, b) in tn.graph._edges or not tn.graph.has_edge(a, b):
        return tn.linkinds((a, b))
    return tn.linkinds((b, a))


def _site_layout(tn: ITensorNetwork, v):
    """(site indices, neighbours in port order, links per port) of vertex v."""
    return tn.siteinds(v), tn.graph.neighbors(v), [_edge_links(tn, v, w) for w in
                                                    tn.graph.neighbors(v)]


def _port_array(t: ITensor, sites, links) -> np.ndarray:
    order = list(sites) + [i for ls in links for i in ls]
    if len(order) != len(t.inds):
        raise ArgumentError("a vertex tensor has indices that are neither site indices nor "
                            "links to its neighbours")
    d = int(np.prod([i.dim for i in sites])) if sites else 1
    dims = [int(np.prod([i.dim for i in ls])) for ls in links]
    data = t.permute(order)
    if np.issubdtype(data.dtype, np.complexfloating):
        raise ArgumentError(f"element type {data.dtype}: the gate kernel applies real gates to "
                            "real networks (Float64 or Float32)")
    return np.ascontiguousarray(data, dtype=np.float64).reshape([d] + dims)


def _gate_vertices(op: Operator, tn: ITensorNetwork) -> List[Hashable]:
    dom = set(op.domain)
    vs = [v for v in tn.vertices() if dom.intersection(tn.siteinds(v))]
    if not vs:
        raise ArgumentError("operator shares no indices with the tensor network")
    return vs


class _Gate:
    """One gate marshalled for the device: descriptor, packed inputs, and what is needed to
    rebuild the tensors afterwards."""

    def __init__(self, op: Operator, tn: ITensorNetwork, env: MessageCache, maxdim: int,
                 normalize: bool):
        self.vs = vs = _gate_vertices(op, tn)
        if len(vs) > 2:
            raise ArgumentError(f"{len(vs)}-site gate decomposition not implemented")
        self.layouts = [_site_layout(tn, v) for v in vs]
        doms = []
        for v, (sites, _, _) in zip(vs, self.layouts):
            if not set(sites) <= set(op.domain):
                raise ArgumentError(f"the gate acts on part of vertex {v!r}'s site indices; "
                                    "a gate's domain must hold all of a vertex's site indices")
            doms += list(sites)
        if set(doms) != set(op.domain):
            raise ArgumentError("the gate's domain holds indices of no vertex")
        desc = np.zeros(L.BP_GATE_DESC, dtype=np.int32)
        parts = []
        self.arrays = []
        for s, (v, (sites, nbrs, links)) in enumerate(zip(vs, self.layouts)):
            if len(nbrs) > KMAX:
                raise ArgumentError(f"vertex {v!r} has {len(nbrs)} neighbours (at most {KMAX})")
            A = _port_array(tn[v], sites, links)
            self.arrays.append(A)
            desc[1 + s] = A.shape[0]
            desc[3 + s] = len(nbrs)
            desc[9 + 8 * s: 9 + 8 * s + len(nbrs)] = A.shape[1:]
        if len(vs) == 2:
            v1, v2 = vs
            if not tn.graph.has_edge(v1, v2):
                raise ArgumentError(f"a two-site gate needs neighbouring vertices, got {v1!r}, "
                                    f"{v2!r}")
            self.ports = [self.layouts[0][1].index(v2), self.layouts[1][1].index(v1)]
            desc[0] = 2
            desc[5], desc[6] = self.ports
        else:
            self.ports = [-1]
            desc[0] = 1
            desc[5] = desc[6] = -1
        desc[7] = maxdim
        desc[8] = int(normalize)
        # element type: the network keeps its own; the kernel computes in FP64, Float32
        # tensors and gates get Float32 results, complex inputs are refused
        msgs = []
        for s, (v, (sites, nbrs, links)) in enumerate(zip(vs, self.layouts)):
            for j, w in enumerate(nbrs):
                if j == self.ports[s] or (len(vs) == 1 and not normalize):
                    continue
                msgs.append(np.asarray(env[(w, v)]))
        # the tensors' and the gate's type decide (BP messages of a Float32 network may come
        # back in FP64); complex messages are refused as well
        dt = np.result_type(*([tn[v].data for v in vs] + [op.data]))
        if any(np.iscomplexobj(m) for m in msgs):
            dt = np.result_type(dt, np.complex128)
        if np.issubdtype(dt, np.complexfloating):
            raise ArgumentError(f"element type {dt}: the gate kernel applies real gates to "
                                "real networks (Float64 or Float32)")
        self.out_dtype = np.float32 if dt == np.float32 else np.float64
        parts = [A.reshape(-1) for A in self.arrays]
        parts += [np.ascontiguousarray(m, dtype=np.float64).reshape(-1) for m in msgs]
        parts.append(np.asarray(op.matrix(doms), dtype=np.float64).reshape(-1))
        self.desc = desc
        self.data = np.concatenate(parts)

    def rebuild(self, out: np.ndarray, chi: int):
        """New ITensors for the gate's vertices and the new bond's messages."""
        new = {}
        if len(self.vs) == 1:
            v = self.vs[0]
            sites, nbrs, links = self.layouts[0]
            A = out[:self.arrays[0].size].astype(self.out_dtype)
            inds = list(sites) + [i for ls in links for i in ls]
            new[v] = ITensor(A.reshape([i.dim for i in inds]), inds)
            return new, None, None
        bond = Index(chi, "Link")
        off = 0
        caps = _out_caps(self.desc)
        for s, v in enumerate(self.vs):
            sites, nbrs, links = self.layouts[s]
            inds = list(sites)
            for j, ls in enumerate(links):
                inds += [bond] if j == self.ports[s] else list(ls)
            n = int(np.prod([i.dim for i in inds]))
            new[v] = ITensor(out[off:off + n].astype(self.out_dtype).reshape(
                [i.dim for i in inds]), inds)
            off += caps[s]
        S = out[off:off + chi].astype(self.out_dtype)
        return new, bond, S


def _out_caps(desc) -> List[int]:
    """Capacities of the packed outputs of one two-site gate: [site 1, site 2]."""
    d1, d2, k1, k2, p1, p2, maxdim = (int(x) for x in desc[1:8])
    dims1, dims2 = desc[9:9 + k1], desc[17:17 + k2]
    Db = int(dims1[p1])
    P1 = int(np.prod([dims1[j] for j in range(k1) if j != p1] or [1]))
    P2 = int(np.prod([dims2[j] for j in range(k2) if j != p2] or [1]))
    q = min(min(P1, d1 * Db) * d1, min(P2, d2 * Db) * d2)
    chimax = min(q, maxdim) if maxdim > 0 else q
    return [P1 * d1 * chimax, P2 * d2 * chimax]


def apply_gates_device(gates: List[_Gate], rtol: float, device: int = 0):
    """One bp_apply_gates call over independent gates; returns [(out, chi)] per gate and the
    kernel's device time in ms."""
    lib = L.load()
    ctx = _context(device)
    desc = np.ascontiguousarray(np.concatenate([g.desc for g in gates]), dtype=np.int32)
    ins = np.ascontiguousarray(np.concatenate([g.data for g in gates]), dtype=np.float64)
    n = len(gates)
    isz = np.zeros(n, dtype=np.int64)
    osz = np.zeros(n, dtype=np.int64)
    ip = C.POINTER(C.c_int)
    L.check(lib.bp_gate_sizes(n, desc.ctypes.data_as(ip),
                              isz.ctypes.data_as(C.POINTER(C.c_longlong)),
                              osz.ctypes.data_as(C.POINTER(C.c_longlong))), "bp_gate_sizes")
    if int(isz.sum()) != ins.size:
        raise ArgumentError(f"packed gate inputs hold {ins.size} values, the descriptors "
                            f"{int(isz.sum())}")
    out = np.zeros(max(int(osz.sum()), 1))
    chi = np.zeros(n, dtype=np.int32)
    L.check(lib.bp_apply_gates(ctx, n, desc.ctypes.data_as(ip),
                               ins.ctypes.data_as(C.POINTER(C.c_double)),
                               out.ctypes.data_as(C.POINTER(C.c_double)),
                               chi.ctypes.data_as(ip), float(rtol)), "bp_apply_gates")
    ms = C.c_double()
    L.check(lib.bp_last_run_ms(ctx, C.byref(ms)), "bp_last_run_ms")
    res, off = [], 0
    for g in range(n):
        res.append((out[off:off + int(osz[g])], int(chi[g])))
        off += int(osz[g])
    return res, float(ms.value)


# -------------------------------------------------------------------- entry points

def _alg_params(alg: BPApplyGate):
    maxdim, rtol = _trunc_params(alg.trunc)
    return maxdim, rtol, alg.normalize


def _apply_batch(ops, state: ITensorNetwork, env: MessageCache, alg: BPApplyGate, device=0):
    maxdim, rtol, normalize = _alg_params(alg)
    gates = [_Gate(op, state, env, maxdim, normalize) for op in ops]
    res, _ = apply_gates_device(gates, rtol, device)
    for g, (out, chi) in zip(gates, res):
        new, bond, S = g.rebuild(out, chi)
        for v, t in new.items():
            state[v] = t
        if bond is not None:
            v1, v2 = g.vs
            D = np.diag(S)
            env[(v1, v2)] = D
            env[(v2, v1)] = D.copy()
    return state, env


def apply_operator(operator, state, env, alg=None, **kwargs):
    """apply_operator(operator, state, env; alg, kwargs...) -> (state, env)
    (apply_operators.jl:167-176): a copy of `state` with the gate applied and a copy of
    `env` with the gate bond's messages replaced."""
    algorithm = select_algorithm(apply_operator, alg, (operator, state, env), **kwargs)
    if not isinstance(algorithm, BPApplyGate):
        raise ArgumentError(f"`apply_operator` algorithm `{algorithm!r}` not implemented.")
    return _apply_batch([operator], state.copy(), env.copy(), algorithm)


def apply_operators(operators, state, env, alg=None, **kwargs):
    """apply_operators(operators, state, env; alg, kwargs...) -> (state, env)
    (apply_operators.jl:29-66): the operators in order; each run of operators on pairwise
    disjoint vertex sets goes to the device as one batch."""
    algorithm = select_algorithm(apply_operator, alg, (operators, state, env), **kwargs)
    if not isinstance(algorithm, BPApplyGate):
        raise ArgumentError(f"`apply_operator` algorithm `{algorithm!r}` not implemented.")
    state, env = state.copy(), env.copy()
    batch, used = [], set()
    for op in operators:
        vs = set(_gate_vertices(op, state))
        if vs & used:
            state, env = _apply_batch(batch, state, env, algorithm)
            batch, used = [], set()
        batch.append(op)
        used |= vs
    if batch:
        state, env = _apply_batch(batch, state, env, algorithm)
    return state, env


register_default(apply_operator, lambda argtypes, **kw: BPApplyGate(**kw))
