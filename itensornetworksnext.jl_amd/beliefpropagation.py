"""beliefpropagation and the message-update strategies (the reference's plugin surface).

Restates src/beliefpropagation/beliefpropagation.jl:
  * beliefpropagation(factors, messages; edges, stopping_criterion,
    message_update_algorithm) (:69-92): the strategy is chosen by
    select_algorithm(message_update!, alg, argtypes); the default is SimpleMessageUpdate;
  * SimpleMessageUpdate (:237-257): each message is recomputed from its incoming
    neighbours and divided by its sum; updates are written into the cache immediately, so
    a sweep over `edges` (default forest_cover_edge_sequence, :14) is Gauss-Seidel;
  * StopAfterIteration(maxiter) | StopWhenConverged(tol) on iterate_diff (:16-55, :261-267).

Strategies here:
  * SimpleMessageUpdate  -- the reference's sequential semantics, each update a HIP kernel
                            (bp_run_sequential / bp_update_slot);
  * JacobiGPUUpdate      -- the MI355X path: every directed message of every network
                            updated per sweep in one launch (Jacobi), convergence tracked
                            on the device, sweeps replayed from a hipGraph.  Same fixed
                            point as the sequential schedule (per-sweep trajectories differ).
  * any user subclass of MessageUpdateAlgorithm that defines `update(cache, factors, edge)`
    runs through the generic per-edge loop, as in the reference.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from . import algorithms as A
from .graphs import forest_cover_edge_sequence
from .itensors import to_topology
from .messagecache import MessageCache, _engine, _flat, incoming_messages, iterate_diff


class MessageUpdateAlgorithm(A.AbstractAlgorithm):
    """Strategy interface (beliefpropagation.jl:218-235).  Subclasses define
    `update(cache, factors, edge)` to recompute cache[edge] in place."""

    def update(self, cache, factors, edge):  # pragma: no cover - interface
        raise A.MethodError(f"message_update_ not defined for {type(self).__name__}")


class SimpleMessageUpdate(MessageUpdateAlgorithm):
    """Contract the source factor with its incoming messages, divide by the sum.

    dtype selects the device precision ("f64" keeps the reference's Float64 semantics)."""

    def __init__(self, normalize: bool = True, contraction_alg=None, dtype: str = "f64"):
        if not normalize:
            raise NotImplementedError("the device update always sum-normalises (normalize=True)")
        self.normalize = normalize
        self.contraction_alg = contraction_alg
        self.dtype = dtype

    def update(self, cache, factors, edge):
        fl, e = _engine(factors, cache, self.dtype)
        try:
            q = fl.slot(edge)
            e.update_slot(q)
            flat = e.messages()
        finally:
            e.close()
        cache[edge] = flat[fl.topo.msg_off[q]:fl.topo.msg_off[q + 1]].reshape(fl.topo.msg_shape(q))
        return cache

    def __repr__(self):
        return f"SimpleMessageUpdate(normalize={self.normalize}, dtype={self.dtype!r})"


class JacobiGPUUpdate(MessageUpdateAlgorithm):
    """Whole Jacobi sweeps on the device.

    dtype: "f32" (MFMA vertex path at D = 8), "f64", or "bf16".
    path: "auto" | "vertex" | "generic".  criterion: "fidelity" (the reference's
    iterate_diff) or "maxabs".  polish: FP64 sweeps appended after convergence."""

    def __init__(self, dtype: str = "f32", path: str = "auto", criterion: str = "fidelity",
                 use_graph: bool = True, polish: int = 0):
        self.dtype = dtype
        self.path = path
        self.criterion = criterion
        self.use_graph = use_graph
        self.polish = int(polish)

    def update(self, cache, factors, edge):
        # one message in isolation is just the sequential update
        return SimpleMessageUpdate(dtype=self.dtype if self.dtype != "bf16" else "f32").update(
            cache, factors, edge)

    def __repr__(self):
        return f"JacobiGPUUpdate(dtype={self.dtype!r}, path={self.path!r})"


def message_update_(*args, alg=None, **kwargs):
    """message_update!(alg, cache, factors, edge) or message_update!(cache, factors, edge;
    alg, kwargs...) (beliefpropagation.jl:218-235)."""
    if args and isinstance(args[0], MessageUpdateAlgorithm):
        if alg is not None or kwargs:
            raise A.ArgumentError("pass either an algorithm instance or alg/kwargs")
        algo, cache, factors, edge = args
        return algo.update(cache, factors, edge)
    cache, factors, edge = args
    algo = A.select_algorithm(message_update_, alg, (type(cache), type(factors), type(edge)),
                              **kwargs)
    return algo.update(cache, factors, edge)


A.register_default(message_update_, lambda argtypes, **kw: SimpleMessageUpdate(**kw))


def default_beliefpropagation_edges(factors):
    g = factors.graph
    return forest_cover_edge_sequence(g)


def beliefpropagation(factors, messages, *, edges=None, stopping_criterion=None,
                      message_update_algorithm=None) -> MessageCache:
    """Run BP on `factors` (a NormNetwork: double-layer, matrix messages; or an
    ITensorNetwork: plain, vector messages) from `messages`; returns the converged
    MessageCache.  `result.state` records iteration, delta and at_iteration."""
    cache = messages.copy() if isinstance(messages, MessageCache) else MessageCache(messages)
    algo = A.select_algorithm(message_update_, message_update_algorithm,
                              (type(cache), type(factors), tuple))
    crit = A.select_beliefpropagation_stopping_criterion(stopping_criterion)
    maxiter, tol = crit.limits()
    fl = _flat(factors)
    topo = fl.topo

    if isinstance(algo, JacobiGPUUpdate):
        if edges is not None and len(edges) != topo.nports:
            raise A.ArgumentError("a Jacobi sweep updates every directed edge once; custom "
                                  "`edges` need a sequential strategy")
        from .engine import Engine

        e = Engine(topo, fl.sites, dtype=algo.dtype, path=algo.path)
        try:
            e.set_messages(cache.to_flat(topo))
            r = e.run(maxiter=maxiter, tol=tol, criterion=algo.criterion, use_graph=algo.use_graph)
            if algo.polish:
                e.polish_fp64(algo.polish)
            out = MessageCache.from_flat(topo, e.messages())
        finally:
            e.close()
        out.state = _state(int(r.sweeps[0]), float(r.delta[0]), int(r.at_iteration[0]), tol)
        return out

    if edges is None:
        edges = default_beliefpropagation_edges(factors)

    if isinstance(algo, SimpleMessageUpdate):
        from .engine import Engine

        order = [fl.slot(ed) for ed in edges]
        e = Engine(topo, fl.sites, dtype=algo.dtype, path="generic")
        try:
            e.set_messages(cache.to_flat(topo))
            r = e.run_sequential(order, maxiter=maxiter, tol=tol)
            out = MessageCache.from_flat(topo, e.messages())
        finally:
            e.close()
        out.state = _state(int(r.sweeps[0]), float(r.delta[0]), int(r.at_iteration[0]), tol)
        return out

    # generic strategy: the reference's nested loop, one message_update! per edge
    prev = cache.copy()
    it, at, delta = 0, -1, float("inf")
    while True:
        if maxiter is not None and it >= maxiter:
            break
        if tol is not None:
            d = iterate_diff(cache, prev)
            prev = cache.copy()
            if it > 0:
                delta = d
                if d < tol:
                    at = it
                    break
        it += 1
        for ed in edges:
            algo.update(cache, factors, ed)
    cache.state = _state(it, delta, at, tol)
    return cache


def _state(iteration, delta, at, tol):
    st = A.StopWhenConvergedState(delta if tol is not None else float("inf"), at)
    st.iteration = iteration
    return st


__all__ = ["MessageUpdateAlgorithm", "SimpleMessageUpdate", "JacobiGPUUpdate",
           "message_update_", "beliefpropagation", "default_beliefpropagation_edges",
           "incoming_messages", "iterate_diff"]
