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
from .messagecache import MessageCache, _flat, context_dtype, incoming_messages, iterate_diff


class MessageUpdateAlgorithm(A.AbstractAlgorithm):
    """Strategy interface (beliefpropagation.jl:218-235).  Subclasses define
    `update(cache, factors, edge)` to recompute cache[edge] in place."""

    def update(self, cache, factors, edge):  # pragma: no cover - interface
        raise A.MethodError(f"message_update_ not defined for {type(self).__name__}")


class JacobiSweep:
    """An `edges` entry that stands for one whole Jacobi sweep: `message_update_` with a
    device strategy and this token updates every directed message at once (the sweep
    problem only needs `edges` to be indexable with a length, beliefpropagation.jl:86,
    205)."""

    def __repr__(self):
        return "JacobiSweep()"


JACOBI_SWEEP = JacobiSweep()


def _factor_engine(factors, dtype: str, normalize: bool = True, path: str = "generic"):
    """One device context per (network, dtype, normalize), kept on the network object so
    per-edge updates do not re-upload the site tensors.  Messages start at the identity;
    callers upload what an update reads."""
    from .engine import Engine

    fl = _flat(factors)
    key = (dtype, bool(normalize), path)
    cache = getattr(factors, "_bp_engines", None)
    if cache is None:
        cache = {}
        try:
            factors._bp_engines = cache
        except AttributeError:
            pass
    e = cache.get(key)
    if e is None:
        e = Engine(fl.topo, fl.sites, dtype=dtype, path=path)
        if not normalize:
            e.set_option("normalize", 0)
        e.init_messages("identity")
        cache[key] = e
    return fl, e


class SimpleMessageUpdate(MessageUpdateAlgorithm):
    """Contract the source factor with its incoming messages, divide by the sum unless
    `normalize` is false (beliefpropagation.jl:237-257).

    dtype selects the device precision ("f64" keeps the reference's Float64 semantics)."""

    def __init__(self, normalize: bool = True, contraction_alg=None, dtype: str = "f64"):
        self.normalize = bool(normalize)
        self.contraction_alg = contraction_alg
        self.dtype = dtype

    def update(self, cache, factors, edge):
        fl = _flat(factors)
        fl, e = _factor_engine(factors, context_dtype(self.dtype, fl.sites, cache),
                               self.normalize)
        t = fl.topo
        q = fl.slot(edge)
        u = int(t.port_owner()[q])
        ins = [int(t.port_rev[j]) for j in range(t.port_off[u], t.port_off[u + 1]) if j != q]
        own = t.port_owner()
        if ins:
            data = []
            for r in ins:
                key = (t.vertices[own[r]], t.vertices[t.port_nbr[r]])
                data.append(np.asarray(cache[key]).reshape(-1))
            e.set_slot_messages(ins, np.concatenate(data))
        e.update_slot(q)
        cache[edge] = e.slot_messages([q]).reshape(t.msg_shape(q))
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
        if isinstance(edge, JacobiSweep):
            # one device sweep over every directed message, written back into the cache
            fl = _flat(factors)
            fl, e = _factor_engine(factors, context_dtype(self.dtype, fl.sites, cache), True,
                                   self.path)
            e.set_messages(cache.to_flat(fl.topo))
            e.sweep()
            new = MessageCache.from_flat(fl.topo, e.messages())
            for k, m in new.items():
                cache[k] = m
            return cache
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
A.register_iterate_diff(MessageCache, iterate_diff)


def default_beliefpropagation_edges(factors):
    g = factors.graph
    return forest_cover_edge_sequence(g)


# ------------------------------------------------------------------ the nested solver
# beliefpropagation.jl:94-210: the outer loop is a NestedAlgorithm whose step solves one
# sweep problem; the sweep's own loop calls message_update! once per entry of `edges`.


class BeliefPropagationProblem(A.Problem):
    def __init__(self, factors):
        self.factors = factors


class BeliefPropagationSweepProblem(A.Problem):
    def __init__(self, factors, edges):
        self.factors = factors
        self.edges = edges


class BeliefPropagationSweepState(A.State):
    pass


class BeliefPropagationSweepAlgorithm(A.Algorithm):
    """One sweep: step i updates edges[i] (beliefpropagation.jl:155-210); stops after
    len(edges) steps."""

    def __init__(self, message_update_algorithm=None, stopping_criterion=None):
        self.message_update_algorithm = (message_update_algorithm
                                         if message_update_algorithm is not None
                                         else SimpleMessageUpdate())
        self.stopping_criterion = stopping_criterion

    def initialize_state(self, problem, iterate=None, iteration: int = 0):
        st = self.stopping_criterion.initialize_state(problem, self, iterate=iterate)
        return BeliefPropagationSweepState(iterate, iteration, st)

    def step_(self, problem, state):
        edge = problem.edges[state.iteration - 1]  # Julia's 1-based edges[state.iteration]
        message_update_(self.message_update_algorithm, state.iterate, problem.factors, edge)
        return state


class BeliefPropagationState(A.NestedState):
    pass


class BeliefPropagationAlgorithm(A.NestedAlgorithm):
    """BP outer loop (beliefpropagation.jl:100-151): any `subalgorithm` AI.Algorithm runs
    one sweep per outer step; `stopping_criterion` decides when BP stops."""

    def __init__(self, edges, subalgorithm, stopping_criterion):
        self.edges = edges
        self.subalgorithm = subalgorithm
        self.stopping_criterion = stopping_criterion

    def initialize_state(self, problem, iterate=None, iteration: int = 0):
        sub = BeliefPropagationSweepProblem(problem.factors, self.edges)
        substate = A.initialize_state(sub, self.subalgorithm, iterate=iterate)
        st = self.stopping_criterion.initialize_state(problem, self, iterate=iterate)
        return BeliefPropagationState(substate, iteration, st)

    def initialize_state_(self, problem, state, iteration: int = 0, **kwargs):
        state.iteration = iteration
        self.stopping_criterion.initialize_state_(problem, self, state.stopping_criterion_state)
        return state

    def initialize_subsolve(self, problem, state):
        sub = BeliefPropagationSweepProblem(problem.factors, self.edges)
        return sub, self.subalgorithm, state.substate


def beliefpropagation(factors, messages, *, edges=None, stopping_criterion=None,
                      message_update_algorithm=None) -> MessageCache:
    """Run BP on `factors` (a NormNetwork: double-layer, matrix messages; or an
    ITensorNetwork: plain, vector messages) from `messages`; returns the converged
    MessageCache.  `result.state` records iteration, delta and at_iteration.

    SimpleMessageUpdate and JacobiGPUUpdate with the default schedule run as one device
    call each (bp_run_sequential / bp_run, the stopping rule applied on the device); any
    other strategy, or a JacobiSweep schedule, runs the reference's nested solver with one
    message_update_ call per entry of `edges`."""
    cache = messages.copy() if isinstance(messages, MessageCache) else MessageCache(messages)
    algo = A.select_algorithm(message_update_, message_update_algorithm,
                              (type(cache), type(factors), tuple))
    crit = A.select_beliefpropagation_stopping_criterion(stopping_criterion)
    maxiter, tol = crit.limits()
    fl = _flat(factors)
    topo = fl.topo
    token = edges is not None and any(isinstance(ed, JacobiSweep) for ed in edges)

    if isinstance(algo, JacobiGPUUpdate) and not token:
        if edges is not None and len(edges) != topo.nports:
            raise A.ArgumentError("a Jacobi sweep updates every directed edge once; custom "
                                  "`edges` need a sequential strategy")
        from .engine import Engine

        dtype = context_dtype(algo.dtype, fl.sites, cache)
        e = Engine(topo, fl.sites, dtype=dtype, path=algo.path)
        try:
            e.set_messages(cache.to_flat(topo))
            r = e.run(maxiter=maxiter, tol=tol, criterion=algo.criterion, use_graph=algo.use_graph)
            if algo.polish and dtype != "c128":  # complex runs are FP64 already
                e.polish_fp64(algo.polish)
            out = MessageCache.from_flat(topo, e.messages())
        finally:
            e.close()
        out.state = _state(int(r.sweeps[0]), float(r.delta[0]), int(r.at_iteration[0]), tol)
        return out

    if edges is None:
        edges = default_beliefpropagation_edges(factors)

    if isinstance(algo, SimpleMessageUpdate) and type(algo) is SimpleMessageUpdate and not token:
        from .engine import Engine

        order = [fl.slot(ed) for ed in edges]
        e = Engine(topo, fl.sites, dtype=context_dtype(algo.dtype, fl.sites, cache),
                   path="generic")
        try:
            if not algo.normalize:
                e.set_option("normalize", 0)
            e.set_messages(cache.to_flat(topo))
            r = e.run_sequential(order, maxiter=maxiter, tol=tol)
            out = MessageCache.from_flat(topo, e.messages())
        finally:
            e.close()
        out.state = _state(int(r.sweeps[0]), float(r.delta[0]), int(r.at_iteration[0]), tol)
        return out

    # the reference's nested loop (beliefpropagation.jl:84-92)
    sub = BeliefPropagationSweepAlgorithm(message_update_algorithm=algo,
                                          stopping_criterion=A.StopAfterIteration(len(edges)))
    alg = BeliefPropagationAlgorithm(edges, sub, crit)
    problem = BeliefPropagationProblem(factors)
    state = A.initialize_state(problem, alg, iterate=cache)
    A.solve_(problem, alg, state)
    out = state.iterate
    st = _criterion_state(state.stopping_criterion_state)
    out.state = _state(state.iteration, st.delta if st else float("inf"),
                       st.at_iteration if st else -1, tol)
    return out


def _criterion_state(st):
    if isinstance(st, A.StopWhenConvergedState):
        return st
    for s in getattr(st, "states", ()):
        if isinstance(s, A.StopWhenConvergedState):
            return s
    return None


def _state(iteration, delta, at, tol):
    st = A.StopWhenConvergedState(delta if tol is not None else float("inf"), at)
    st.iteration = iteration
    return st


__all__ = ["MessageUpdateAlgorithm", "SimpleMessageUpdate", "JacobiGPUUpdate", "JacobiSweep",
           "JACOBI_SWEEP", "message_update_", "beliefpropagation",
           "default_beliefpropagation_edges", "incoming_messages", "iterate_diff",
           "BeliefPropagationProblem", "BeliefPropagationAlgorithm", "BeliefPropagationState",
           "BeliefPropagationSweepProblem", "BeliefPropagationSweepAlgorithm",
           "BeliefPropagationSweepState"]
