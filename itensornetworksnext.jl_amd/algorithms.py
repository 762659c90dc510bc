"""Algorithm selection, the solver loop and stopping criteria, as the reference's plugin
surface.

select_algorithm / default_algorithm restate src/select_algorithm.jl:7-57: an algorithm
instance passes through, `None` resolves to the registered default, a dict (the
NamedTuple form) resolves to the default with those keyword arguments, and keyword
arguments beside an instance or a dict are an error.

The solver loop restates the part of AlgorithmsInterface that the reference builds on
(`solve`, `solve!`, `initialize_state(!)`, `is_finished!`, `increment!`, `step!`) and
AlgorithmsInterfaceExtensions (src/AlgorithmsInterfaceExtensions/
AlgorithmsInterfaceExtensions.jl): `NestedAlgorithm` / `NestedState` /
`initialize_subsolve` / `finalize_substate!` (:7-53) and `StopWhenConverged` (:57-119),
including the skipped check at iteration 0.  Julia's dispatch on (problem, algorithm,
state) becomes methods on the algorithm and criterion objects; the module-level functions
keep the reference's names (a trailing `_` stands for Julia's `!`).
"""
from __future__ import annotations

import math
from typing import Any, Callable, Dict, Optional


class ArgumentError(ValueError):
    """Julia's ArgumentError."""


class MethodError(TypeError):
    """Julia's MethodError (no method for these argument types)."""


class AbstractAlgorithm:
    """Base of every strategy type (select_algorithm.jl:7)."""


_DEFAULTS: Dict[Callable, Callable[..., AbstractAlgorithm]] = {}


def register_default(f: Callable, factory: Callable[..., AbstractAlgorithm]) -> None:
    """`default_algorithm(::typeof(f), ::Type{<:Tuple}; kwargs...)` method definition."""
    _DEFAULTS[f] = factory


def default_algorithm(f: Callable, argtypes: Any = (), **kwargs) -> AbstractAlgorithm:
    if f not in _DEFAULTS:
        raise MethodError(f"no default algorithm for {getattr(f, '__name__', f)} with "
                          f"arguments {argtypes}")
    return _DEFAULTS[f](argtypes, **kwargs)


def select_algorithm(f: Callable, alg, args: Any = (), **kwargs) -> AbstractAlgorithm:
    if alg is None:
        return default_algorithm(f, args, **kwargs)
    if isinstance(alg, dict):
        if kwargs:
            raise ArgumentError("Additional keyword arguments are not allowed when `alg` is a "
                                "`NamedTuple`.")
        return default_algorithm(f, args, **alg)
    if isinstance(alg, AbstractAlgorithm):
        if kwargs:
            raise ArgumentError("Additional keyword arguments are not allowed when `alg` is an "
                                "`AbstractAlgorithm` instance.")
        return alg
    raise MethodError(f"cannot select an algorithm from {alg!r}")


def initialize_output(f, *args, **kwargs):
    raise MethodError(f"initialize_output not defined for {getattr(f, '__name__', f)}")


# ------------------------------------------------------------------ solver loop
# AlgorithmsInterface (external; v0.1 as the reference's Project.toml pins) restated:
#   solve(problem, algorithm; kwargs...)  = solve!(problem, algorithm,
#                                                  initialize_state(problem, algorithm; kwargs...))
#   solve!(problem, algorithm, state; kwargs...):
#       initialize_state!(problem, algorithm, state; kwargs...)
#       while !is_finished!(problem, algorithm, state)
#           increment!(problem, algorithm, state); step!(problem, algorithm, state)
# `solve` returns the final iterate (beliefpropagation.jl:91 returns the cache it solved).


class Problem:
    """AI.Problem."""


class Algorithm:
    """AI.Algorithm.  Subclasses define `initialize_state(problem, **kw)`, optionally
    `initialize_state_(problem, state, **kw)`, and `step_(problem, state)`; they carry a
    `stopping_criterion`."""

    stopping_criterion: "StoppingCriterion"

    def initialize_state(self, problem, **kwargs):
        raise MethodError(f"initialize_state not defined for {type(self).__name__}")

    def initialize_state_(self, problem, state, iteration: int = 0, **kwargs):
        for k, v in kwargs.items():
            setattr(state, k, v)
        state.iteration = iteration
        self.stopping_criterion.initialize_state_(problem, self,
                                                  state.stopping_criterion_state)
        return state

    def increment_(self, problem, state):
        state.iteration += 1
        return state

    def step_(self, problem, state):
        raise MethodError(f"step! not defined for {type(self).__name__}")


class State:
    """AI.State: iterate, iteration, stopping_criterion_state."""

    def __init__(self, iterate=None, iteration: int = 0, stopping_criterion_state=None):
        self.iterate = iterate
        self.iteration = iteration
        self.stopping_criterion_state = stopping_criterion_state


class StoppingCriterionState:
    """AI.StoppingCriterionState."""


def initialize_state(problem, algorithm, **kwargs):
    return algorithm.initialize_state(problem, **kwargs)


def initialize_state_(problem, algorithm, state, **kwargs):
    return algorithm.initialize_state_(problem, state, **kwargs)


def increment_(problem, algorithm, state):
    return algorithm.increment_(problem, state)


def step_(problem, algorithm, state):
    return algorithm.step_(problem, state)


def is_finished_(problem, algorithm, state) -> bool:
    return algorithm.stopping_criterion.is_finished_(problem, algorithm, state,
                                                     state.stopping_criterion_state)


def solve_(problem, algorithm, state, **kwargs):
    initialize_state_(problem, algorithm, state, **kwargs)
    while not is_finished_(problem, algorithm, state):
        increment_(problem, algorithm, state)
        step_(problem, algorithm, state)
    return state


def solve(problem, algorithm, **kwargs):
    state = initialize_state(problem, algorithm, **kwargs)
    solve_(problem, algorithm, state)
    return state.iterate


def _copy(x):
    return x.copy() if hasattr(x, "copy") else x


# --------------------------------------------------------------- nested algorithms


class NestedAlgorithm(Algorithm):
    """AIE.NestedAlgorithm (AlgorithmsInterfaceExtensions.jl:7-32): one outer step solves
    the subproblem that `initialize_subsolve` returns, then `finalize_substate_` copies the
    substate's iterate back."""

    def step_(self, problem, state):
        subproblem, subalgorithm, substate = initialize_subsolve(problem, self, state)
        solve_(subproblem, subalgorithm, substate)
        finalize_substate_(problem, self, state, substate)
        return state


def initialize_subsolve(problem, algorithm, state):
    """Returns (subproblem, subalgorithm, substate); NestedAlgorithm subtypes define
    `initialize_subsolve(problem, state)` (AIE.jl:15-19: the bare default is a
    MethodError)."""
    f = getattr(algorithm, "initialize_subsolve", None)
    if f is None:
        raise MethodError(f"initialize_subsolve not defined for ({type(problem).__name__}, "
                          f"{type(algorithm).__name__}, {type(state).__name__})")
    return f(problem, state)


def finalize_substate_(problem, algorithm, state, substate):
    """AIE.jl:21-26: the parent's iterate becomes the substate's (the same object)."""
    f = getattr(algorithm, "finalize_substate_", None)
    if f is not None:
        return f(problem, state, substate)
    state.iterate = substate.iterate
    return state


class NestedState(State):
    """AIE.NestedState (AIE.jl:34-53): `iterate` reads and writes the substate's, so the
    inner loop's iterate is shared without a second copy.  Subclasses set `substate`."""

    def __init__(self, substate, iteration: int = 0, stopping_criterion_state=None):
        object.__setattr__(self, "substate", substate)
        self.iteration = iteration
        self.stopping_criterion_state = stopping_criterion_state

    @property
    def iterate(self):
        return self.substate.iterate

    @iterate.setter
    def iterate(self, value):
        self.substate.iterate = value


# ------------------------------------------------------------------ stopping criteria

_ITERATE_DIFF: Dict[type, Callable] = {}


def register_iterate_diff(T: type, f: Callable) -> None:
    """An `AIE.iterate_diff(::T, ::T)` method."""
    _ITERATE_DIFF[T] = f


def iterate_diff(a, b) -> float:
    """AIE.iterate_diff (AIE.jl:57-61): a MethodError unless the iterate type defines it."""
    for T, f in _ITERATE_DIFF.items():
        if isinstance(a, T):
            return f(a, b)
    raise MethodError(f"iterate_diff not defined for ({type(a).__name__}, {type(b).__name__})")


class StoppingCriterion:
    def __or__(self, other: "StoppingCriterion") -> "StopWhenAny":
        return StopWhenAny(self, other)

    # (maxiter, tol) the device loop enforces for this criterion
    def limits(self):
        raise NotImplementedError

    def initialize_state(self, problem, algorithm, iterate=None):
        return StoppingCriterionState()

    def initialize_state_(self, problem, algorithm, st):
        return st

    def is_finished_(self, problem, algorithm, state, st) -> bool:
        raise NotImplementedError


class StopAfterIterationState(StoppingCriterionState):
    def __init__(self):
        self.at_iteration = -1


class StopAfterIteration(StoppingCriterion):
    """Finished once `state.iteration >= maxiter`."""

    def __init__(self, maxiter: int):
        self.maxiter = int(maxiter)

    def limits(self):
        return self.maxiter, None

    def initialize_state(self, problem, algorithm, iterate=None):
        return StopAfterIterationState()

    def initialize_state_(self, problem, algorithm, st):
        st.at_iteration = -1
        return st

    def is_finished_(self, problem, algorithm, state, st) -> bool:
        if state.iteration >= self.maxiter:
            st.at_iteration = state.iteration
            return True
        return False

    def __repr__(self):
        return f"StopAfterIteration({self.maxiter})"


class StopWhenConvergedState(StoppingCriterionState):
    """delta / at_iteration / previous_iterate as the reference records them (AIE.jl:67-71)."""

    def __init__(self, delta: float = math.inf, at_iteration: int = -1, previous_iterate=None):
        self.delta = delta
        self.at_iteration = at_iteration
        self.previous_iterate = previous_iterate

    def __repr__(self):
        return f"StopWhenConvergedState(delta={self.delta}, at_iteration={self.at_iteration})"


class StopWhenConverged(StoppingCriterion):
    """Fires once iterate_diff(iterate, previous) < tol; never at iteration 0."""

    def __init__(self, tol: float):
        self.tol = float(tol)

    def limits(self):
        return None, self.tol

    def initialize_state(self, problem, algorithm, iterate=None):
        return StopWhenConvergedState(previous_iterate=_copy(iterate))

    def initialize_state_(self, problem, algorithm, st):
        st.delta = math.inf
        return st

    def is_finished_(self, problem, algorithm, state, st) -> bool:
        # AIE.jl:84-109: the difference is taken and the previous iterate refreshed every
        # time; at iteration 0 (delta would be 0) nothing is recorded
        delta = iterate_diff(state.iterate, st.previous_iterate)
        st.previous_iterate = _copy(state.iterate)
        if state.iteration == 0:
            return False
        st.delta = delta
        if st.delta < self.tol:
            st.at_iteration = state.iteration
            return True
        return False

    def __repr__(self):
        return f"StopWhenConverged(tol={self.tol})"


class StopWhenAnyState(StoppingCriterionState):
    def __init__(self, states):
        self.states = tuple(states)

    def __getattr__(self, name):
        # delta / at_iteration of the convergence criterion, as the BP state reports them
        for st in self.__dict__.get("states", ()):
            if isinstance(st, StopWhenConvergedState) and hasattr(st, name):
                return getattr(st, name)
        raise AttributeError(name)


class StopWhenAny(StoppingCriterion):
    """`c1 | c2`: finished when any criterion is; every criterion is evaluated (so each
    keeps its own state current)."""

    def __init__(self, *crits: StoppingCriterion):
        flat = []
        for c in crits:
            flat += list(c.criteria) if isinstance(c, StopWhenAny) else [c]
        self.criteria = tuple(flat)

    def limits(self):
        maxiter, tol = None, None
        for c in self.criteria:
            m, t = c.limits()
            if m is not None:
                maxiter = m if maxiter is None else min(maxiter, m)
            if t is not None:
                tol = t if tol is None else max(tol, t)
        return maxiter, tol

    def initialize_state(self, problem, algorithm, iterate=None):
        return StopWhenAnyState(c.initialize_state(problem, algorithm, iterate=iterate)
                                for c in self.criteria)

    def initialize_state_(self, problem, algorithm, st):
        for c, s in zip(self.criteria, st.states):
            c.initialize_state_(problem, algorithm, s)
        return st

    def is_finished_(self, problem, algorithm, state, st) -> bool:
        done = [c.is_finished_(problem, algorithm, state, s)
                for c, s in zip(self.criteria, st.states)]
        return any(done)

    def __repr__(self):
        return " | ".join(map(repr, self.criteria))


def select_beliefpropagation_stopping_criterion(spec=None, **kwargs) -> StoppingCriterion:
    """beliefpropagation.jl:16-55: a criterion passes through; a dict/kwargs with maxiter and/or
    tol builds StopAfterIteration | StopWhenConverged; anything else is an error."""
    if isinstance(spec, StoppingCriterion):
        return spec
    if spec is None and not kwargs:
        raise ArgumentError(
            "`stopping_criterion` must be specified, e.g.\n"
            "  `stopping_criterion = (; maxiter = 10)`,\n"
            "  `stopping_criterion = (; maxiter = 10, tol = 1.0e-10)`, or\n"
            "  `stopping_criterion = AI.StopAfterIteration(10) | StopWhenConverged(1.0e-10)`.")
    opts = dict(spec or {})
    opts.update(kwargs)
    maxiter = opts.pop("maxiter", None)
    tol = opts.pop("tol", None)
    if opts:
        raise ArgumentError(f"Unrecognized `stopping_criterion` kwargs: {tuple(opts)}. "
                            "Supported: `maxiter`, `tol`.")
    if maxiter is None and tol is None:
        raise ArgumentError("At least one of `maxiter` or `tol` must be specified.")
    crit: Optional[StoppingCriterion] = None
    if maxiter is not None:
        crit = StopAfterIteration(maxiter)
    if tol is not None:
        conv = StopWhenConverged(tol)
        crit = conv if crit is None else crit | conv
    return crit
