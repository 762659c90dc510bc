"""Algorithm selection and stopping criteria, as the reference's plugin surface.

select_algorithm / default_algorithm restate src/select_algorithm.jl:7-57: an algorithm
instance passes through, `None` resolves to the registered default, a dict (the
NamedTuple form) resolves to the default with those keyword arguments, and keyword
arguments beside an instance or a dict are an error.  Stopping criteria restate
AlgorithmsInterface's StopAfterIteration, the `|` combinator and
AlgorithmsInterfaceExtensions' StopWhenConverged (AlgorithmsInterfaceExtensions.jl:57-119),
including the skipped check at iteration 0.
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


# ------------------------------------------------------------------ stopping criteria


class StoppingCriterion:
    def __or__(self, other: "StoppingCriterion") -> "StopWhenAny":
        return StopWhenAny(self, other)

    # (maxiter, tol) the device loop enforces for this criterion
    def limits(self):
        raise NotImplementedError


class StopAfterIteration(StoppingCriterion):
    def __init__(self, maxiter: int):
        self.maxiter = int(maxiter)

    def limits(self):
        return self.maxiter, None

    def __repr__(self):
        return f"StopAfterIteration({self.maxiter})"


class StopWhenConverged(StoppingCriterion):
    """Fires once iterate_diff(iterate, previous) < tol; never at iteration 0."""

    def __init__(self, tol: float):
        self.tol = float(tol)

    def limits(self):
        return None, self.tol

    def __repr__(self):
        return f"StopWhenConverged(tol={self.tol})"


class StopWhenAny(StoppingCriterion):
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

    def __repr__(self):
        return " | ".join(map(repr, self.criteria))


class StopWhenConvergedState:
    """delta / at_iteration as the reference records them (AIE.jl:67-71)."""

    def __init__(self, delta: float = math.inf, at_iteration: int = -1):
        self.delta = delta
        self.at_iteration = at_iteration

    def __repr__(self):
        return f"StopWhenConvergedState(delta={self.delta}, at_iteration={self.at_iteration})"


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
