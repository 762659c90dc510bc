"""Host-side mirrors of reference tests that need no arithmetic on the device (CPU only).

  test/test_algorithmsinterfaceextensions.jl:95-139  NestedAlgorithm defaults and the nested
                                                     solve (iteration 2, iterate +3)
  test/test_beliefpropagation.jl:40-82               MessageCache basics: string messages,
                                                     setindex, copyto! with and without edges
  test/test_beliefpropagation.jl:117-132             subgraph of a MessageCache
  test/test_normnetwork.jl:106-134                   KetView / BraView
"""
import numpy as np
import pytest


# --------------------------------------------------------------- nested algorithms

def _nested_fixtures(pkg):
    A = pkg.algorithms

    class TestProblem(A.Problem):
        pass

    class TestChildState(A.State):
        pass

    class TestChildAlgorithm(A.Algorithm):
        def __init__(self, stopping_criterion=None):
            self.stopping_criterion = stopping_criterion or A.StopAfterIteration(2)

        def initialize_state(self, problem, iterate=None, **kw):
            sc = self.stopping_criterion.initialize_state(problem, self, iterate=iterate)
            return TestChildState(iterate, 0, sc)

        def step_(self, problem, state):
            state.iterate += 1.0  # in place, as `state.iterate .+= 1`
            return state

    class TestNestedAlgorithm(A.NestedAlgorithm):
        def __init__(self, algorithms, stopping_criterion=None):
            self.algorithms = list(algorithms)
            self.stopping_criterion = stopping_criterion or A.StopAfterIteration(len(self.algorithms))

        def initialize_state(self, problem, iterate=None, **kw):
            sc = self.stopping_criterion.initialize_state(problem, self, iterate=iterate)
            return TestChildState(iterate, 0, sc)

        def initialize_subsolve(self, problem, state):
            sub = self.algorithms[state.iteration - 1]  # algorithms[state.iteration]
            return problem, sub, A.initialize_state(problem, sub, iterate=state.iterate)

    return A, TestProblem, TestChildAlgorithm, TestNestedAlgorithm


def test_nested_algorithm_defaults(pkg):
    A, TestProblem, TestChildAlgorithm, _ = _nested_fixtures(pkg)
    problem = TestProblem()
    alg = TestChildAlgorithm()
    state = A.initialize_state(problem, alg, iterate=np.array([0.0]))
    with pytest.raises(pkg.MethodError):
        A.initialize_subsolve(problem, alg, state)
    sub = A.initialize_state(problem, alg, iterate=np.array([42.0]))
    A.finalize_substate_(problem, alg, state, sub)
    assert np.array_equal(state.iterate, [42.0])


def test_nested_solve_counts(pkg):
    A, TestProblem, TestChildAlgorithm, TestNestedAlgorithm = _nested_fixtures(pkg)
    problem = TestProblem()
    nested = TestNestedAlgorithm([TestChildAlgorithm(A.StopAfterIteration(1)),
                                  TestChildAlgorithm(A.StopAfterIteration(2))])
    assert isinstance(nested, A.NestedAlgorithm)
    state = A.initialize_state(problem, nested, iterate=np.array([0.0, 0.0]))
    A.solve_(problem, nested, state, iterate=np.array([0.0, 0.0]))
    # 1 inner step + 2 inner steps = 3 increments of the shared iterate
    assert state.iteration == 2
    np.testing.assert_allclose(state.iterate, [3.0, 3.0])


def test_nested_state_forwards_iterate(pkg):
    A = pkg.algorithms
    inner = A.State(iterate=np.zeros(2))
    outer = A.NestedState(inner)
    assert outer.iterate is inner.iterate
    outer.iterate = np.ones(3)
    assert inner.iterate.shape == (3,)


def test_stop_when_converged_state_machine(pkg):
    """AIE.jl:84-109 on a scalar iterate: the check at iteration 0 is skipped, the previous
    iterate is refreshed every call, at_iteration records the stopping iteration."""
    A = pkg.algorithms

    class Box:
        def __init__(self, x):
            self.x = x

        def copy(self):
            return Box(self.x)

    A.register_iterate_diff(Box, lambda a, b: abs(a.x - b.x))

    class Halve(A.Algorithm):
        def __init__(self, crit):
            self.stopping_criterion = crit

        def initialize_state(self, problem, iterate=None, **kw):
            return A.State(iterate, 0, self.stopping_criterion.initialize_state(problem, self,
                                                                                iterate=iterate))

        def step_(self, problem, state):
            state.iterate.x *= 0.5
            return state

    crit = A.StopAfterIteration(100) | A.StopWhenConverged(0.1)
    alg = Halve(crit)
    st = A.initialize_state(A.Problem(), alg, iterate=Box(1.0))
    A.solve_(A.Problem(), alg, st)
    # differences 0.5, 0.25, 0.125, 0.0625 -> stops at iteration 4
    assert st.iteration == 4
    conv = st.stopping_criterion_state.states[1]
    assert conv.at_iteration == 4 and conv.delta == pytest.approx(0.0625)
    with pytest.raises(pkg.MethodError):
        A.iterate_diff(object(), object())


# --------------------------------------------------------------- MessageCache

def test_messagecache_string_messages(pkg):
    g = pkg.named_grid((3, 3))
    bpc = pkg.messagecache(lambda e: f"{e[0]} => {e[1]}", g.all_edges())
    assert pkg.valtype(bpc) is str and pkg.edge_data_type(bpc) is str
    assert len(pkg.edge_data(bpc)) == 2 * g.ne()
    assert bpc[((1, 1), (1, 2))] == "(1, 1) => (1, 2)"
    bpc[((1, 1), (1, 2))] = "new message"
    assert bpc[((1, 1), (1, 2))] == "new message"
    pairs = [(((1, 2), (2, 2)), "m1"), (((2, 2), (2, 3)), "m2")]
    new = pkg.copyto_(bpc.copy(), dict(pairs))
    assert new[((1, 1), (1, 2))] == "new message"
    assert new[((1, 2), (2, 2))] == "m1" and new[((2, 2), (2, 3))] == "m2"
    assert bpc[((1, 2), (2, 2))] == "(1, 2) => (2, 2)"  # the copy was written, not bpc
    dst = pkg.messagecache(lambda e: "", g.all_edges())
    pkg.copyto_(dst, bpc, [((1, 2), (2, 2)), ((2, 2), (2, 3))])
    assert dst[((1, 1), (1, 2))] == ""
    assert dst[((1, 2), (2, 2))] == "(1, 2) => (2, 2)"
    assert dst[((2, 2), (2, 3))] == "(2, 2) => (2, 3)"


def test_messagecache_graph_edits(pkg):
    g = pkg.named_path_graph(3)
    c = pkg.messagecache(lambda e: np.ones((2, 2)), g.all_edges())
    assert pkg.has_edge(c, (1, 2)) and c.has_edge((2, 1))
    pkg.rem_edge_(c, (1, 2))
    assert not pkg.has_edge(c, (1, 2)) and (1, 2) not in c and len(c) == 3
    pkg.set_edge_data_(c, np.eye(2), (1, 2))  # set_edge_data! adds a missing edge
    assert pkg.has_edge(c, (1, 2))
    with pytest.raises(KeyError):
        pkg.insert_edge_data_(c, (1, 2), np.eye(2))
    pkg.add_vertex_(c, 7)
    assert 7 in c.vertices()
    doubled = c.map(lambda m: 2 * m)
    np.testing.assert_array_equal(doubled[(2, 3)], 2 * np.ones((2, 2)))
    pkg.map_(lambda m: -m, doubled, c)
    np.testing.assert_array_equal(doubled[(2, 3)], -np.ones((2, 2)))


def test_messagecache_subgraph(pkg):
    g = pkg.named_grid((3,))
    bpc = pkg.messagecache(lambda e: np.ones((2, 2)), g.all_edges())
    sub = pkg.subgraph(bpc, [(1,), (2,)])
    assert isinstance(sub, pkg.MessageCache)
    assert set(sub.vertices()) == {(1,), (2,)}
    assert pkg.has_edge(sub, ((1,), (2,))) and pkg.has_edge(sub, ((2,), (1,)))
    assert not pkg.has_edge(sub, ((2,), (3,)))
    np.testing.assert_array_equal(sub[((1,), (2,))], np.ones((2, 2)))


# --------------------------------------------------------------- KetView / BraView

def test_ket_and_bra_views(pkg):
    g = pkg.named_path_graph(3)
    l = {frozenset(e): pkg.Index(2) for e in g.edges()}
    s = {v: pkg.Index(2) for v in g.vertices()}
    tn = pkg.tensornetwork(lambda v: pkg.random_itensor(
        [s[v]] + [l[frozenset(e)] for e in g.incident_edges(v)]), g.vertices())
    nn = pkg.NormNetwork(tn)
    kv, bv = pkg.KetView(nn), pkg.BraView(nn)
    assert set(kv.vertices()) == set(tn.vertices()) == set(bv.vertices())
    assert set(map(frozenset, kv.edges())) == set(map(frozenset, tn.edges()))
    assert set(map(frozenset, bv.edges())) == set(map(frozenset, tn.edges()))
    for v in tn.vertices():
        assert kv[v] is pkg.kettensor(nn, v)
        assert list(bv[v].inds) == list(pkg.bratensor(nn, v).inds)
    assert pkg.is_vertex_assigned(kv, 1) and pkg.is_vertex_assigned(bv, 1)
    for view in (kv, bv):
        assert not pkg.issettable(view) and not pkg.isinsertable(view)
    # the bra view's link names are the ket names mapped by braname (messagecache.jl:205-218)
    e = (1, 2)
    ket = pkg.linknames(kv, e)
    bra = pkg.linknames(bv, e)
    assert bra == [pkg.braname(nn, i) for i in ket] and set(bra).isdisjoint(ket)
