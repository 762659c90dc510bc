"""Host logic of the device contraction (no GPU): the nested-order walker, the order
algorithm names and the algorithm dispatch of contract_network (contract_network.jl:12-17)."""
import pytest


def test_order_pairs_postfix(pkg):
    walk = pkg.contract._order_pairs
    # ((0 1) (2 3)): 0*1 -> 4, 2*3 -> 5, 4*5 -> 6
    assert walk(((0, 1), (2, 3)), [0, 1, 2, 3]) == [0, 1, 2, 3, 4, 5]
    # left fold of a node with three children
    assert walk((0, 1, 2), [0, 1, 2]) == [0, 1, 3, 2]
    # tuple-valued keys (grid vertices) are keys, not sub-products
    keys = [(1, 1), (1, 2), (2, 1)]
    assert walk((((1, 1), (1, 2)), (2, 1)), keys) == [0, 1, 3, 2]
    with pytest.raises(pkg.ArgumentError):
        walk(((0, 7),), [0, 1])


def test_order_algorithm_codes(pkg):
    code = pkg.contract._order_alg_code
    L = pkg._lib
    assert code(pkg.LeftAssociative()) == L.BP_CN_ORDER_LEFT
    assert code(pkg.Flat()) == L.BP_CN_ORDER_LEFT
    assert code(pkg.Greedy()) == L.BP_CN_ORDER_GREEDY
    assert code(pkg.Cheapest()) == L.BP_CN_ORDER_CHEAPEST
    assert code("greedy") == L.BP_CN_ORDER_GREEDY
    with pytest.raises(pkg.ArgumentError):
        code("optimal")
    assert isinstance(pkg.Exact().order_alg, pkg.Cheapest)


def test_contract_network_rejects_unknown_algorithm(pkg):
    t = pkg.random_itensor([pkg.Index(2)])
    with pytest.raises(pkg.ArgumentError, match="not implemented"):
        pkg.contract_network([t], alg="bp")
    with pytest.raises(pkg.ArgumentError):
        pkg.contract_network([t], alg=pkg.Greedy())
