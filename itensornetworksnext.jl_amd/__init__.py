"""MI355X-native belief-propagation engine for ITensorNetworksNext-style networks.

The package mirrors the reference's BP API (src/beliefpropagation/*.jl,
src/select_algorithm.jl, src/AlgorithmsInterfaceExtensions, src/normnetwork.jl,
src/tensornetwork.jl) in Python and runs every message update, normalisation, residual,
vertex contraction and adjoint in hand-written HIP kernels for gfx950 behind the C-ABI
library `csrc/libbpmi355x.so` (declared in include/bp_mi355x.h).  There is no CPU
fallback: the compute entry points raise when the library or a GPU is missing.
"""
from . import (_lib, algorithms, apply, build, contract, engine, graphs, itensors,  # noqa: F401
               peps, resident)
from .algorithms import (AbstractAlgorithm, ArgumentError, MethodError,  # noqa: F401
                         NestedAlgorithm, NestedState, StopAfterIteration, StopWhenAny,
                         StopWhenConverged, default_algorithm, finalize_substate_,
                         initialize_subsolve, select_algorithm,
                         select_beliefpropagation_stopping_criterion)
from .beliefpropagation import (JACOBI_SWEEP, BeliefPropagationAlgorithm,  # noqa: F401
                                BeliefPropagationProblem, BeliefPropagationSweepAlgorithm,
                                BeliefPropagationSweepProblem, JacobiGPUUpdate, JacobiSweep,
                                MessageUpdateAlgorithm, SimpleMessageUpdate, beliefpropagation,
                                default_beliefpropagation_edges, message_update_)
from .apply import (BPApplyGate, Operator, apply_operator, apply_operators,  # noqa: F401
                    operator, truncrank, trunctol)
from .contract import (Cheapest, ContractionPlan, Exact, Flat, Greedy,  # noqa: F401
                       LeftAssociative, contract_labeled, contract_network)
from .engine import Engine, RunResult  # noqa: F401
from .resident import DeviceNetwork  # noqa: F401
from .graphs import (NamedGraph, Topology, build_topology, concat_topologies,  # noqa: F401
                     forest_cover_edge_sequence, heavy_hex_lattice, named_comb_tree,
                     named_grid, named_path_graph)
from .itensors import (BraView, Index, ITensor, ITensorNetwork, KetView,  # noqa: F401
                       NormNetwork, braname, bratensor, conj_bratensor, is_vertex_assigned,
                       isinsertable, issettable, kettensor, linkinds, linknames, normnetwork,
                       random_itensor, siteinds, tensornetwork, to_topology, uniquename)
from .messagecache import (MessageCache, add_vertex_, bethe_free_energy,  # noqa: F401
                           copyto_, edge_data, edge_data_type, edge_scalar, edge_scalars,
                           expect, expect2, has_edge, incoming_edge_data, incoming_messages,
                           insert_edge_data_, iterate_diff, map_, message_environment,
                           messagecache, one, region_scalar, rem_edge_, set_edge_data_,
                           similar_message_environment, subgraph, valtype, vertex_scalar,
                           vertex_scalars)

__all__ = [
    "AbstractAlgorithm", "ArgumentError", "BPApplyGate", "Cheapest", "DeviceNetwork",
    "Operator",
    "apply_operator", "apply_operators", "operator", "truncrank", "trunctol", "ContractionPlan", "Engine", "Exact", "Flat",
    "Greedy", "LeftAssociative", "contract_labeled", "contract_network", "ITensor", "ITensorNetwork", "Index",
    "JacobiGPUUpdate", "MessageCache", "MessageUpdateAlgorithm", "MethodError", "NamedGraph",
    "NormNetwork", "RunResult", "SimpleMessageUpdate", "StopAfterIteration", "StopWhenAny",
    "StopWhenConverged", "Topology", "beliefpropagation", "bethe_free_energy",
    "build_topology", "concat_topologies", "default_algorithm",
    "default_beliefpropagation_edges", "edge_scalar", "edge_scalars", "expect", "expect2",
    "forest_cover_edge_sequence", "heavy_hex_lattice", "incoming_edge_data",
    "incoming_messages", "iterate_diff", "linkinds", "message_environment", "message_update_",
    "messagecache", "named_comb_tree", "named_grid", "named_path_graph", "normnetwork", "one",
    "random_itensor", "region_scalar", "select_algorithm",
    "select_beliefpropagation_stopping_criterion", "similar_message_environment", "siteinds",
    "tensornetwork", "to_topology", "uniquename", "vertex_scalar", "vertex_scalars",
    "NestedAlgorithm", "NestedState", "initialize_subsolve", "finalize_substate_",
    "JACOBI_SWEEP", "JacobiSweep", "BeliefPropagationAlgorithm", "BeliefPropagationProblem",
    "BeliefPropagationSweepAlgorithm", "BeliefPropagationSweepProblem", "BraView", "KetView",
    "braname", "bratensor", "conj_bratensor", "kettensor", "is_vertex_assigned",
    "isinsertable", "issettable", "linknames", "add_vertex_", "copyto_", "edge_data",
    "edge_data_type", "has_edge", "insert_edge_data_", "map_", "rem_edge_", "set_edge_data_",
    "subgraph", "valtype",
]
