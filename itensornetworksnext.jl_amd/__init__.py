"""MI355X-native belief-propagation engine for ITensorNetworksNext-style networks.

The package mirrors the reference's BP API (src/beliefpropagation/*.jl,
src/select_algorithm.jl, src/AlgorithmsInterfaceExtensions) in Python and runs every
message update, normalisation, residual and vertex contraction in hand-written HIP
kernels for gfx950 behind the C-ABI library `csrc/libbpmi355x.so` (declared in
include/bp_mi355x.h).  There is no CPU fallback: the compute entry points raise when the
library or a GPU is missing.
"""
from . import _lib, build, engine, graphs, peps  # noqa: F401
from .engine import Engine, RunResult  # noqa: F401
from .graphs import (NamedGraph, Topology, build_topology, concat_topologies,  # noqa: F401
                     forest_cover_edge_sequence, heavy_hex_lattice, named_comb_tree,
                     named_grid, named_path_graph)

__all__ = ["Engine", "RunResult", "engine", "graphs", "peps", "NamedGraph", "Topology", "build_topology", "concat_topologies",
           "forest_cover_edge_sequence", "heavy_hex_lattice", "named_comb_tree", "named_grid",
           "named_path_graph"]
