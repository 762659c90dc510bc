"""A tensor network kept on the device between BP runs and gate layers.

The reference marshals nothing: its ITensorNetwork and MessageCache are the data its BP and
gate application work on (tensornetwork.jl:138-173, messagecache.jl:205-229,
apply_operators.jl:167-283).  Here a network is flattened to the engine's CSR port layout
once (`DeviceNetwork(tn)`), and then

  * `beliefpropagation(...)` runs BP on the device-resident sites and messages;
  * `apply_operators(gates, ...)` applies BP-gauged gates in place on the device
    (bp_apply_gates_resident): no site tensor or message crosses PCIe, so a simple-update
    loop (BP, gate layer, BP, ...) stays on the GPU;
  * `itensornetwork()` / `messagecache()` export the current state on demand.

The flattened layout is fixed: every bond keeps its dimension.  A two-site gate keeps at
most that many singular values (`trunc` -- truncrank, trunctol -- may lower the count
further) and zero-pads the rest, which represents the same state; the exported network
carries the original link indices.  Gate matrices are the only per-layer upload.
"""
from __future__ import annotations

from typing import Hashable, List

import numpy as np

from .algorithms import ArgumentError
from .apply import Operator, _trunc_params
from .engine import Engine, RunResult
from .itensors import ITensor, ITensorNetwork, normnetwork, to_topology
from .messagecache import MessageCache


class DeviceNetwork:
    """An ITensorNetwork (the ket of its NormNetwork) resident on the device with its BP
    messages.  dtype: the BP arithmetic ("f64", "f32", "bf16"); gates always run in FP64
    on the FP64 master copy of the site tensors."""

    def __init__(self, tn: ITensorNetwork, messages: MessageCache | None = None,
                 dtype: str = "f64", device: int = 0, path: str = "auto"):
        self.ket = tn.copy()
        self.flat = to_topology(normnetwork(self.ket))
        self.topo = self.flat.topo
        self.engine = Engine(self.topo, self.flat.sites, dtype=dtype, device=device, path=path)
        if messages is None:
            self.engine.init_messages("identity")
        else:
            self.engine.set_messages(messages.to_flat(self.topo))
        self._vidx = {v: i for i, v in enumerate(self.topo.vertices)}
        self._site_of = {}
        for v in self.topo.vertices:
            for i in self.ket.siteinds(v):
                self._site_of[i] = v

    # -- BP
    def beliefpropagation(self, maxiter: int | None = None, tol: float | None = None,
                          criterion: str = "fidelity") -> RunResult:
        return self.engine.run(maxiter=maxiter, tol=tol, criterion=criterion)

    # -- gates
    def _slot(self, u: int, w: int) -> int:
        t = self.topo
        for q in range(t.port_off[u], t.port_off[u + 1]):
            if int(t.port_nbr[q]) == w:
                return q
        raise ArgumentError("a two-site gate needs neighbouring vertices")

    def _gate(self, op: Operator):
        vs: List[Hashable] = []
        for i in op.domain:
            if i not in self._site_of:
                raise ArgumentError("operator shares no indices with the tensor network")
            v = self._site_of[i]
            if v not in vs:
                vs.append(v)
        if len(vs) > 2:
            raise ArgumentError(f"{len(vs)}-site gate decomposition not implemented")
        doms = []
        for v in vs:
            sites = list(self.ket.siteinds(v))
            if not set(sites) <= set(op.domain):
                raise ArgumentError(f"the gate acts on part of vertex {v!r}'s site indices")
            doms += sites
        mat = op.matrix(doms)
        if len(vs) == 1:
            return 1, self._vidx[vs[0]], mat, vs
        u1, u2 = self._vidx[vs[0]], self._vidx[vs[1]]
        return 2, self._slot(u1, u2), mat, vs

    def apply_operators(self, operators, trunc=None, normalize: bool = False) -> List[np.ndarray]:
        """Apply the gates in order, in place on the device; runs of gates on disjoint
        vertices go in one call.  Returns the singular values of every two-site gate (None
        for one-site gates)."""
        maxdim, rtol = _trunc_params(trunc)
        out: List[np.ndarray] = []
        batch, used = [], set()

        def flush():
            if not batch:
                return
            chi, S = self.engine.apply_gates([b[0] for b in batch], [b[1] for b in batch],
                                             [b[2] for b in batch], normalize, rtol, maxdim)
            for b, c, s in zip(batch, chi, S):
                out.append(s[:c].copy() if b[0] == 2 else None)
            batch.clear()
            used.clear()

        for op in operators:
            kind, where, mat, vs = self._gate(op)
            if used.intersection(vs):
                flush()
            batch.append((kind, where, mat))
            used.update(vs)
        flush()
        return out

    def apply_operator(self, operator, trunc=None, normalize: bool = False):
        return self.apply_operators([operator], trunc=trunc, normalize=normalize)[0]

    # -- export
    def messagecache(self) -> MessageCache:
        return MessageCache.from_flat(self.topo, self.engine.messages())

    def itensornetwork(self) -> ITensorNetwork:
        """The current ket network: the original indices, the device's site tensors."""
        sites = self.engine.sites()
        t, fl = self.topo, self.flat
        out = {}
        for ui, v in enumerate(t.vertices):
            inds = list(self.ket.siteinds(v))
            for q in range(t.port_off[ui], t.port_off[ui + 1]):
                w = t.vertices[t.port_nbr[q]]
                inds += list(fl.port_inds[(v, w)])
            a = sites[t.site_off[ui]:t.site_off[ui + 1]]
            out[v] = ITensor(a.reshape([i.dim for i in inds]), inds)
        return ITensorNetwork(out)

    def close(self):
        self.engine.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
