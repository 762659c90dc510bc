"""Engine: one device context holding a network (or a batch), its site tensors and
its messages, driven through the C ABI.

This is the plumbing under the reference-shaped API in `beliefpropagation.py`; every
call lands in HIP kernels of libbpmi355x.so.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass
from typing import Optional

import numpy as np

from . import _lib as L
from .graphs import Topology


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ip(a):
    return a.ctypes.data_as(C.POINTER(C.c_int)) if a is not None else None


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double)) if a is not None else None


@dataclass
class RunResult:
    sweeps: np.ndarray       # per network
    delta: np.ndarray        # last delta per network
    at_iteration: np.ndarray  # converged-at iteration per network (-1: never)
    ms: float                # device time of the run (HIP events)

    @property
    def converged(self) -> np.ndarray:
        return self.at_iteration >= 0


class Engine:
    """Device-resident BP state for one topology.

    dtype: "f64" | "f32" | "bf16" (contraction precision) | "c128" (complex FP64: sites,
    messages and scalars are complex128 arrays).  device: HIP ordinal.
    """

    def __init__(self, topo: Topology, sites=None, dtype: str = "f64", device: int = 0,
                 path: str = "auto"):
        self.lib = L.load()
        self.topo = topo
        self._run_bufs = None
        self.dtype = dtype
        self.cplx = L.DTYPES[dtype] == L.BP_C128
        self.device = int(device)
        h = C.c_void_p()
        L.check(self.lib.bp_create(C.byref(h), int(device), L.DTYPES[dtype]), "bp_create")
        self.h = h
        self._keep = {k: _i32(getattr(topo, k)) for k in
                      ("port_off", "port_nbr", "port_edge", "port_rev", "edge_dim", "phys_dim")}
        self._keep["vnet"] = _i32(topo.vertex_net)
        k = self._keep
        L.check(self.lib.bp_set_topology(self.h, topo.nv, topo.ne, _ip(k["port_off"]),
                                         _ip(k["port_nbr"]), _ip(k["port_edge"]),
                                         _ip(k["port_rev"]), _ip(k["edge_dim"]),
                                         _ip(k["phys_dim"]), _ip(k["vnet"]), int(topo.nnet),
                                         int(topo.plain)), "bp_set_topology")
        self.set_path(path)
        if sites is not None:
            self.set_sites(sites)

    # -- lifecycle
    def close(self):
        if getattr(self, "h", None):
            self.lib.bp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # -- data
    def _host(self, a):
        """A host buffer in the context's element type: float64, or complex128 passed to
        the library as interleaved (re, im) doubles."""
        if self.cplx:
            return np.ascontiguousarray(a, dtype=np.complex128).reshape(-1)
        if np.iscomplexobj(a):
            raise TypeError(f"complex data for a {self.dtype} context (use dtype='c128')")
        return _f64(a).reshape(-1)

    def _zeros(self, n):
        return np.zeros(n, dtype=np.complex128 if self.cplx else np.float64)

    @staticmethod
    def _ptr(a):
        return _dp(a.view(np.float64)) if np.iscomplexobj(a) else _dp(a)

    def set_sites(self, sites) -> None:
        s = self._host(sites)
        if s.size != int(self.topo.site_off[-1]):
            raise ValueError(f"sites has {s.size} entries, topology needs {self.topo.site_off[-1]}")
        L.check(self.lib.bp_upload_sites(self.h, self._ptr(s)), "bp_upload_sites")

    def set_messages(self, msgs) -> None:
        m = self._host(msgs)
        if m.size != int(self.topo.msg_off[-1]):
            raise ValueError("message buffer size does not match the topology")
        L.check(self.lib.bp_init_messages(self.h, L.BP_INIT_UPLOAD, self._ptr(m)),
                "bp_init_messages")

    def init_messages(self, kind: str = "identity") -> None:
        code = {"identity": L.BP_INIT_IDENTITY, "one": L.BP_INIT_IDENTITY,
                "ones": L.BP_INIT_ONES}[kind]
        L.check(self.lib.bp_init_messages(self.h, code, None), "bp_init_messages")

    def set_slot_messages(self, slots, data) -> None:
        """Write the messages of `slots` (packed FP64, slot-list order)."""
        sl = _i32(slots)
        d = self._host(data)
        need = int(sum(int(self.topo.msg_off[q + 1] - self.topo.msg_off[q]) for q in sl))
        if d.size != need:
            raise ValueError(f"{d.size} values for slots holding {need}")
        L.check(self.lib.bp_set_slot_messages(self.h, len(sl), _ip(sl), self._ptr(d)),
                "bp_set_slot_messages")

    def slot_messages(self, slots) -> np.ndarray:
        """The messages of `slots`, packed FP64 in slot-list order."""
        sl = _i32(slots)
        need = int(sum(int(self.topo.msg_off[q + 1] - self.topo.msg_off[q]) for q in sl))
        out = self._zeros(need)
        L.check(self.lib.bp_get_slot_messages(self.h, len(sl), _ip(sl), self._ptr(out)),
                "bp_get_slot_messages")
        return out

    def messages(self) -> np.ndarray:
        out = self._zeros(int(self.topo.msg_off[-1]))
        L.check(self.lib.bp_download_messages(self.h, self._ptr(out)), "bp_download_messages")
        return out

    # -- path
    def set_path(self, path: str) -> None:
        L.check(self.lib.bp_set_path(self.h, L.PATHS[path]), "bp_set_path")

    def path(self) -> str:
        p = C.c_int(0)
        L.check(self.lib.bp_get_path(self.h, C.byref(p)), "bp_get_path")
        return {L.BP_PATH_GENERIC: "generic", L.BP_PATH_VERTEX: "vertex",
                L.BP_PATH_MIXED: "mixed"}[p.value]

    # -- BP
    def sweep(self) -> None:
        L.check(self.lib.bp_sweep(self.h), "bp_sweep")

    def run(self, maxiter: Optional[int] = None, tol: Optional[float] = None,
            criterion: str = "fidelity", use_graph: bool = True) -> RunResult:
        # the output buffers and their ctypes pointers are made once per engine (building
        # the pointers costs about 15 us a call, 2% of a 20-sweep run at 32x32 D=8); the
        # result holds copies
        if self._run_bufs is None:
            nn = int(self.topo.nnet)
            sw = np.zeros(nn, dtype=np.int32)
            dl = np.zeros(nn)
            at = np.zeros(nn, dtype=np.int32)
            ms = C.c_double(0)
            self._run_bufs = (sw, dl, at, _ip(sw), _dp(dl), _ip(at), ms, C.byref(ms))
        sw, dl, at, psw, pdl, pat, ms, pms = self._run_bufs
        crit = L.BP_CRIT_FIDELITY if criterion == "fidelity" else L.BP_CRIT_MAXABS
        L.check(self.lib.bp_run(self.h, int(maxiter) if maxiter is not None else 0,
                                float(tol) if tol is not None else -1.0, crit,
                                int(bool(use_graph)), psw, pdl, pat), "bp_run")
        L.check(self.lib.bp_last_run_ms(self.h, pms), "bp_last_run_ms")
        return RunResult(sw.copy(), dl.copy(), at.copy(), ms.value)

    def prepare_run(self, maxiter: int) -> None:
        """Capture the sweep graphs that `run(maxiter, use_graph=True)` replays (chunks of 16
        and the last, shorter chunk, from either message buffer) without running a sweep."""
        L.check(self.lib.bp_prepare_run(self.h, int(maxiter)), "bp_prepare_run")

    def run_sequential(self, order, maxiter: Optional[int] = None, tol: Optional[float] = None,
                       criterion: str = "fidelity") -> RunResult:
        """Gauss-Seidel sweeps over `order` (message slots), reference semantics."""
        o = _i32(order)
        sw = np.zeros(1, dtype=np.int32)
        dl = np.zeros(1)
        at = np.zeros(1, dtype=np.int32)
        crit = L.BP_CRIT_FIDELITY if criterion == "fidelity" else L.BP_CRIT_MAXABS
        L.check(self.lib.bp_run_sequential(self.h, _ip(o), len(o),
                                           int(maxiter) if maxiter is not None else 0,
                                           float(tol) if tol is not None else -1.0, crit,
                                           _ip(sw), _dp(dl), _ip(at)), "bp_run_sequential")
        ms = C.c_double(0)
        L.check(self.lib.bp_last_run_ms(self.h, C.byref(ms)), "bp_last_run_ms")
        return RunResult(sw, dl, at, ms.value)

    def time_sweeps(self, n: int) -> float:
        """Average device milliseconds of one sweep's kernels over n back-to-back sweeps."""
        ms = C.c_double(0)
        L.check(self.lib.bp_time_sweeps(self.h, int(n), C.byref(ms)), "bp_time_sweeps")
        return ms.value

    def update_slot(self, q: int) -> None:
        L.check(self.lib.bp_update_slot(self.h, int(q)), "bp_update_slot")

    def residuals(self):
        fid = np.zeros(self.topo.nports)
        mab = np.zeros(self.topo.nports)
        L.check(self.lib.bp_residuals(self.h, _dp(fid), _dp(mab)), "bp_residuals")
        return fid, mab

    def iterate_diff(self, other) -> float:
        o = self._host(other)
        if o.size != int(self.topo.msg_off[-1]):
            raise ValueError("message buffer size does not match the topology")
        out = C.c_double(0)
        L.check(self.lib.bp_iterate_diff(self.h, self._ptr(o), C.byref(out)), "bp_iterate_diff")
        return out.value

    # -- scalars
    def vertex_scalars(self) -> np.ndarray:
        out = self._zeros(self.topo.nv)
        L.check(self.lib.bp_vertex_scalars(self.h, self._ptr(out)), "bp_vertex_scalars")
        return out

    def edge_scalars(self) -> np.ndarray:
        out = self._zeros(max(self.topo.ne, 1))
        L.check(self.lib.bp_edge_scalars(self.h, self._ptr(out)), "bp_edge_scalars")
        return out[: self.topo.ne]

    def lognorms(self) -> np.ndarray:
        nn = int(self.topo.nnet)
        re = np.zeros(nn)
        im = np.zeros(nn)
        L.check(self.lib.bp_lognorm(self.h, _dp(re), _dp(im)), "bp_lognorm")
        if self.cplx:  # complex element types give a complex log-norm
            return re + 1j * im
        return np.where(im != 0.0, re + 1j * im, re) if np.any(im != 0.0) else re

    def lognorm(self):
        v = self.lognorms()[0]
        return complex(v) if isinstance(v, complex) or np.iscomplexobj(v) else float(v)

    def polish_fp64(self, nsweeps: int = 3) -> np.ndarray:
        out = np.zeros(int(self.topo.nnet))
        L.check(self.lib.bp_polish_fp64(self.h, int(nsweeps), _dp(out)), "bp_polish_fp64")
        return out

    # -- gradients
    def expect_vertex(self, ops) -> np.ndarray:
        """<O_v>_BP for every vertex at the current messages; ops: (nv, ld, ld) real
        operators on the physical index (top-left d x d block used)."""
        o = _f64(ops)
        if o.ndim != 3 or o.shape[0] != self.topo.nv or o.shape[1] != o.shape[2]:
            raise ValueError("ops must have shape (nv, ld, ld)")
        out = np.zeros(self.topo.nv)
        L.check(self.lib.bp_expect_vertex(self.h, _dp(o), int(o.shape[1]), _dp(out)),
                "bp_expect_vertex")
        return out

    def expect_edge(self, ops_first, ops_second) -> np.ndarray:
        """<A_e B_e>_BP for every topology edge e: ops_first[e] on the owner of the edge's
        lower message slot, ops_second[e] on the other end; shapes (ne, ld, ld)."""
        a, b = _f64(ops_first), _f64(ops_second)
        ne = self.topo.ne
        if a.shape != b.shape or a.ndim != 3 or a.shape[0] != ne or a.shape[1] != a.shape[2]:
            raise ValueError("ops must have shape (ne, ld, ld)")
        out = np.zeros(ne)
        L.check(self.lib.bp_expect_edge(self.h, _dp(a), _dp(b), int(a.shape[1]), _dp(out)),
                "bp_expect_edge")
        return out

    def last_run_ms(self) -> float:
        """Device time of the last timed call (bp_last_run_ms)."""
        ms = C.c_double()
        L.check(self.lib.bp_last_run_ms(self.h, C.byref(ms)), "bp_last_run_ms")
        return float(ms.value)

    def sites(self) -> np.ndarray:
        """The FP64 site tensors (site layout), as held on the device."""
        out = self._zeros(int(self.topo.site_off[-1]))
        L.check(self.lib.bp_download_sites(self.h, self._ptr(out)), "bp_download_sites")
        return out

    def apply_gates(self, kinds, wheres, gates, normalize: bool = False, rtol: float = -1.0,
                    maxdim: int = 0):
        """Gates on the resident network (bp_apply_gates_resident): kinds[g] 2 with
        wheres[g] a message slot (site 1 its owner), or 1 with a vertex index; gates[g] the
        matrix [(t1 t2)][(s1 s2)] (or [t][s]).  Returns (chi per gate, S as rows padded to
        the largest gated bond dimension).  Bonds keep their dimension; at most
        min(bond dimension, maxdim) singular values are kept (maxdim <= 0: no cap)."""
        n = len(kinds)
        k = _i32(kinds)
        w = _i32(wheres)
        g = _f64(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x in gates])
                 if n else np.zeros(1))
        stride = 1
        for kk, ww in zip(kinds, wheres):
            if kk == 2:
                stride = max(stride, int(self.topo.edge_dim[self.topo.port_edge[ww]]))
        chi = np.zeros(max(n, 1), dtype=np.int32)
        S = np.zeros((max(n, 1), stride))
        L.check(self.lib.bp_apply_gates_resident(self.h, n, _ip(k), _ip(w), _dp(g),
                                                 int(maxdim), int(normalize), float(rtol),
                                                 _ip(chi), _dp(S), stride),
                "bp_apply_gates_resident")
        return chi[:n], S[:n]

    def set_option(self, name: str, value: int) -> None:
        """Kernel options (bp_set_option): "vertex_variant" (D <= 8 sweep kernel, 0..5; 2..5
        FP32 only) and
        "vertex_bond_mask" (bit D lets networks of largest bond dimension D use the vertex
        kernels)."""
        L.check(self.lib.bp_set_option(self.h, name.encode(), int(value)), "bp_set_option")

    def get_option(self, name: str) -> int:
        """The current value of a kernel option (bp_get_option): after autotune, the tuned
        choice."""
        v = C.c_int(0)
        L.check(self.lib.bp_get_option(self.h, name.encode(), C.byref(v)), "bp_get_option")
        return v.value

    def autotune(self, cache: str | None = None, mode: str = "missing") -> int:
        """Time the kernel choices on this engine's data and apply the fastest (see
        bp_autotune in include/bp_mi355x.h).  cache: JSON file of earlier decisions (read,
        and rewritten unless mode == "cached").  mode: "cached" (apply the file only),
        "missing" (tune what the file lacks) or "all" (retune).  Returns the number of
        decisions timed; the messages are left as they were."""
        modes = {"cached": 0, "missing": 1, "all": 2}
        n = C.c_int(0)
        path = None if cache is None else os.fsencode(cache)
        L.check(self.lib.bp_autotune(self.h, path, modes[mode], C.byref(n)), "bp_autotune")
        return n.value

    def _device_grad_buffer(self):
        return L.DeviceArray(int(self.topo.site_off[-1]),
                             np.float64 if self.dtype == "f64" else np.float32, self.device)

    def grad_lognorm(self, device: bool = False, out=None):
        """d log Z_BP / dA at the current fixed point: FP64 numpy array, or with
        device=True a DeviceArray on the context's GPU in the arithmetic type (float64 for
        f64 engines, float32 otherwise) that stays in device memory."""
        if device:
            out = out if out is not None else self._device_grad_buffer()
            L.check(self.lib.bp_grad_lognorm_device(self.h, C.c_void_p(out.ptr)),
                    "bp_grad_lognorm_device")
            return out
        out = np.zeros(int(self.topo.site_off[-1]))
        L.check(self.lib.bp_grad_lognorm(self.h, _dp(out)), "bp_grad_lognorm")
        return out

    def local_energy(self, ops_v=None, ops_first=None, ops_second=None, grad: bool = False,
                     adj_tol: float = 1e-8, adj_maxiter: int = 1000, device: bool = False,
                     out=None):
        """E = sum_v <O_v>_BP + sum_e <A_e B_e>_BP with operator arrays as expect_vertex /
        expect_edge take (either sum may be omitted); with grad=True also (gradient,
        adjoint sweeps, adjoint residual), on the device with device=True."""
        if ops_v is None and (ops_first is None or ops_second is None):
            raise ValueError("give vertex operators, edge operator pairs, or both")
        ov = _f64(ops_v) if ops_v is not None else None
        oa = _f64(ops_first) if ops_first is not None else None
        ob = _f64(ops_second) if ops_second is not None else None
        ref = ov if ov is not None else oa
        ld = int(ref.shape[-1])
        if ov is not None and ov.shape != (self.topo.nv, ld, ld):
            raise ValueError("ops_v must have shape (nv, ld, ld)")
        if oa is not None and (oa.shape != (self.topo.ne, ld, ld) or ob is None or
                               ob.shape != oa.shape):
            raise ValueError("edge operators must both have shape (ne, ld, ld)")
        E = C.c_double(0)
        it = C.c_int(0)
        res = C.c_double(0)
        if grad and device:
            g = out if out is not None else self._device_grad_buffer()
            L.check(self.lib.bp_local_energy_device(self.h, _dp(ov), _dp(oa), _dp(ob), ld,
                                                    C.byref(E), C.c_void_p(g.ptr),
                                                    float(adj_tol), int(adj_maxiter),
                                                    C.byref(it), C.byref(res)),
                    "bp_local_energy_device")
            return E.value, g, it.value, res.value
        g = np.zeros(int(self.topo.site_off[-1])) if grad else None
        L.check(self.lib.bp_local_energy(self.h, _dp(ov), _dp(oa), _dp(ob), ld, C.byref(E),
                                         _dp(g), float(adj_tol), int(adj_maxiter),
                                         C.byref(it), C.byref(res)), "bp_local_energy")
        if grad:
            return E.value, g, it.value, res.value
        return E.value

    def tfim_energy(self, J: float = 1.0, h: float = 3.0, grad: bool = False,
                    adj_tol: float = 1e-8, adj_maxiter: int = 1000, device: bool = False,
                    out=None):
        """BP energy of H = -J sum ZZ - h sum X; with grad=True also (gradient, adjoint
        sweeps, adjoint residual).  device=True returns the gradient as a DeviceArray (see
        grad_lognorm); `out` reuses a DeviceArray from an earlier call."""
        E = C.c_double(0)
        it = C.c_int(0)
        res = C.c_double(0)
        if grad and device:
            g = out if out is not None else self._device_grad_buffer()
            L.check(self.lib.bp_tfim_energy_device(self.h, float(J), float(h), C.byref(E),
                                                   C.c_void_p(g.ptr), float(adj_tol),
                                                   int(adj_maxiter), C.byref(it),
                                                   C.byref(res)), "bp_tfim_energy_device")
            return E.value, g, it.value, res.value
        g = np.zeros(int(self.topo.site_off[-1])) if grad else None
        L.check(self.lib.bp_tfim_energy(self.h, float(J), float(h), C.byref(E), _dp(g),
                                        float(adj_tol), int(adj_maxiter), C.byref(it),
                                        C.byref(res)), "bp_tfim_energy")
        if grad:
            return E.value, g, it.value, res.value
        return E.value
