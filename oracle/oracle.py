"""ctypes front end of the FP64 CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by the test suite, the smoke check and the
benchmark's CPU-baseline leg for reference values; the product package never imports it.
Each wrapper names the reference function it restates (see bp_oracle.c for lines).

`topo` arguments are duck-typed: any object with the flat CSR arrays of the engine's
topology (port_off, port_nbr, port_edge, port_rev, edge_dim, phys_dim, site_off,
msg_off, nv, ne, plain).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (gcc); returns the library path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE, "liboracle.so"],
                       check=True, stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


class _Graph(C.Structure):
    _fields_ = [("nv", C.c_int), ("ne", C.c_int),
                ("port_off", C.POINTER(C.c_int)), ("port_nbr", C.POINTER(C.c_int)),
                ("port_edge", C.POINTER(C.c_int)), ("port_rev", C.POINTER(C.c_int)),
                ("edge_dim", C.POINTER(C.c_int)), ("phys_dim", C.POINTER(C.c_int)),
                ("site_off", C.POINTER(C.c_longlong)), ("msg_off", C.POINTER(C.c_longlong))]


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


def _declare(L):
    G = C.POINTER(_Graph)
    L.orc_norm_message.argtypes = [G, _dp, _dp, C.c_int, _dp]
    L.orc_plain_message.argtypes = [G, _dp, _dp, C.c_int, _dp]
    L.orc_sweep_jacobi.argtypes = [G, _dp, _dp, _dp, C.c_int, C.c_int, C.c_int, _dp, _dp]
    L.orc_sweep_sequential.argtypes = [G, _dp, _dp, C.c_int, C.c_int, _ip, C.c_int]
    L.orc_bp.argtypes = [G, _dp, _dp, C.c_int, C.c_int, _ip, C.c_int, C.c_int, C.c_double,
                         C.c_int, C.c_int, _ip, _dp, _ip]
    L.orc_iterate_diff.argtypes = [G, _dp, _dp, C.c_int]
    L.orc_iterate_diff.restype = C.c_double
    L.orc_vertex_scalars.argtypes = [G, _dp, _dp, C.c_int, _dp]
    L.orc_edge_scalars.argtypes = [G, _dp, C.c_int, _dp]
    L.orc_bethe_free_energy.argtypes = [G, _dp, _dp, C.c_int, _dp, _dp]
    L.orc_contract_labeled.argtypes = [C.c_int, _ip, _ip, _ip, _ip, _dp,
                                       C.POINTER(C.c_longlong), C.c_longlong, _ip, _ip, _dp,
                                       C.c_longlong]
    L.orc_norm_vjp.argtypes = [G, _dp, _dp, _dp, _dp, _dp]
    L.orc_set_num_threads.argtypes = [C.c_int]
    L.orc_set_num_threads.restype = None
    L.orc_grad_lognorm_explicit.argtypes = [G, _dp, _dp, _dp]
    L.orc_tfim_energy.argtypes = [G, _dp, _dp, C.c_double, C.c_double, _dp, _dp, _dp]
    L.orc_expect_vertex.argtypes = [G, _dp, _dp, _dp, C.c_int, _dp]
    L.orc_expect_edge.argtypes = [G, _dp, _dp, _dp, _dp, C.c_int, _dp]
    L.orc_local_energy.argtypes = [G, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp, _dp]
    L.orc_fixed_point_gradient.argtypes = [G, _dp, _dp, _dp, _dp, C.c_int, C.c_double, _dp,
                                           _ip, _dp]
    for name in ("orc_norm_message", "orc_plain_message", "orc_sweep_jacobi",
                 "orc_sweep_sequential", "orc_bp", "orc_vertex_scalars", "orc_edge_scalars",
                 "orc_bethe_free_energy", "orc_contract_labeled", "orc_norm_vjp",
                 "orc_grad_lognorm_explicit", "orc_tfim_energy", "orc_fixed_point_gradient",
              "orc_expect_vertex", "orc_expect_edge", "orc_local_energy"):
        getattr(L, name).restype = C.c_int


def _p(a, ct=C.c_double):
    return a.ctypes.data_as(C.POINTER(ct)) if a is not None else None


class OracleGraph:
    """Keeps the int arrays alive and exposes the C struct."""

    def __init__(self, topo):
        self.topo = topo
        self._arrs = {k: np.ascontiguousarray(getattr(topo, k), dtype=np.int32)
                      for k in ("port_off", "port_nbr", "port_edge", "port_rev", "edge_dim",
                                "phys_dim")}
        self._arrs["site_off"] = np.ascontiguousarray(topo.site_off, dtype=np.int64)
        self._arrs["msg_off"] = np.ascontiguousarray(topo.msg_off, dtype=np.int64)
        a = self._arrs
        self.s = _Graph(int(topo.nv), int(topo.ne), _p(a["port_off"], C.c_int),
                        _p(a["port_nbr"], C.c_int), _p(a["port_edge"], C.c_int),
                        _p(a["port_rev"], C.c_int), _p(a["edge_dim"], C.c_int),
                        _p(a["phys_dim"], C.c_int), _p(a["site_off"], C.c_longlong),
                        _p(a["msg_off"], C.c_longlong))
        self.plain = int(bool(getattr(topo, "plain", False)))

    @property
    def ref(self):
        return C.byref(self.s)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _check(rc, what):
    if rc != 0:
        raise RuntimeError(f"oracle {what} failed with code {rc}")


def message(topo, sites, msgs, q):
    """Unnormalised message for slot q (SimpleMessageUpdate before the division)."""
    g = OracleGraph(topo)
    sites, msgs = _f64(sites), _f64(msgs)
    n = int(topo.msg_off[q + 1] - topo.msg_off[q])
    out = np.zeros(n)
    f = lib().orc_plain_message if g.plain else lib().orc_norm_message
    _check(f(g.ref, _p(sites), _p(msgs), int(q), _p(out)), "message")
    return out


def sweep_jacobi(topo, sites, msgs, normalize=True, nthreads=1):
    g = OracleGraph(topo)
    sites, msgs = _f64(sites), _f64(msgs)
    new = np.zeros_like(msgs)
    fid = np.zeros(topo.nports)
    mab = np.zeros(topo.nports)
    _check(lib().orc_sweep_jacobi(g.ref, _p(sites), _p(msgs), _p(new), g.plain,
                                  int(normalize), int(nthreads), _p(fid), _p(mab)), "sweep")
    return new, fid, mab


def sweep_sequential(topo, sites, msgs, order, normalize=True):
    g = OracleGraph(topo)
    sites = _f64(sites)
    out = _f64(msgs).copy()
    order = np.ascontiguousarray(order, dtype=np.int32)
    _check(lib().orc_sweep_sequential(g.ref, _p(sites), _p(out), g.plain, int(normalize),
                                      _p(order, C.c_int), len(order)), "sweep_sequential")
    return out


def bp(topo, sites, msgs, maxiter=None, tol=None, schedule="jacobi", order=None,
       criterion="fidelity", nthreads=1):
    """beliefpropagation(...) restated: returns (messages, sweeps, delta, at_iteration)."""
    g = OracleGraph(topo)
    sites = _f64(sites)
    out = _f64(msgs).copy()
    order_arr = np.ascontiguousarray(order if order is not None else [], dtype=np.int32)
    sw = C.c_int(0)
    dl = C.c_double(0)
    at = C.c_int(0)
    _check(lib().orc_bp(g.ref, _p(sites), _p(out), g.plain, 0 if schedule == "jacobi" else 1,
                        _p(order_arr, C.c_int), len(order_arr),
                        int(maxiter) if maxiter is not None else 0,
                        float(tol) if tol is not None else -1.0,
                        0 if criterion == "fidelity" else 1, int(nthreads),
                        C.byref(sw), C.byref(dl), C.byref(at)), "bp")
    return out, sw.value, dl.value, at.value


def iterate_diff(topo, m1, m2):
    g = OracleGraph(topo)
    return lib().orc_iterate_diff(g.ref, _p(_f64(m1)), _p(_f64(m2)), g.plain)


def vertex_scalars(topo, sites, msgs):
    g = OracleGraph(topo)
    out = np.zeros(topo.nv)
    _check(lib().orc_vertex_scalars(g.ref, _p(_f64(sites)), _p(_f64(msgs)), g.plain, _p(out)),
           "vertex_scalars")
    return out


def edge_scalars(topo, msgs):
    g = OracleGraph(topo)
    out = np.zeros(topo.ne)
    _check(lib().orc_edge_scalars(g.ref, _p(_f64(msgs)), g.plain, _p(out)), "edge_scalars")
    return out


def bethe_free_energy(topo, sites, msgs):
    g = OracleGraph(topo)
    re = C.c_double(0)
    im = C.c_double(0)
    _check(lib().orc_bethe_free_energy(g.ref, _p(_f64(sites)), _p(_f64(msgs)), g.plain,
                                       C.byref(re), C.byref(im)), "bethe")
    return complex(re.value, im.value) if im.value != 0.0 else re.value


def contract_labeled(tensors, labels, max_elems=1 << 26):
    """Exact contraction of labelled tensors (contract_network restated).  `labels[t]` is a
    tuple of hashable labels for tensors[t]; returns (array, open_labels)."""
    names = {}
    for ls in labels:
        for l in ls:
            names.setdefault(l, len(names))
    label_dim = np.zeros(len(names), dtype=np.int32)
    for t, ls in zip(tensors, labels):
        t = np.asarray(t)
        if t.ndim != len(ls):
            raise ValueError("rank / label mismatch")
        for d, l in zip(t.shape, ls):
            if label_dim[names[l]] and label_dim[names[l]] != d:
                raise ValueError(f"dimension mismatch on label {l}")
            label_dim[names[l]] = d
    ranks = np.array([len(ls) for ls in labels], dtype=np.int32)
    lab_off = np.zeros(len(labels), dtype=np.int32)
    lab_off[1:] = np.cumsum(ranks)[:-1]
    flat_labels = np.array([names[l] for ls in labels for l in ls] or [0], dtype=np.int32)
    data = np.concatenate([_f64(t).ravel() for t in tensors])
    data_off = np.zeros(len(tensors), dtype=np.int64)
    data_off[1:] = np.cumsum([np.asarray(t).size for t in tensors])[:-1]
    counts = {}
    for ls in labels:
        for l in ls:
            counts[l] = counts.get(l, 0) + 1
    open_labels = [l for l in names if counts[l] == 1]
    cap = int(np.prod([label_dim[names[l]] for l in open_labels])) if open_labels else 1
    out = np.zeros(cap)
    orank = C.c_int(0)
    olabs = np.zeros(64, dtype=np.int32)
    _check(lib().orc_contract_labeled(len(tensors), _p(ranks, C.c_int), _p(lab_off, C.c_int),
                                      _p(flat_labels, C.c_int), _p(label_dim, C.c_int),
                                      _p(data), _p(data_off, C.c_longlong), int(max_elems),
                                      C.byref(orank), _p(olabs, C.c_int), _p(out), cap),
           "contract_labeled")
    inv = {v: k for k, v in names.items()}
    olist = [inv[int(i)] for i in olabs[:orank.value]]
    shape = tuple(int(label_dim[names[l]]) for l in olist)
    return out[:cap].reshape(shape) if shape else out[0], olist


def norm_exact(topo, sites):
    """<psi|psi> of a PEPS by exact contraction of the double layer (ket tensors and
    conj bra tensors, sharing only the site index), absorbed vertex by vertex."""
    tensors, labels = [], []
    for u in range(topo.nv):
        shp = topo.site_shape(u)
        A = _f64(sites[topo.site_off[u]:topo.site_off[u + 1]]).reshape(shp)
        ports = range(topo.port_off[u], topo.port_off[u + 1])
        ket = (("s", u),) + tuple(("k", int(topo.port_edge[q])) for q in ports)
        bra = (("s", u),) + tuple(("b", int(topo.port_edge[q])) for q in ports)
        tensors += [A, A.conj()]
        labels += [ket, bra]
    val, _ = contract_labeled(tensors, labels)
    return float(val)


def plain_exact(topo, sites):
    """Full contraction of a plain (single-layer) network."""
    tensors, labels = [], []
    for u in range(topo.nv):
        shp = topo.site_shape(u)[1:]
        A = _f64(sites[topo.site_off[u]:topo.site_off[u + 1]]).reshape(shp)
        ports = range(topo.port_off[u], topo.port_off[u + 1])
        tensors.append(A)
        labels.append(tuple(("e", int(topo.port_edge[q])) for q in ports))
    val, _ = contract_labeled(tensors, labels)
    return float(val)


def set_num_threads(n):
    """Threads of norm_vjp / fixed_point_gradient (OpenMP over vertices)."""
    lib().orc_set_num_threads(int(n))


def norm_vjp(topo, sites, msgs, gbar, want_m=True, want_a=True, nthreads=1):
    set_num_threads(nthreads)
    g = OracleGraph(topo)
    mbar = np.zeros(int(topo.msg_off[-1])) if want_m else None
    abar = np.zeros(int(topo.site_off[-1])) if want_a else None
    _check(lib().orc_norm_vjp(g.ref, _p(_f64(sites)), _p(_f64(msgs)), _p(_f64(gbar)),
                              _p(mbar), _p(abar)), "norm_vjp")
    return mbar, abar


def grad_lognorm_explicit(topo, sites, msgs):
    g = OracleGraph(topo)
    out = np.zeros(int(topo.site_off[-1]))
    _check(lib().orc_grad_lognorm_explicit(g.ref, _p(_f64(sites)), _p(_f64(msgs)), _p(out)),
           "grad_lognorm")
    return out


def tfim_energy(topo, sites, msgs, J=1.0, h=3.0, partials=False):
    g = OracleGraph(topo)
    E = C.c_double(0)
    dm = np.zeros(int(topo.msg_off[-1])) if partials else None
    da = np.zeros(int(topo.site_off[-1])) if partials else None
    _check(lib().orc_tfim_energy(g.ref, _p(_f64(sites)), _p(_f64(msgs)), float(J), float(h),
                                 C.byref(E), _p(dm), _p(da)), "tfim_energy")
    return (E.value, dm, da) if partials else E.value


def expect_vertex(topo, sites, msgs, ops):
    """<O_v>_BP for every vertex; ops: (nv, ld, ld) operators on the physical index."""
    g = OracleGraph(topo)
    ops = _f64(ops)
    out = np.zeros(topo.nv)
    _check(lib().orc_expect_vertex(g.ref, _p(_f64(sites)), _p(_f64(msgs)), _p(ops),
                                   int(ops.shape[-1]), _p(out)), "expect_vertex")
    return out


def expect_edge(topo, sites, msgs, ops_first, ops_second):
    """<A_e B_e>_BP for every edge (topology edge id); ops_first[e] acts on the owner of
    the edge's lower port, ops_second[e] on the other end."""
    g = OracleGraph(topo)
    a, b = _f64(ops_first), _f64(ops_second)
    out = np.zeros(int(topo.ne))
    _check(lib().orc_expect_edge(g.ref, _p(_f64(sites)), _p(_f64(msgs)), _p(a), _p(b),
                                 int(a.shape[-1]), _p(out)), "expect_edge")
    return out


def local_energy(topo, sites, msgs, ops_v=None, ops_first=None, ops_second=None,
                 partials=False):
    """E = sum_v <O_v>_BP + sum_e <A_e B_e>_BP (either sum may be omitted), optionally with
    the explicit partials (dE/dm, dE/dA)."""
    g = OracleGraph(topo)
    ref = ops_v if ops_v is not None else ops_first
    ld = int(np.shape(ref)[-1])
    E = C.c_double(0)
    dm = np.zeros(int(topo.msg_off[-1])) if partials else None
    da = np.zeros(int(topo.site_off[-1])) if partials else None
    ov = _f64(ops_v) if ops_v is not None else None
    oa = _f64(ops_first) if ops_first is not None else None
    ob = _f64(ops_second) if ops_second is not None else None
    _check(lib().orc_local_energy(g.ref, _p(_f64(sites)), _p(_f64(msgs)), _p(ov), _p(oa), _p(ob),
                                  ld, C.byref(E), _p(dm), _p(da)), "local_energy")
    return (E.value, dm, da) if partials else E.value


def fixed_point_gradient(topo, sites, msgs, dL_dm, dL_dA, maxiter=1000, tol=1e-12, nthreads=1):
    set_num_threads(nthreads)
    g = OracleGraph(topo)
    grad = np.zeros(int(topo.site_off[-1]))
    it = C.c_int(0)
    res = C.c_double(0)
    _check(lib().orc_fixed_point_gradient(g.ref, _p(_f64(sites)), _p(_f64(msgs)),
                                          _p(_f64(dL_dm) if dL_dm is not None else None),
                                          _p(_f64(dL_dA) if dL_dA is not None else None),
                                          int(maxiter), float(tol), _p(grad), C.byref(it),
                                          C.byref(res)), "fixed_point_gradient")
    return grad, it.value, res.value


# ------------------------------------------------------------ gate application (numpy)
# BPApplyGate / apply_gate_bp_nsite! (src/apply/apply_operators.jl:214-283) restated with
# numpy's LAPACK routines in the roles of the external factorisations: the message square
# root X (X X^T = M) and its pseudo-inverse from numpy.linalg.eigh (gram_eigh_full_with_pinv,
# TensorAlgebra), the site factorisation from numpy.linalg.qr (qr_compact,
# MatrixAlgebraKit), the split from numpy.linalg.svd (svd_trunc).  Tensors are
# [phys][port 0]...[port k-1]; `msgs` maps a port to its incoming D x D message.  Pinned in
# test_oracle.py by the reference's gauge-invariant properties (test_apply_operator.jl):
# an untruncated gate reproduces the gated two-site tensor exactly, and on a tree with
# exact BP environments a truncated gate is the global optimal truncation.

def _gate_gauge(M):
    lam, V = np.linalg.eigh(0.5 * (M + M.T))
    lmax = lam.max() if lam.size else 0.0
    lam = np.where(lam > 1e-13 * lmax, lam, 0.0)
    X = V * np.sqrt(lam)
    inv = np.zeros_like(lam)
    nz = lam > 0
    inv[nz] = 1.0 / np.sqrt(lam[nz])
    return X, (V * inv).T


def _gate_mode(T, j, M):
    return np.moveaxis(np.tensordot(T, M, axes=([j + 1], [0])), -1, j + 1)


def apply_gate1(A, O, msgs=None, normalize=False):
    """One-site gate (:214-227): O [t][s] on the physical index; with normalize, divided by
    the norm of the result with every port's square root X absorbed."""
    B = np.tensordot(O, A, axes=([1], [0]))
    if normalize:
        C_ = B
        for j, M in (msgs or {}).items():
            C_ = _gate_mode(C_, j, _gate_gauge(M)[0])
        B = B / np.linalg.norm(C_)
    return B


def apply_gate2(A1, A2, p1, p2, msgs1, msgs2, G, maxdim=0, rtol=-1.0, normalize=False):
    """Two-site gate (:229-283).  G is [t1][t2][s1][s2] (any shape of d1*d2 x d1*d2 entries).
    Returns (new A1, new A2, S) with the shared port's dimension replaced by len(S)."""
    k1, k2 = A1.ndim - 1, A2.ndim - 1
    d1, d2 = A1.shape[0], A2.shape[0]
    Db = A1.shape[1 + p1]
    g1 = {j: _gate_gauge(M) for j, M in msgs1.items()}
    g2 = {j: _gate_gauge(M) for j, M in msgs2.items()}
    B1, B2 = A1, A2
    for j, (X, _) in g1.items():
        B1 = _gate_mode(B1, j, X)
    for j, (X, _) in g2.items():
        B2 = _gate_mode(B2, j, X)
    o1 = [j for j in range(k1) if j != p1]
    o2 = [j for j in range(k2) if j != p2]
    X1 = np.transpose(B1, [1 + j for j in o1] + [0, 1 + p1]).reshape(-1, d1 * Db)
    X2 = np.transpose(B2, [1 + j for j in o2] + [0, 1 + p2]).reshape(-1, d2 * Db)
    Q1, R1 = np.linalg.qr(X1)
    Q2, R2 = np.linalg.qr(X2)
    r1, r2 = R1.shape[0], R2.shape[0]
    T = np.einsum("iab,jcb->iajc", R1.reshape(r1, d1, Db), R2.reshape(r2, d2, Db))
    Tp = np.einsum("xyac,iajc->ixjy", np.asarray(G).reshape(d1, d2, d1, d2), T)
    U, S, Vt = np.linalg.svd(Tp.reshape(r1 * d1, r2 * d2), full_matrices=False)
    chi = len(S) if maxdim <= 0 else min(len(S), maxdim)
    if rtol >= 0:
        while chi > 1 and S[chi - 1] <= rtol * S[0]:
            chi -= 1
    U, S, Vt = U[:, :chi], S[:chi], Vt[:chi]
    if normalize:
        S = S / np.linalg.norm(S)
    N1 = (Q1 @ (U * np.sqrt(S)).reshape(r1, d1 * chi)).reshape(
        [A1.shape[1 + j] for j in o1] + [d1, chi])
    N2 = (Q2 @ (Vt.T * np.sqrt(S)).reshape(r2, d2 * chi)).reshape(
        [A2.shape[1 + j] for j in o2] + [d2, chi])

    def back(N, o, p, k):
        perm = [len(o)] + [0] * k
        for a, j in enumerate(o):
            perm[1 + j] = a
        perm[1 + p] = len(o) + 1
        return np.transpose(N, perm)

    A1n, A2n = back(N1, o1, p1, k1), back(N2, o2, p2, k2)
    for j, (_, Xi) in g1.items():
        A1n = _gate_mode(A1n, j, Xi)
    for j, (_, Xi) in g2.items():
        A2n = _gate_mode(A2n, j, Xi)
    return A1n, A2n, S
