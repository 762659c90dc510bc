"""Complex (ComplexF64) restatement of the BP path in numpy, the reference values for the
engine's complex contexts.

TEST INFRASTRUCTURE ONLY: imported by the test suite for reference values; the product package
never imports it.  The real FP64 oracle (bp_oracle.c) covers real element types; this file
restates the same functions for complex ones, which the reference runs for ComplexF64
(test/test_beliefpropagation.jl:154, test/test_normnetwork.jl:148):

  * message update, src/beliefpropagation/beliefpropagation.jl:242-257 with the double-layer
    factor of src/normnetwork.jl:49-54 (ket A, bra conj(A)):
        m_{u->v}[a, a'] = sum A[s, ..a.., X] conj(A[s, ..a'.., X']) prod_w m_{w->u}[x_w, x'_w]
    messages [ket][bra]; plain networks: m_{u->v}[a] = sum A[..a.., X] prod_w m_{w->u}[x_w];
    then division by the (complex) sum when it is nonzero;
  * iterate_diff, beliefpropagation.jl:261-267: max_e 1 - |<m1^_e, m2^_e>|^2, the first
    argument conjugated (LinearAlgebra.dot);
  * the loop of beliefpropagation (:69-92) under StopAfterIteration | StopWhenConverged
    (AlgorithmsInterfaceExtensions.jl:84-109): iteration 0 is skipped by the convergence
    test, as in bp_oracle.c's orc_bp;
  * vertex_scalar (messagecache.jl:139-143), edge_scalar (:153-157: both names contracted,
    no conjugate) and bethe_free_energy (:185-201: principal logarithms, -Inf on a zero
    edge term).

Every function loops over messages in Python and contracts with numpy.einsum: sized for
small networks (trees, lattices up to about 6x6 at D = 3).

`topo` is duck-typed like oracle.py's: port_off, port_nbr, port_edge, port_rev, edge_dim,
site_off, msg_off, nv, ne, plain.
"""
from __future__ import annotations

import string

import numpy as np

_L = string.ascii_letters


def _c(a):
    return np.ascontiguousarray(a, dtype=np.complex128).reshape(-1)


def _deg(topo, u):
    return int(topo.port_off[u + 1] - topo.port_off[u])


def _site(topo, sites, u):
    """A_u with dims (phys, D_0, ..., D_{k-1}) in CSR port order."""
    b = int(topo.port_off[u])
    Ds = [int(topo.edge_dim[topo.port_edge[b + j]]) for j in range(_deg(topo, u))]
    n = int(topo.site_off[u + 1] - topo.site_off[u])
    d = n // int(np.prod(Ds)) if Ds else n
    return sites[int(topo.site_off[u]):int(topo.site_off[u + 1])].reshape([d] + Ds)


def _msg(topo, msgs, q):
    D = int(topo.edge_dim[topo.port_edge[q]])
    m = msgs[int(topo.msg_off[q]):int(topo.msg_off[q + 1])]
    return m.reshape(D) if topo.plain else m.reshape(D, D)


def owner(topo, q):
    return int(np.searchsorted(np.asarray(topo.port_off), q, side="right") - 1)


def message(topo, sites, msgs, q, normalize=True):
    """The updated message of slot q = (u -> v), port p of u (SimpleMessageUpdate)."""
    sites, msgs = _c(sites), _c(msgs)
    u = owner(topo, q)
    b = int(topo.port_off[u])
    p, k = q - b, _deg(topo, u)
    A = _site(topo, sites, u)
    ket = [_L[1 + j] for j in range(k)]
    ops, subs = [A], ["a" + "".join(ket)]
    if topo.plain:
        for j in range(k):
            if j != p:
                ops.append(_msg(topo, msgs, int(topo.port_rev[b + j])))
                subs.append(ket[j])
        out = np.einsum(",".join(subs) + "->" + ket[p], *ops)
    else:
        bra = [_L[26 + j] for j in range(k)]
        ops.append(np.conj(A))
        subs.append("a" + "".join(bra))
        for j in range(k):
            if j != p:
                ops.append(_msg(topo, msgs, int(topo.port_rev[b + j])))
                subs.append(ket[j] + bra[j])
        out = np.einsum(",".join(subs) + "->" + ket[p] + bra[p], *ops)
    out = np.asarray(out, dtype=np.complex128).reshape(-1)
    if normalize:
        S = out.sum()
        if S != 0:
            out = out / S
    return out


def sweep_jacobi(topo, sites, msgs, normalize=True):
    msgs = _c(msgs)
    new = np.empty_like(msgs)
    for q in range(int(topo.port_off[topo.nv])):
        new[int(topo.msg_off[q]):int(topo.msg_off[q + 1])] = message(topo, sites, msgs, q, normalize)
    return new


def sweep_sequential(topo, sites, msgs, order, normalize=True):
    out = _c(msgs).copy()
    for q in order:
        out[int(topo.msg_off[q]):int(topo.msg_off[q + 1])] = message(topo, sites, out, q, normalize)
    return out


def _fidelity(a, b):
    na, nb = np.linalg.norm(a), np.linalg.norm(b)
    return 1.0 - abs(np.vdot(a / na, b / nb)) ** 2


def iterate_diff(topo, m1, m2):
    m1, m2 = _c(m1), _c(m2)
    nports = int(topo.port_off[topo.nv])
    if nports == 0:
        return -np.inf
    return max(_fidelity(m1[int(topo.msg_off[q]):int(topo.msg_off[q + 1])],
                         m2[int(topo.msg_off[q]):int(topo.msg_off[q + 1])]) for q in range(nports))


def _maxabs(m1, m2):
    return float(np.max(np.abs(m1 - m2))) if m1.size else 0.0


def bp(topo, sites, msgs, maxiter=None, tol=None, schedule="jacobi", order=None,
       criterion="fidelity", normalize=True):
    """beliefpropagation(...) restated: returns (messages, sweeps, delta, at_iteration)."""
    msgs = _c(msgs).copy()
    prev = msgs.copy()
    it, at, delta = 0, -1, np.inf
    while True:
        stop = maxiter is not None and maxiter > 0 and it >= maxiter
        if tol is not None:
            d = iterate_diff(topo, msgs, prev) if criterion == "fidelity" else _maxabs(msgs, prev)
            prev = msgs.copy()
            if it > 0:
                delta = d
                if d < tol:
                    at, stop = it, True
        if stop:
            break
        it += 1
        if schedule == "jacobi":
            msgs = sweep_jacobi(topo, sites, msgs, normalize)
        else:
            msgs = sweep_sequential(topo, sites, msgs, order, normalize)
        if (maxiter is None or maxiter <= 0) and tol is None:
            break
    return msgs, it, delta, at


def vertex_scalars(topo, sites, msgs):
    """Z_v: the factor of v contracted with every incoming message."""
    sites, msgs = _c(sites), _c(msgs)
    out = np.empty(topo.nv, dtype=np.complex128)
    for u in range(topo.nv):
        b, k = int(topo.port_off[u]), _deg(topo, u)
        A = _site(topo, sites, u)
        ket = [_L[1 + j] for j in range(k)]
        ops, subs = [A], ["a" + "".join(ket)]
        if not topo.plain:
            bra = [_L[26 + j] for j in range(k)]
            ops.append(np.conj(A))
            subs.append("a" + "".join(bra))
        for j in range(k):
            ops.append(_msg(topo, msgs, int(topo.port_rev[b + j])))
            subs.append(ket[j] if topo.plain else ket[j] + bra[j])
        out[u] = np.einsum(",".join(subs) + "->", *ops)
    return out


def edge_scalars(topo, msgs):
    """Z_e = sum_i m_uv[i] m_vu[i], one per undirected edge (edge order of the topology)."""
    msgs = _c(msgs)
    out = np.empty(topo.ne, dtype=np.complex128)
    for q in range(int(topo.port_off[topo.nv])):
        r = int(topo.port_rev[q])
        if q < r:
            out[int(topo.port_edge[q])] = np.sum(
                msgs[int(topo.msg_off[q]):int(topo.msg_off[q + 1])] *
                msgs[int(topo.msg_off[r]):int(topo.msg_off[r + 1])])
    return out


def bethe_free_energy(topo, sites, msgs):
    """sum log Z_v - sum log Z_e with principal logarithms; -Inf on a zero edge term."""
    zv = vertex_scalars(topo, sites, msgs)
    ze = edge_scalars(topo, msgs)
    if np.any(ze == 0):
        return -np.inf
    with np.errstate(divide="ignore"):
        return complex(np.sum(np.log(zv)) - np.sum(np.log(ze)))


def _network_args(topo, sites, open_phys):
    """einsum operands in sublist form: one integer label per edge, one per open physical
    index (numpy allows 52 labels)."""
    args, outs = [], []
    for u in range(topo.nv):
        b, k = int(topo.port_off[u]), _deg(topo, u)
        A = _site(topo, sites, u)
        lab = [int(topo.port_edge[b + j]) for j in range(k)]
        if open_phys:
            outs.append(topo.ne + u)
            args += [A, [topo.ne + u] + lab]
        else:
            args += [A.reshape(A.shape[1:]) if A.shape[0] == 1 else A, lab]
    if topo.ne + (topo.nv if open_phys else 0) > 52:
        raise ValueError("network too large for the einsum restatement")
    return args, outs


def plain_exact(topo, sites):
    """The full contraction of a plain network (every index shared by two tensors)."""
    args, _ = _network_args(topo, _c(sites), False)
    return complex(np.einsum(*args, [], optimize="greedy"))


def norm_exact(topo, sites):
    """<psi|psi> = sum |psi|^2 of the state the ket layer defines (physical indices open)."""
    args, outs = _network_args(topo, _c(sites), True)
    psi = np.einsum(*args, outs, optimize="greedy")
    return float(np.vdot(psi, psi).real)
