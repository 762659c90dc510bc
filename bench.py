"""Benchmark: BP message updates/s on the 32x32, d=2, D=8 PEPS norm network (FP32).

One step = one Jacobi sweep over all 3968 directed messages (update, sum normalisation,
residual) with inputs resident in HBM.  The timed region runs K sweeps through the
engine's sweep sequence (hipGraph replay with per-network convergence bookkeeping,
tol disabled so exactly K sweeps run), bracketed by torch.cuda.synchronize().

Also reported on the same JSON line:
  roofline      the dominant kernel (the message-update sweep kernel): algorithmic FLOPs
                per launch / its average launch duration, measured live with HIP events
                on the engine's stream;
  cpu_baseline  the FP64 CPU oracle (a C restatement of the reference's algorithm; the
                reference itself is Julia and cannot run here) on the host cores, over a
                bounded sample of sweeps of the same workload;
  extras        the plain FP64 GPU path (generic kernels) on the same workload, and the
                forward + adjoint timing of the differentiation workload when available.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# HBM bytes per launch of the dominant kernel, measured with two rocprofv3 --pmc passes
# (FETCH_SIZE, WRITE_SIZE) over this bench and scaled by a 256 MiB device copy of known
# traffic (fetch x2.00, write x1.00); the summary and its commands are in
# profiles/k_vertex_d8*_32x32_D8_fp32_hbm_traffic.md.  Keyed by (path, dtype, L, D, FP32
# vertex kernel variant).
MEASURED_TRAFFIC = {("vertex", "f32", 32, 8, 0): 37.54e6, ("vertex", "f32", 32, 8, 1): 41.14e6,
                    ("vertex", "f32", 32, 8, 4): 37.97e6, ("vertex", "f32", 32, 8, 5): 42.63e6,
                    ("vertex", "f64", 32, 8, 0): 210.3e6}
# the sweep kernel that dominates each configuration (NS: the physical dimension)
SWEEP_KERNEL = {("vertex", "f32", 0): "k_vertex_d8<NS, false>",
                ("vertex", "f32", 1): "k_vertex_d8_lo<NS, false>",
                ("vertex", "bf16", 0): "k_vertex_d8<NS, true>",
                ("vertex", "bf16", 1): "k_vertex_d8_lo<NS, true>",
                ("vertex", "f32", 2): "k_vertex_d8_mo<NS, false>",
                ("vertex", "f32", 3): "k_vertex_d8_mo<NS, true>",
                ("vertex", "f32", 4): "k_vertex_d8_hy<NS, false, false>",
                ("vertex", "f32", 5): "k_vertex_d8_hy<NS, true, false>",
                ("vertex", "bf16", 5): "k_vertex_d8_hy<NS, true, true>",
                ("vertex", "f64", 0): "k_vertex_d8_f64<NS>"}

PEAK = {  # MI355X published dense peaks
    "hbm_GBps": 8000.0,
    "fp32_mfma_TFLOPs": 157.3,
    "fp64_TFLOPs": 78.6,
    "fp32_valu_TFLOPs": 78.6,  # non-packed FP32 FMA on the vector units
    "bf16_mfma_TFLOPs": 2500.0,
}


def algorithmic_flops_per_sweep(topo) -> float:
    """sum over directed messages of 2 * k * d * D^(k+1) (k = degree of the sending vertex):
    k-1 absorptions of [d D^(k-1) x D] x [D x D] plus one [D x d D^(k-1)] x [d D^(k-1) x D]."""
    own = topo.port_owner()
    tot = 0.0
    for q in range(topo.nports):
        u = own[q]
        k = topo.degree(u)
        D = int(topo.edge_dim[topo.port_edge[q]])
        d = int(topo.phys_dim[u])
        tot += 2.0 * k * d * float(D) ** (k + 1)
    return tot


def algorithmic_bytes_per_sweep(topo, elt: int) -> float:
    """every site tensor read once + every message read once and written once."""
    return (float(topo.site_off[-1]) + 2.0 * float(topo.msg_off[-1])) * elt


def adjoint_timing(pkg, torch, L, D, dtype, seed):
    """BP to convergence, then the BP-estimated TFIM energy (J = 1, h = 3) and its gradient
    with respect to every site tensor by fixed-point adjoint sweeps; wall times of both."""
    import time

    g = pkg.named_grid((L, L))
    topo = pkg.build_topology(g, D, 2)
    sites = pkg.peps.random_sites(topo, seed, scale=pkg.peps.peps_scale(2, D))
    e = pkg.Engine(topo, sites, dtype=dtype)
    e.init_messages("identity")
    # steady state of a repeated call: one untimed run of each timed call first (the
    # context keeps its captured sweep graph, the adjoint scratch and the gradient buffer)
    e.run(maxiter=500, tol=1e-8)
    _, gbuf, _, _ = e.tfim_energy(1.0, 3.0, grad=True, adj_tol=1e-6, adj_maxiter=2000,
                                  device=True)
    e.init_messages("identity")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = e.run(maxiter=500, tol=1e-8)
    torch.cuda.synchronize()
    t_bp = time.perf_counter() - t0
    t1 = time.perf_counter()
    E, grad, it, res = e.tfim_energy(1.0, 3.0, grad=True, adj_tol=1e-6, adj_maxiter=2000,
                                     device=True, out=gbuf)
    torch.cuda.synchronize()
    t_adj = time.perf_counter() - t1
    t2 = time.perf_counter()
    e.tfim_energy(1.0, 3.0, grad=True, adj_tol=1e-6, adj_maxiter=2000)
    t_host = time.perf_counter() - t2
    msgs = int(topo.nports) * (int(r.sweeps[0]) + int(it))
    out = {"workload": f"{L}x{L} D={D} {dtype}: BP to fidelity 1e-8, then TFIM energy + "
                       f"gradient (adjoint to relative 1e-6), gradient left in HBM",
           "bp_sweeps": int(r.sweeps[0]), "bp_s": t_bp, "adjoint_sweeps": it,
           "adjoint_residual": res, "adjoint_s": t_adj, "energy": E,
           "bp_plus_adjoint_over_bp": (t_bp + t_adj) / t_bp,
           "forward_plus_adjoint_msg_per_s": msgs / (t_bp + t_adj),
           "adjoint_s_with_fp64_host_gradient": t_host, "path": e.path()}
    e.close()
    return out


def uneven_batch_timing(pkg, torch, seed, sweeps=20):
    """The batched uneven workload (64 networks, L in 8..48, D in {4,6,8,10,12}, square or
    heavy-hex) in one engine: autotune, then FP32 Jacobi sweeps with per-network masks."""
    import time

    tops, sites = pkg.peps.uneven_batch(seed=seed, nnet=64)
    b = pkg.concat_topologies(tops)
    e = pkg.Engine(b, np.concatenate(sites), dtype="f32")
    e.init_messages("identity")
    t0 = time.perf_counter()
    ntuned = e.autotune(mode="all")
    t_tune = time.perf_counter() - t0
    e.run(maxiter=sweeps, use_graph=True)  # warm up (captures the sweep graph)
    e.init_messages("identity")
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    r = e.run(maxiter=sweeps, use_graph=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t1
    flops = algorithmic_flops_per_sweep(b)
    nbytes = algorithmic_bytes_per_sweep(b, 4)
    per = wall / sweeps
    out = {"workload": f"64 uneven PEPS norm networks ({b.nv} vertices, {b.nports} messages), "
                       f"{sweeps} FP32 Jacobi sweeps, autotuned",
           "path": e.path(), "msg_per_s": b.nports * sweeps / wall,
           "ms_per_sweep": 1e3 * per, "tuned_decisions": ntuned, "tune_s": t_tune,
           "algorithmic_flops_per_sweep": flops, "algorithmic_bytes_per_sweep": nbytes,
           "algorithmic_TFLOPs": flops / per / 1e12,
           "frac_fp32_mfma_peak": flops / per / 1e12 / PEAK["fp32_mfma_TFLOPs"],
           "achieved_GBs": nbytes / per / 1e9}
    e.close()
    return out


def local_energy_timing(pkg, torch, seed, L=16, D=6):
    """Expectation values and a general local energy with its gradient on the C4 lattice
    (16x16, d=2, D=6, FP32), random one-site and nearest-neighbour operators; wall times
    after one untimed call of each (generic kernels), beside the TFIM energy + gradient on
    the MFMA kernels."""
    import time

    t = pkg.build_topology(pkg.named_grid((L, L)), D, 2)
    s = pkg.peps.random_sites(t, seed, scale=pkg.peps.peps_scale(2, D))
    rng = np.random.default_rng(seed)
    ov = rng.standard_normal((t.nv, 2, 2))
    oa = rng.standard_normal((t.ne, 2, 2))
    ob = rng.standard_normal((t.ne, 2, 2))
    e = pkg.Engine(t, s, dtype="f32")
    e.init_messages("identity")
    e.run(maxiter=500, tol=1e-8)

    def timed(f):
        f()
        t0 = time.perf_counter()
        out = f()
        return out, time.perf_counter() - t0

    _, t_v = timed(lambda: e.expect_vertex(ov))
    _, t_e = timed(lambda: e.expect_edge(oa, ob))
    buf = e.local_energy(ov, oa, ob, grad=True, adj_tol=1e-6, adj_maxiter=2000, device=True)[1]
    (E, _, it, res), t_l = timed(lambda: e.local_energy(ov, oa, ob, grad=True, adj_tol=1e-6,
                                                         adj_maxiter=2000, device=True, out=buf))
    (_, _, it_t, _), t_t = timed(lambda: e.tfim_energy(1.0, 3.0, grad=True, adj_tol=1e-6,
                                                        adj_maxiter=2000, device=True, out=buf))
    buf.free()
    e.close()
    return {"workload": f"{L}x{L} D={D} f32, random 2x2 one-site and nearest-neighbour "
                        f"operators", "expect_vertex_s": t_v, "expect_edge_s": t_e,
            "local_energy_plus_gradient_s": t_l, "local_energy_adjoint_sweeps": it,
            "local_energy": E, "tfim_plus_gradient_mfma_s": t_t, "tfim_adjoint_sweeps": it_t}


def norm_network_labels(pkg, t, s):
    """The double-layer tensors of <psi|psi> (ket A_u and bra A_u per vertex, sharing the
    site label, links labelled per layer), as the oracle's exact norm builds them."""
    tensors, labels = [], []
    for u in range(t.nv):
        A = s[t.site_off[u]:t.site_off[u + 1]].reshape(t.site_shape(u))
        ports = range(t.port_off[u], t.port_off[u + 1])
        tensors += [A, A]
        labels += [(("s", u),) + tuple(("k", int(t.port_edge[q])) for q in ports),
                   (("s", u),) + tuple(("b", int(t.port_edge[q])) for q in ports)]
    return tensors, labels


def contract_timing(pkg, seed, cases=((6, 3), (5, 4)), cpu_case=(6, 3), reps=5):
    """Exact <psi|psi> of L x L PEPS norm networks (d=2) on the device (ContractionPlan with
    the cheapest order, inputs resident in device memory): device time of the plan's launches
    (HIP events, the best of `reps` executions after one untimed), the algorithmic bytes
    (operands read once, results written once) per second against the HBM peak (the products
    are skinny: HBM-bound) and the FP64 rate; the oracle's frontier contraction of `cpu_case`
    timed beside it on one core, with the relative difference of the two values."""
    import time

    out = []
    for L, D in cases:
        t = pkg.build_topology(pkg.named_grid((L, L)), D, 2)
        s = pkg.peps.random_sites(t, seed, scale=pkg.peps.peps_scale(2, D))
        tensors, labels = norm_network_labels(pkg, t, s)
        t0 = time.perf_counter()
        plan = pkg.ContractionPlan([x.shape for x in tensors], labels, order_alg="cheapest")
        t_plan = time.perf_counter() - t0
        flat = np.concatenate([x.ravel() for x in tensors])
        din = pkg._lib.DeviceArray(flat.size, np.float64)
        dout = pkg._lib.DeviceArray(1, np.float64)
        din.upload(flat)
        plan.execute_device(din, dout)
        ms = []
        for _ in range(reps):
            plan.execute_device(din, dout)
            ms.append(plan.last_ms)
        z = float(dout.numpy()[0])
        t1 = time.perf_counter()
        zh = plan(flat=flat)
        t_host = time.perf_counter() - t1
        best = min(ms)
        rec = {"workload": f"exact <psi|psi>, {L}x{L} PEPS norm network, d=2, D={D}, "
                           f"{len(tensors)} tensors, cheapest order",
               "steps": len(tensors) - 1, "GFLOP": plan.flops / 1e9,
               "workspace_MB": plan.workspace_bytes / 1e6, "plan_s": t_plan,
               "device_ms_best": best, "device_ms_median": float(np.median(ms)),
               "algorithmic_GB": plan.bytes / 1e9,
               "achieved_TBps": plan.bytes / (best * 1e-3) / 1e12,
               "frac_hbm_peak": plan.bytes / (best * 1e-3) / 8e12,
               "fp64_TFLOPs": plan.flops / (best * 1e-3) / 1e12,
               "frac_fp64_peak": plan.flops / (best * 1e-3) / 78.6e12,
               "host_buffers_call_s": t_host, "value": z, "host_value_equal": zh == z}
        if (L, D) == cpu_case:
            orc = __import__("__graft_entry__").import_oracle()
            t2 = time.perf_counter()
            zo = orc.norm_exact(t, s)
            t_cpu = time.perf_counter() - t2
            rec.update({"cpu_oracle_s": t_cpu, "cpu_oracle_order": "left-associative frontier",
                        "cpu_cores": 1, "rel_diff_vs_oracle": abs(z - zo) / abs(zo)})
        din.free()
        dout.free()
        plan.close()
        out.append(rec)
    return out


class _PackedGate:
    """A gate already in the packed C-ABI form (descriptor and inputs)."""

    def __init__(self, desc, data):
        self.desc = desc
        self.data = data


def gate_layer(pkg, t, s, m, gate_mats, maxdim, pairs):
    """Packed two-site gates on the bonds `pairs` (vertex-index pairs) of topology t with
    site tensors s and messages m (flat, engine layout); the same inputs in the oracle's
    form."""
    packed, plain = [], []
    for (u1, u2), G in zip(pairs, gate_mats):
        desc = np.zeros(pkg._lib.BP_GATE_DESC, dtype=np.int32)
        desc[0], desc[1], desc[2] = 2, t.phys_dim[u1], t.phys_dim[u2]
        parts, tens, msgs, ports = [], [], [], []
        for side, (u, w) in enumerate(((u1, u2), (u2, u1))):
            qs = list(range(t.port_off[u], t.port_off[u + 1]))
            dims = [int(t.edge_dim[t.port_edge[q]]) for q in qs]
            p = [int(t.port_nbr[q]) for q in qs].index(w)
            desc[3 + side], desc[5 + side] = len(qs), p
            desc[9 + 8 * side: 9 + 8 * side + len(qs)] = dims
            A = s[t.site_off[u]:t.site_off[u + 1]].reshape(t.site_shape(u))
            tens.append(A)
            mm = {}
            for j, q in enumerate(qs):
                if j != p:
                    sl = int(t.port_rev[q])
                    mm[j] = m[t.msg_off[sl]:t.msg_off[sl + 1]].reshape(dims[j], dims[j])
            msgs.append(mm)
            ports.append(p)
        desc[7] = maxdim
        parts = [tens[0].ravel(), tens[1].ravel()]
        parts += [msgs[0][j].ravel() for j in sorted(msgs[0])]
        parts += [msgs[1][j].ravel() for j in sorted(msgs[1])]
        parts.append(G.ravel())
        packed.append(_PackedGate(desc, np.concatenate(parts)))
        plain.append((tens[0], tens[1], ports[0], ports[1], msgs[0], msgs[1], G))
    return packed, plain


def gate_timing(pkg, seed, L=32, D=8, reps=3, cpu_sample=16):
    """A simple-update layer on the L x L PEPS (d=2, bond dimension D): random two-site gates
    on the L * L/2 vertex-disjoint horizontal bonds (i, 2j-1)-(i, 2j), truncated to D, with
    the converged BP messages of the state (FP32 vertex path, downloaded as FP64) as the
    environments; one bp_apply_gates call per layer.  Device time of the gate kernel (HIP
    events, best of `reps` after one untimed call) and the call's wall time including the
    PCIe copies; the numpy oracle timed on `cpu_sample` gates of the same layer; singular
    values of those gates against the oracle's."""
    import time

    t = pkg.build_topology(pkg.named_grid((L, L)), D, 2)
    s = pkg.peps.random_sites(t, seed, scale=pkg.peps.peps_scale(2, D))
    e = pkg.Engine(t, s, dtype="f32")
    e.init_messages("identity")
    r = e.run(maxiter=200, tol=1e-10)
    m = e.messages()
    idx = {v: i for i, v in enumerate(t.vertices)}
    pairs = [(idx[(i, j)], idx[(i, j + 1)]) for i in range(1, L + 1) for j in range(1, L, 2)]
    rng = np.random.default_rng(seed)
    mats = [rng.standard_normal((4, 4)) for _ in pairs]
    packed, plain = gate_layer(pkg, t, s, m, mats, D, pairs)
    res, _ = pkg.apply.apply_gates_device(packed, -1.0)
    ms, walls = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        res, kms = pkg.apply.apply_gates_device(packed, -1.0)
        walls.append(time.perf_counter() - t0)
        ms.append(kms)
    orc = __import__("__graft_entry__").import_oracle()
    t1 = time.perf_counter()
    errs = []
    for pg, (A1, A2, p1, p2, m1, m2, G), (out, chi) in zip(packed[:cpu_sample],
                                                          plain[:cpu_sample],
                                                          res[:cpu_sample]):
        _, _, So = orc.apply_gate2(A1, A2, p1, p2, m1, m2, G, maxdim=D)
        caps = pkg.apply._out_caps(pg.desc)
        Sd = out[caps[0] + caps[1]:caps[0] + caps[1] + chi]
        errs.append(float(np.max(np.abs(Sd - So)) / So[0]))
    t_cpu = (time.perf_counter() - t1) / cpu_sample
    # the same layer in place on the engine's resident network (no site tensor or message
    # crosses PCIe; the FP32 messages are read and written in place)
    slots = []
    for u1, u2 in pairs:
        slots.append(next(q for q in range(t.port_off[u1], t.port_off[u1 + 1])
                          if int(t.port_nbr[q]) == u2))
    kinds = [2] * len(pairs)
    e.apply_gates(kinds, slots, mats, maxdim=D)
    rms, rwall = [], []
    for _ in range(reps):
        t2 = time.perf_counter()
        e.apply_gates(kinds, slots, mats, maxdim=D)
        rwall.append(time.perf_counter() - t2)
        rms.append(e.last_run_ms())
    e.close()
    best = min(ms)
    n = len(pairs)
    return {"workload": f"simple-update layer, {L}x{L} PEPS d=2 D={D}: {n} two-site gates on "
                        f"disjoint bonds, truncated to {D}, converged BP messages "
                        f"({int(r.sweeps[0])} FP32 sweeps)",
            "gates": n, "kernel_ms_best": best, "kernel_ms_median": float(np.median(ms)),
            "gates_per_s": n / (best * 1e-3), "call_wall_s": float(min(walls)),
            "resident_device_ms_best": min(rms), "resident_call_wall_s": float(min(rwall)),
            "cpu_oracle_s_per_gate": t_cpu, "cpu_oracle_gates_per_s": 1.0 / t_cpu,
            "cpu_cores": 1, "cpu_sample_gates": cpu_sample,
            "max_rel_diff_S_vs_oracle": max(errs)}


def large_bond_timing(pkg, torch, seed, L=24, D=16):
    """The mixed-precision workload at large bond dimension: BF16 MFMA sweeps (FP32 sums,
    per-message max-norm scaling) to fidelity 1e-5, then 3 FP64 polishing sweeps; the
    message and log-norm change from each stage to the next is reported (the oracle check
    of both stages against FP64 runs on 5x5 at this bond dimension in the tests)."""
    import time

    t = pkg.build_topology(pkg.named_grid((L, L)), D, 2)
    s = pkg.peps.random_sites(t, seed, scale=pkg.peps.peps_scale(2, D))
    e = pkg.Engine(t, s, dtype="bf16")
    e.init_messages("identity")
    e.autotune(mode="all")
    kms = e.time_sweeps(10)
    e.init_messages("identity")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = e.run(maxiter=200, tol=1e-5)
    torch.cuda.synchronize()
    t_bf = time.perf_counter() - t0
    m_bf, ln_bf = e.messages(), e.lognorm()
    t1 = time.perf_counter()
    e.polish_fp64(3)
    t_pol = time.perf_counter() - t1
    m_pol, ln_pol = e.messages(), e.lognorm()
    e.polish_fp64(3)
    m_pol2, ln_pol2 = e.messages(), e.lognorm()
    flops = algorithmic_flops_per_sweep(t)
    bf_path = e.path()
    e.close()
    # the FP64 sweep alone (the polish's kernel), against the generic FP64 kernel
    fp64 = {}
    for path in ("auto", "generic"):
        with pkg.Engine(t, s, dtype="f64", path=path) as e64:
            e64.init_messages("identity")
            fp64[path] = (e64.path(), e64.time_sweeps(3 if path == "generic" else 10))

    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    out = {"workload": f"{L}x{L} PEPS norm network, d=2, D={D}, BF16 MFMA operands / FP32 sums, "
                       f"then 3 FP64 polishing sweeps",
           "path": bf_path, "kernel_ms_per_sweep": kms,
           "msg_per_s_kernel": t.nports / (kms * 1e-3),
           "algorithmic_TFLOPs": flops / (kms * 1e-3) / 1e12,
           "bf16_sweeps": int(r.sweeps[0]), "bf16_s": t_bf, "polish3_s": t_pol,
           "msg_change_bf16_to_polished": rel(m_bf, m_pol),
           "msg_change_polished_3_to_6": rel(m_pol, m_pol2),
           "lognorm_change_bf16_to_polished": abs(ln_bf - ln_pol) / abs(ln_pol),
           "lognorm_change_polished_3_to_6": abs(ln_pol - ln_pol2) / abs(ln_pol2),
           "fp64_sweep_path": fp64["auto"][0], "fp64_sweep_ms": fp64["auto"][1],
           "fp64_sweep_algorithmic_TFLOPs": flops / (fp64["auto"][1] * 1e-3) / 1e12,
           "fp64_generic_sweep_ms": fp64["generic"][1]}
    return out


def small_bond_timing(pkg, seed, cases=((512, 2), (256, 3), (256, 4)), d=2):
    """Bond dimensions below 8: the small-bond kernel (k_vertex_s) on an L x L FP32 lattice per
    (L, D), against the same sweep zero-padded to 8 on the D <= 8 MFMA kernel and on the
    generic kernel.  Algorithmic bytes: FP32 sites read once, messages read and written once;
    algorithmic FLOPs: sum over directed messages of 2 k d D^(k+1)."""
    out = []
    for L, D in cases:
        t = pkg.build_topology(pkg.named_grid((L, L)), D, d)
        s = pkg.peps.random_sites(t, seed, scale=pkg.peps.peps_scale(d, D))
        flops = algorithmic_flops_per_sweep(t)
        nbytes = algorithmic_bytes_per_sweep(t, 4)
        row = {"workload": f"{L}x{L} PEPS norm network, d={d}, D={D}, FP32 Jacobi sweep",
               "messages_per_sweep": int(t.nports), "algorithmic_bytes": nbytes,
               "algorithmic_flops": flops}
        for name, opts, path, reps in (("small_bond", {"small_bond": 1}, "auto", 50),
                                       ("padded_to_8", {"small_bond": 0}, "auto", 5),
                                       ("generic", {}, "generic", 5)):
            with pkg.Engine(t, s, dtype="f32", path=path) as e:
                for k, v in opts.items():
                    e.set_option(k, v)
                e.init_messages("identity")
                ms = e.time_sweeps(reps)
            row[f"{name}_kernel_ms"] = ms
            if name == "small_bond":
                gbs = nbytes / (ms * 1e-3) / 1e9
                row.update({"path": path, "achieved_GBs": gbs,
                            "frac_hbm_peak": gbs / PEAK["hbm_GBps"],
                            "algorithmic_TFLOPs": flops / (ms * 1e-3) / 1e12,
                            "frac_fp32_valu_peak": flops / (ms * 1e-3) / 1e12 / PEAK["fp32_valu_TFLOPs"],
                            "msg_per_s": t.nports / (ms * 1e-3)})
        row["speedup_vs_padded"] = row["padded_to_8_kernel_ms"] / row["small_bond_kernel_ms"]
        row["speedup_vs_generic"] = row["generic_kernel_ms"] / row["small_bond_kernel_ms"]
        out.append(row)
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def precision_vs_oracle(pkg, oracle, threads, seed):
    """Relative errors (max-abs over the largest entry) of the GPU results against the FP64
    oracle's converged solution: the 32x32 D=8 FP32 forward at fidelity 1e-7 and 1e-12,
    its log-norm and TFIM gradients, FP64 polishing; the 24x24 D=16 BF16 solve and each
    polishing stage."""
    def rel(a, b):
        return float(np.max(np.abs(a - b)) / np.max(np.abs(b)))

    out = {}
    t = pkg.build_topology(pkg.named_grid((32, 32)), 8, 2)
    s = pkg.peps.random_sites(t, seed, scale=pkg.peps.peps_scale(2, 8))
    m0 = pkg.peps.identity_messages(t)
    m64, _, _, _ = oracle.bp(t, s, m0, maxiter=400, tol=1e-15, criterion="maxabs",
                             nthreads=threads)
    ln64 = oracle.bethe_free_energy(t, s, m64)
    gl64 = oracle.grad_lognorm_explicit(t, s, m64)
    E64, dm, da = oracle.tfim_energy(t, s, m64, 1.0, 3.0, partials=True)
    g64, _, _ = oracle.fixed_point_gradient(t, s, m64, dm, da, maxiter=500, tol=1e-13,
                                            nthreads=threads)
    with pkg.Engine(t, s, dtype="f32") as e:
        for tol in (1e-7, 1e-12):
            e.set_messages(m0)
            r = e.run(maxiter=200, tol=tol)
            entry = {"sweeps": int(r.sweeps[0]), "messages": rel(e.messages(), m64),
                     "lognorm": abs(e.lognorm() - ln64) / abs(ln64)}
            if tol == 1e-12:
                E, g, it, res = e.tfim_energy(1.0, 3.0, grad=True, adj_tol=1e-6,
                                              adj_maxiter=200)
                entry.update({"lognorm_gradient": rel(e.grad_lognorm(), gl64),
                              "tfim_energy": abs(E - E64) / abs(E64), "tfim_gradient": rel(g, g64),
                              "adjoint_sweeps": it})
            out[f"c2_fp32_fidelity_{tol:g}"] = entry
        e.polish_fp64(10)
        out["c2_fp32_then_10_fp64_polish"] = {"messages": rel(e.messages(), m64),
                                              "lognorm": abs(e.lognorm() - ln64) / abs(ln64)}
    t = pkg.build_topology(pkg.named_grid((24, 24)), 16, 2)
    s = pkg.peps.random_sites(t, seed, scale=pkg.peps.peps_scale(2, 16))
    m0 = pkg.peps.identity_messages(t)
    m64, _, _, _ = oracle.bp(t, s, m0, maxiter=60, tol=1e-13, criterion="maxabs",
                             nthreads=threads)
    ln64 = oracle.bethe_free_energy(t, s, m64)
    with pkg.Engine(t, s, dtype="bf16") as e:
        e.set_messages(m0)
        r = e.run(maxiter=200, tol=1e-5)
        stages = {"bf16": (int(r.sweeps[0]), 0)}
        errs = {}
        for name, n in (("bf16", 0), ("bf16_then_3_fp64_polish", 3),
                        ("bf16_then_6_fp64_polish", 3)):
            if n:
                e.polish_fp64(n)
            errs[name] = {"messages": rel(e.messages(), m64),
                          "lognorm": abs(e.lognorm() - ln64) / abs(ln64)}
        errs["bf16"]["sweeps"] = stages["bf16"][0]
    out["c5_24x24_D16"] = errs
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--dtype", default="f32")
    ap.add_argument("--path", default="auto")
    ap.add_argument("--L", type=int, default=32)
    ap.add_argument("--D", type=int, default=8)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extras", action="store_true")
    ap.add_argument("--only", choices=["adjoint", "large_bond", "uneven", "local_energy",
                                       "contract", "gates", "small_bond"],
                    default=None,
                    help="adjoint: run only the forward + adjoint workload (L x L, D; three "
                         "energy + gradient calls); large_bond: only the 24x24 D=16 BF16 + "
                         "FP64-polish workload; uneven: only the 64-network batched "
                         "workload; local_energy: only the expectation-value / local-energy "
                         "timings; contract: only the exact-contraction timings; gates: only "
                         "the gate-layer timing; small_bond: only the bond-dimension <= 4 "
                         "timings; print the timings, for profiling")
    ap.add_argument("--vertex-variant", type=int, choices=[0, 1, 2, 3, 4, 5], default=None,
                    help="FP32 vertex sweep kernel to run instead of the tuned one (0: two LDS "
                         "buffers per wave, 1: one buffer and two waves per SIMD, 2: MFMA open "
                         "contractions, bra in LDS, 3: MFMA open contractions, bra from HBM, "
                         "4: half of the open contractions on MFMA in place, 5: all of them)")
    ap.add_argument("--copy-probe-mb", type=int, default=0,
                    help="first run one device-to-device copy of this many MiB (a kernel of "
                         "known traffic, used to calibrate rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args(argv)

    if args.gpus != 1:
        print(json.dumps({"error": "this engine runs on one GPU; --gpus must be 1 "
                          "(the multi-GPU domain decomposition is design-only)"}))
        return 2

    import numpy as np
    import torch
    import __graft_entry__ as ge

    pkg = ge.import_package()
    if args.only == "adjoint":
        torch.cuda.init()
        runs = [adjoint_timing(pkg, torch, args.L, args.D, args.dtype, args.seed)
                for _ in range(3)]
        print(json.dumps({"adjoint_only": runs}), flush=True)
        return 0
    if args.only == "gates":
        torch.cuda.init()
        print(json.dumps({"gates_only": gate_timing(pkg, args.seed)}), flush=True)
        return 0
    if args.only == "contract":
        torch.cuda.init()
        print(json.dumps({"contract_only": contract_timing(pkg, args.seed)}), flush=True)
        return 0
    if args.only == "local_energy":
        torch.cuda.init()
        print(json.dumps({"local_energy_only": local_energy_timing(pkg, torch, args.seed)}),
              flush=True)
        return 0
    if args.only == "uneven":
        torch.cuda.init()
        print(json.dumps({"uneven_only": uneven_batch_timing(pkg, torch, args.seed)}),
              flush=True)
        return 0
    if args.only == "small_bond":
        torch.cuda.init()
        print(json.dumps({"small_bond_only": small_bond_timing(pkg, args.seed)}), flush=True)
        return 0
    if args.only == "large_bond":
        torch.cuda.init()
        print(json.dumps({"large_bond_only": large_bond_timing(pkg, torch, args.seed)}),
              flush=True)
        return 0
    L, D, d = args.L, args.D, 2
    g = pkg.named_grid((L, L))
    topo = pkg.build_topology(g, D, d)
    sites = pkg.peps.random_sites(topo, args.seed, scale=pkg.peps.peps_scale(d, D))
    nmsg = topo.nports
    flops = algorithmic_flops_per_sweep(topo)
    elt = 8 if args.dtype == "f64" else 4
    byts = algorithmic_bytes_per_sweep(topo, elt)

    torch.cuda.init()
    if args.copy_probe_mb > 0:
        n = args.copy_probe_mb * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    eng = pkg.Engine(topo, sites, dtype=args.dtype, path=args.path)
    eng.init_messages("identity")
    t_tune = time.perf_counter()
    ntuned = eng.autotune(mode="all")  # kernel choices for this shape, outside the timed region
    t_tune = time.perf_counter() - t_tune
    path = eng.path()
    fp32_vertex = path == "vertex" and args.dtype == "f32"
    if fp32_vertex and args.vertex_variant is not None:
        eng.set_option("vertex_variant", args.vertex_variant)
    variant = eng.get_option("vertex_variant") if fp32_vertex else 0
    # capture the sweep graphs the timed run replays (no sweep runs), then warm up
    eng.prepare_run(args.steps)
    if args.warmup > 0:
        eng.run(maxiter=args.warmup, use_graph=True)
    eng.init_messages("identity")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    res = eng.run(maxiter=args.steps, use_graph=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    assert int(res.sweeps[0]) == args.steps, res
    ms_per_step = 1e3 * wall / args.steps
    value = nmsg * args.steps / wall

    # dominant kernel: the sweep kernel(s), averaged over back-to-back launches
    kms = eng.time_sweeps(max(20, min(args.steps, 200)))
    if args.dtype == "bf16":
        bound, peak, unit = "mfma", PEAK["bf16_mfma_TFLOPs"], "TFLOP/s"
    elif args.dtype == "f32" and path == "vertex":
        bound, peak, unit = "mfma", PEAK["fp32_mfma_TFLOPs"], "TFLOP/s"
    elif args.dtype == "f64":
        bound, peak, unit = "mfma", PEAK["fp64_TFLOPs"], "TFLOP/s"
    else:
        bound, peak, unit = "mfma", PEAK["fp32_mfma_TFLOPs"], "TFLOP/s"
    if D < 8:
        bound, peak, unit = "hbm", PEAK["hbm_GBps"], "GB/s"
        achieved = byts / (kms * 1e-3) / 1e9
    else:
        achieved = flops / (kms * 1e-3) / 1e12
    roofline = {"bound": bound, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
                "frac": round(achieved / peak, 4),
                "traffic": MEASURED_TRAFFIC.get((path, args.dtype, L, D, variant)),
                "kernel_ms": round(kms, 5), "algorithmic_flops_per_launch": flops,
                "algorithmic_bytes_per_launch": byts}

    extras = {"path": path, "sweep_kernel": SWEEP_KERNEL.get((path, args.dtype, variant)),
              "device_ms_per_step": round(res.ms / args.steps, 5),
              "autotune": {"decisions": ntuned, "seconds": round(t_tune, 3)}}
    if path == "vertex" and args.dtype == "f32":
        # three FP32 sweep kernels, timed the same way (the run above used the tuned one)
        for v in (0, 1, 5):
            e2 = pkg.Engine(topo, sites, dtype="f32")
            e2.set_option("vertex_variant", v)
            e2.init_messages("identity")
            extras[f"vertex_variant_{v}_kernel_ms"] = e2.time_sweeps(100)
            e2.close()
    if not args.no_extras and args.dtype != "f64":
        # the single-network FP64 GPU paths on the same workload: the FP64 vertex kernel
        # (the engine's FP64 path at this shape) and the generic FP64 kernel
        for name, p64 in (("fp64_vertex", "auto"), ("fp64_generic", "generic")):
            e64 = pkg.Engine(topo, sites, dtype="f64", path=p64)
            e64.init_messages("identity")
            e64.autotune(mode="all")  # each FP64 path at its best launch configuration
            e64.run(maxiter=3)
            e64.init_messages("identity")
            n64 = 100 if p64 == "auto" else 20
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            r64 = e64.run(maxiter=n64, use_graph=True)
            torch.cuda.synchronize()
            w64 = time.perf_counter() - t1
            k64 = e64.time_sweeps(50)
            extras[name] = {"path": e64.path(), "msg_per_s": nmsg * int(r64.sweeps[0]) / w64,
                            "kernel_ms": k64,
                            "algorithmic_TFLOPs": flops / (k64 * 1e-3) / 1e12,
                            "frac_fp64_peak": flops / (k64 * 1e-3) / 1e12 / PEAK["fp64_TFLOPs"]}
            e64.close()
        extras["speedup_vs_fp64_vertex_gpu"] = value / extras["fp64_vertex"]["msg_per_s"]
        extras["speedup_vs_fp64_generic_gpu"] = value / extras["fp64_generic"]["msg_per_s"]
    eng.close()
    if not args.no_extras and args.dtype == "f32" and D == 8:
        eb = pkg.Engine(topo, sites, dtype="bf16")
        eb.init_messages("identity")
        eb.autotune(mode="all")  # the BF16 sweep kernel variant, outside the timed region
        eb.run(maxiter=10)
        eb.init_messages("identity")
        torch.cuda.synchronize()
        t3 = time.perf_counter()
        eb.run(maxiter=args.steps, use_graph=True)
        torch.cuda.synchronize()
        wb = time.perf_counter() - t3
        kb = eb.time_sweeps(max(20, min(args.steps, 200)))
        vb = eb.get_option("vertex_variant")
        extras["bf16"] = {"msg_per_s": nmsg * args.steps / wb, "kernel_ms": kb,
                          "algorithmic_TFLOPs": flops / (kb * 1e-3) / 1e12,
                          "sweep_kernel": SWEEP_KERNEL.get(("vertex", "bf16", vb))}
        for v in (0, 1, 5):
            eb.set_option("vertex_variant", v)
            extras["bf16"][f"vertex_variant_{v}_kernel_ms"] = eb.time_sweeps(100)
        eb.close()
    if not args.no_extras:
        extras["adjoint"] = [adjoint_timing(pkg, torch, L, D, args.dtype, args.seed),
                             adjoint_timing(pkg, torch, 16, 6, args.dtype, args.seed)]
        if "fp64_vertex" in extras:
            # the target ratio: FP32 forward + adjoint message updates/s over the
            # single-network FP64 GPU path's (both of its kernels)
            fa = extras["adjoint"][0]["forward_plus_adjoint_msg_per_s"]
            extras["forward_plus_adjoint_over_fp64_vertex"] = fa / extras["fp64_vertex"]["msg_per_s"]
            extras["forward_plus_adjoint_over_fp64_generic"] = fa / extras["fp64_generic"]["msg_per_s"]

    if not args.no_extras and args.dtype == "f32":
        extras["uneven_batch"] = uneven_batch_timing(pkg, torch, args.seed)
        extras["large_bond_bf16"] = large_bond_timing(pkg, torch, args.seed)
        extras["small_bond"] = small_bond_timing(pkg, args.seed)
        extras["local_energy"] = local_energy_timing(pkg, torch, args.seed)
        extras["contract_network"] = contract_timing(pkg, args.seed)
        extras["gate_layer"] = gate_timing(pkg, args.seed)

    cpu = None
    if not args.no_cpu:
        oracle = ge.import_oracle()
        threads = min(os.cpu_count() or 1, 16)

        def cpu_rate(nthreads, seconds):
            cur = pkg.peps.identity_messages(topo)
            n = 0
            t2 = time.perf_counter()
            while True:
                cur, _, _ = oracle.sweep_jacobi(topo, sites, cur, nthreads=nthreads)
                n += 1
                if time.perf_counter() - t2 > seconds:
                    break
            return n, time.perf_counter() - t2

        n1, w1 = cpu_rate(1, max(2.0, args.cpu_seconds / 3))
        n, wc = cpu_rate(threads, args.cpu_seconds)
        cpu = {"value": nmsg * n / wc, "unit": "msg updates/s", "cores": threads,
               "kind": "port",
               "sample": f"{n} FP64 Jacobi sweeps of the same {L}x{L} D={D} network "
                         f"({n * nmsg} message updates) by the C oracle with OpenMP over "
                         f"{threads} threads; the 1-thread leg ran {n1} sweeps",
               "value_1_thread": nmsg * n1 / w1, "cpu_model": cpu_model()}
        if not args.no_extras and args.dtype == "f32" and (L, D) == (32, 8):
            # the oracle as the reference values: errors of the GPU results against FP64
            extras["precision_vs_fp64_oracle"] = precision_vs_oracle(pkg, oracle, threads,
                                                                     args.seed)

    line = {
        "metric": "BP msg updates/s, 32x32 D=8 PEPS norm net",
        "value": value,
        "unit": "msg updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"f32": "fp32", "f64": "fp64", "bf16": "bf16"}[args.dtype],
        "data": "synthetic (seeded Philox N(0,1)/sqrt(d*D^4) site tensors, identity messages)",
        "config": {"workload": f"{L}x{L} open square-lattice PEPS norm network, d={d}, D={D}, "
                               f"Jacobi BP sweep over {nmsg} directed messages",
                   "L": L, "D": D, "d": d, "messages_per_sweep": nmsg},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "extras": extras,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
