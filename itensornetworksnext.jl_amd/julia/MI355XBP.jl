"""
    MI355XBP

Julia binding of libbpmi355x.so (include/bp_mi355x.h) for ITensorNetworksNext.jl.

Adds two strategies to the package's BP plugin surface
(src/beliefpropagation/beliefpropagation.jl:214-257, src/select_algorithm.jl:41-51):

  * `DeviceJacobiUpdate(; dtype = :f32, path = :auto, polish = 0)` — passed as
    `message_update_algorithm`, it makes `beliefpropagation` run whole Jacobi sweeps on the
    GPU: the edge schedule collapses to one `JacobiSweep()` token per sweep, and
    `message_update!(alg, cache, factors, ::JacobiSweep)` performs the sweep on the device
    (bp_sweep), so the reference's outer loop, StopWhenConverged and iterate_diff keep
    their meaning;
  * `device_beliefpropagation(nn, messages; stopping_criterion, alg)` — the same solve in one
    call with convergence tracked on the device (bp_run), returning a `MessageCache`.

Also expectation values and local energies with their adjoint gradients, and the exact
contraction `device_contract_network` (bp_contract_network, the counterpart of
`contract_network(tn; alg = Exact())`).

This file is not exercised in the build container (no Julia there); the same C ABI is
exercised by the Python ctypes binding and its tests.
"""
module MI355XBP

using ITensorNetworksNext: ITensorNetworksNext, MessageCache, MessageUpdateAlgorithm,
    NormNetwork, beliefpropagation, braname, kettensor, linkinds, siteinds
using ITensorNetworksNext.AlgorithmsInterfaceExtensions: AlgorithmsInterfaceExtensions as AIE
using ITensorBase: dimnames, unnamed, name
using Graphs: vertices, edges, src, dst, neighbors
import ITensorNetworksNext: message_update!, bethe_free_energy

const libbp = Ref{String}(joinpath(@__DIR__, "..", "csrc", "libbpmi355x.so"))

const BP_F64, BP_F32, BP_BF16, BP_C128 = Cint(0), Cint(1), Cint(2), Cint(3)
const BP_INIT_UPLOAD = Cint(2)
const BP_CRIT_FIDELITY = Cint(0)

function check(rc::Cint, what)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:bp_last_error, libbp[]), Cstring, ()))
    return error("$what failed (status $rc): $msg")
end

"""A device context holding one flattened NormNetwork."""
mutable struct DeviceContext
    handle::Ptr{Cvoid}
    nn::NormNetwork
    vertices::Vector{Any}
    slots::Dict{Any, UnitRange{Int}}   # directed edge => range in the flat message buffer
    dims::Dict{Any, Int}
    nmsg::Int
    T::DataType   # element type of the host buffers: Float64, or ComplexF64 (BP_C128)
end

function DeviceContext(nn::NormNetwork; dtype::Symbol = :f32, device::Integer = 0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    # complex site tensors (ComplexF32 promotes) run as BP_C128 whatever `dtype` asks:
    # sites, messages and scalars then cross the ABI as interleaved (re, im) doubles, which
    # is the memory layout of a Vector{ComplexF64}
    iscplx = any(v -> eltype(unnamed(kettensor(nn, v))) <: Complex, vertices(nn))
    T = iscplx ? ComplexF64 : Float64
    dt = iscplx ? BP_C128 : dtype === :f64 ? BP_F64 : dtype === :bf16 ? BP_BF16 : BP_F32
    check(ccall((:bp_create, libbp[]), Cint, (Ref{Ptr{Cvoid}}, Cint, Cint), h, device, dt),
        "bp_create")
    vs = collect(vertices(nn))
    vidx = Dict(v => i - 1 for (i, v) in enumerate(vs))
    es = collect(edges(nn))
    eidx = Dict{Any, Int}()
    for (k, e) in enumerate(es)
        eidx[(src(e), dst(e))] = k - 1
        eidx[(dst(e), src(e))] = k - 1
    end
    port_off = Cint[0]
    port_nbr = Cint[]
    port_edge = Cint[]
    nbrs = Dict(v => sort(collect(neighbors(nn, v)); by = w -> vidx[w]) for v in vs)
    for v in vs
        for w in nbrs[v]
            push!(port_nbr, vidx[w])
            push!(port_edge, eidx[(v, w)])
        end
        push!(port_off, length(port_nbr))
    end
    port_rev = similar(port_nbr)
    for (i, v) in enumerate(vs), (j, w) in enumerate(nbrs[v])
        q = port_off[i] + j - 1
        k = findfirst(==(v), nbrs[w])
        port_rev[q + 1] = port_off[vidx[w] + 1] + k - 1
    end
    edge_dim = Cint[prod(length ∘ unnamed, linkinds(nn.ket, (src(e), dst(e)))) for e in es]
    phys_dim = Cint[prod(length ∘ unnamed, siteinds(nn.ket, v)) for v in vs]
    check(ccall((:bp_set_topology, libbp[]), Cint,
            (Ptr{Cvoid}, Cint, Cint, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Ptr{Cint}, Ptr{Cint},
                Ptr{Cint}, Ptr{Cint}, Cint, Cint),
            h[], length(vs), length(es), port_off, port_nbr, port_edge, port_rev, edge_dim,
            phys_dim, C_NULL, 1, 0), "bp_set_topology")
    # site tensors permuted to [site][port 0]...[port k-1], row-major, Float64 / ComplexF64
    sites = T[]
    for v in vs
        t = kettensor(nn, v)
        order = vcat(collect(siteinds(nn.ket, v)),
            reduce(vcat, [collect(linkinds(nn.ket, (v, w))) for w in nbrs[v]]; init = []))
        a = permutedims(unnamed(t, order...), reverse(1:length(order)))  # row-major flatten
        append!(sites, vec(Array{T}(a)))
    end
    GC.@preserve sites check(ccall((:bp_upload_sites, libbp[]), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}), h[], sites), "bp_upload_sites")
    slots = Dict{Any, UnitRange{Int}}()
    dims = Dict{Any, Int}()
    off = 0
    for (i, v) in enumerate(vs), w in nbrs[v]
        D = edge_dim[eidx[(v, w)] + 1]
        slots[(v, w)] = (off + 1):(off + D * D)
        dims[(v, w)] = D
        off += D * D
    end
    ctx = DeviceContext(h[], nn, vs, slots, dims, off, T)
    finalizer(c -> ccall((:bp_destroy, libbp[]), Cint, (Ptr{Cvoid},), c.handle), ctx)
    return ctx
end

"""The names of the message of edge v => w in the engine's axis order [ket][bra]: the ket
link name the edge's tensors share and its bra partner (`braname`,
src/normnetwork.jl:74-81).  The reference builds each message with named bra and ket axes
(src/beliefpropagation/messagecache.jl:205-225); permuting by these names, not by the
message's storage order, keeps the orientation right whatever order `similar_operator`
chose."""
function message_names(nn::NormNetwork, v, w)
    k = name(only(linkinds(nn.ket, (v, w))))
    return (k, braname(nn, k))
end

# the message of edge v => w as a D x D matrix of element type T indexed [ket, bra]
function ketbra_matrix(nn::NormNetwork, cache, v, w, T::DataType = Float64)
    m = cache[v => w]
    return Array{T}(unnamed(m, message_names(nn, v, w)...))
end

"""Messages [ket][bra] as row-major D x D blocks of the flat buffer (the message of
edge v => w is produced at v and consumed at w)."""
function upload!(ctx::DeviceContext, cache)
    flat = zeros(ctx.T, ctx.nmsg)
    for (e, r) in ctx.slots
        m = ketbra_matrix(ctx.nn, cache, e[1], e[2], ctx.T)
        flat[r] = vec(permutedims(m, (2, 1)))  # row-major [ket][bra]
    end
    GC.@preserve flat check(ccall((:bp_init_messages, libbp[]), Cint,
        (Ptr{Cvoid}, Cint, Ptr{Cvoid}), ctx.handle, BP_INIT_UPLOAD, flat), "bp_init_messages")
    return ctx
end

function download!(cache, ctx::DeviceContext)
    flat = zeros(ctx.T, ctx.nmsg)
    GC.@preserve flat check(ccall((:bp_download_messages, libbp[]), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}), ctx.handle, flat), "bp_download_messages")
    for (e, r) in ctx.slots
        D = ctx.dims[e]
        msg = cache[e[1] => e[2]]
        # [ket, bra] matrix, written through the message's own axis names
        copyto!(unnamed(msg, message_names(ctx.nn, e[1], e[2])...),
            permutedims(reshape(flat[r], D, D), (2, 1)))
    end
    return cache
end

struct JacobiSweep end

Base.@kwdef struct DeviceJacobiUpdate <: MessageUpdateAlgorithm
    dtype::Symbol = :f32
    path::Symbol = :auto
    polish::Int = 0
end

const CONTEXTS = IdDict{Any, DeviceContext}()

context(alg::DeviceJacobiUpdate, nn) = get!(() -> DeviceContext(nn; alg.dtype), CONTEXTS, nn)

# One device Jacobi sweep per outer iteration: `edges = [JacobiSweep()]`.
function message_update!(alg::DeviceJacobiUpdate, cache, factors::NormNetwork, ::JacobiSweep)
    ctx = context(alg, factors)
    upload!(ctx, cache)
    check(ccall((:bp_sweep, libbp[]), Cint, (Ptr{Cvoid},), ctx.handle), "bp_sweep")
    download!(cache, ctx)
    return cache
end

"""
    DeviceMessageCache(ctx)

Messages that stay in the device context: `message_update!(alg, cache::DeviceMessageCache,
factors, ::JacobiSweep)` sweeps without moving a message over PCIe; `copy` takes the host
snapshot that `StopWhenConverged` keeps as its previous iterate
(AlgorithmsInterfaceExtensions.jl:74, 96); `AIE.iterate_diff` and `bethe_free_energy`
evaluate on the device (the methods the reference's generic code calls,
beliefpropagation.jl:261-267, messagecache.jl:185-201).
"""
struct DeviceMessageCache
    ctx::DeviceContext
end

"""A host snapshot of a device cache's messages (flat, the engine's slot layout)."""
struct DeviceSnapshot{T}
    flat::Vector{T}
end

function Base.copy(c::DeviceMessageCache)
    flat = zeros(c.ctx.T, c.ctx.nmsg)
    GC.@preserve flat check(ccall((:bp_download_messages, libbp[]), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}), c.ctx.handle, flat), "bp_download_messages")
    return DeviceSnapshot(flat)
end

# iterate_diff of the device messages against the previous iterate's snapshot
function AIE.iterate_diff(c::DeviceMessageCache, prev::DeviceSnapshot)
    out = Ref{Float64}(0.0)
    GC.@preserve prev check(ccall((:bp_iterate_diff, libbp[]), Cint,
        (Ptr{Cvoid}, Ptr{Cvoid}, Ref{Float64}), c.ctx.handle, prev.flat, out),
        "bp_iterate_diff")
    return out[]
end

bethe_free_energy(::NormNetwork, c::DeviceMessageCache) = device_lognorm(c.ctx)

"""A device cache for `nn` holding `messages` (uploaded once)."""
function device_messagecache(alg::DeviceJacobiUpdate, nn::NormNetwork, messages)
    ctx = context(alg, nn)
    upload!(ctx, MessageCache(messages))
    return DeviceMessageCache(ctx)
end

function message_update!(::DeviceJacobiUpdate, cache::DeviceMessageCache, ::NormNetwork,
        ::JacobiSweep)
    check(ccall((:bp_sweep, libbp[]), Cint, (Ptr{Cvoid},), cache.ctx.handle), "bp_sweep")
    return cache
end

"""Whole solve on the device with the reference's stopping semantics."""
function device_beliefpropagation(nn::NormNetwork, messages; stopping_criterion,
        alg::DeviceJacobiUpdate = DeviceJacobiUpdate())
    ctx = context(alg, nn)
    cache = MessageCache(messages)
    upload!(ctx, cache)
    maxiter = get(stopping_criterion, :maxiter, 0)
    tol = get(stopping_criterion, :tol, -1.0)
    sweeps = Ref{Cint}(0)
    delta = Ref{Float64}(0.0)
    at = Ref{Cint}(-1)
    check(ccall((:bp_run, libbp[]), Cint,
            (Ptr{Cvoid}, Cint, Cdouble, Cint, Cint, Ref{Cint}, Ref{Float64}, Ref{Cint}),
            ctx.handle, maxiter, tol, BP_CRIT_FIDELITY, 1, sweeps, delta, at), "bp_run")
    if alg.polish > 0 && !(ctx.T <: Complex)  # complex contexts compute in FP64 already
        check(ccall((:bp_polish_fp64, libbp[]), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}),
            ctx.handle, alg.polish, C_NULL), "bp_polish_fp64")
    end
    download!(cache, ctx)
    return cache, (; sweeps = sweeps[], delta = delta[], at_iteration = at[])
end

"""BP log-norm of the device's current messages (bethe_free_energy)."""
function device_lognorm(ctx::DeviceContext)
    re = Ref{Float64}(0.0)
    im = Ref{Float64}(0.0)
    check(ccall((:bp_lognorm, libbp[]), Cint, (Ptr{Cvoid}, Ref{Float64}, Ref{Float64}),
        ctx.handle, re, im), "bp_lognorm")
    # complex element types give a complex log-norm, as sum(log.(terms)) does
    return (ctx.T <: Complex || im[] != 0) ? complex(re[], im[]) : re[]
end

"""Gradient of the BP log-norm with respect to every site tensor (flat, site layout)."""
function device_grad_lognorm(ctx::DeviceContext, nsite::Integer)
    g = zeros(Float64, nsite)
    GC.@preserve g check(ccall((:bp_grad_lognorm, libbp[]), Cint,
        (Ptr{Cvoid}, Ptr{Float64}), ctx.handle, g), "bp_grad_lognorm")
    return g
end

"""BP-estimated TFIM energy and its gradient by the device's fixed-point adjoint."""
function device_tfim_energy_and_grad(ctx::DeviceContext, nsite::Integer; J = 1.0, h = 3.0,
        adj_tol = 1.0e-8, adj_maxiter = 1000)
    E = Ref{Float64}(0.0)
    g = zeros(Float64, nsite)
    it = Ref{Cint}(0)
    res = Ref{Float64}(0.0)
    GC.@preserve g check(ccall((:bp_tfim_energy, libbp[]), Cint,
        (Ptr{Cvoid}, Cdouble, Cdouble, Ref{Float64}, Ptr{Float64}, Cdouble, Cint, Ref{Cint},
            Ref{Float64}), ctx.handle, J, h, E, g, adj_tol, adj_maxiter, it, res),
        "bp_tfim_energy")
    return E[], g
end

"""
    device_tfim_energy_and_grad!(ctx, grad_ptr; J, h, adj_tol, adj_maxiter)

The same with the gradient written into device memory at `grad_ptr` (`nsite` entries of
Float32 for :f32/:bf16 contexts, Float64 for :f64), e.g. `pointer(g)` of an AMDGPU.jl
`ROCArray` (ROCm's HIP runtime, the one the library links).  Nothing crosses PCIe.
"""
function device_tfim_energy_and_grad!(ctx::DeviceContext, grad_ptr::Ptr{Cvoid}; J = 1.0,
        h = 3.0, adj_tol = 1.0e-8, adj_maxiter = 1000)
    E = Ref{Float64}(0.0)
    it = Ref{Cint}(0)
    res = Ref{Float64}(0.0)
    check(ccall((:bp_tfim_energy_device, libbp[]), Cint,
        (Ptr{Cvoid}, Cdouble, Cdouble, Ref{Float64}, Ptr{Cvoid}, Cdouble, Cint, Ref{Cint},
            Ref{Float64}), ctx.handle, J, h, E, grad_ptr, adj_tol, adj_maxiter, it, res),
        "bp_tfim_energy_device")
    return E[]
end

# operators as an ld x ld x n array (Julia column-major): block i, row a, column b is
# op[a, b, i]; the C side reads row-major ld x ld blocks, so pass the transposes
_rowmajor(ops::AbstractArray{Float64, 3}) = permutedims(ops, (2, 1, 3))

"""
    device_expect_vertex(ctx, ops) -> Vector{Float64}

`<O_v>_BP` for every vertex (topology order), `ops[:, :, v]` a `d x d` operator on the
physical index of the ket layer (`ld x ld` blocks, top-left corner used).
"""
function device_expect_vertex(ctx::DeviceContext, ops::AbstractArray{Float64, 3})
    rm = _rowmajor(ops)
    out = zeros(Float64, size(ops, 3))
    GC.@preserve rm out check(ccall((:bp_expect_vertex, libbp[]), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Cint, Ptr{Float64}), ctx.handle, rm, size(ops, 1), out),
        "bp_expect_vertex")
    return out
end

"""`<A_e B_e>_BP` for every topology edge; `A[:, :, e]` acts on the vertex owning the edge's
lower message slot, `B[:, :, e]` on the other end."""
function device_expect_edge(ctx::DeviceContext, A::AbstractArray{Float64, 3},
        B::AbstractArray{Float64, 3})
    ra, rb = _rowmajor(A), _rowmajor(B)
    out = zeros(Float64, size(A, 3))
    GC.@preserve ra rb out check(ccall((:bp_expect_edge, libbp[]), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Float64}), ctx.handle, ra, rb,
        size(A, 1), out), "bp_expect_edge")
    return out
end

"""Local energy `sum_v <O_v> + sum_e <A_e B_e>` and its gradient (flat site layout) by the
device's fixed-point adjoint."""
function device_local_energy_and_grad(ctx::DeviceContext, nsite::Integer, O, A, B;
        adj_tol = 1.0e-8, adj_maxiter = 1000)
    ro, ra, rb = _rowmajor(O), _rowmajor(A), _rowmajor(B)
    E = Ref{Float64}(0.0)
    g = zeros(Float64, nsite)
    it = Ref{Cint}(0)
    res = Ref{Float64}(0.0)
    GC.@preserve ro ra rb g check(ccall((:bp_local_energy, libbp[]), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint, Ref{Float64},
            Ptr{Float64}, Cdouble, Cint, Ref{Cint}, Ref{Float64}), ctx.handle, ro, ra, rb,
        size(O, 1), E, g, adj_tol, adj_maxiter, it, res), "bp_local_energy")
    return E[], g
end

const BP_CN_ORDER = Dict(:left_associative => Cint(0), :greedy => Cint(1), :cheapest => Cint(2))

"""Exact contraction on the device (`contract_network(tn; alg = Exact())`,
src/contract_network.jl:32-37): `tensors[t]` carries the 0-based integer labels
`labels[t]` (in Julia index order; a label on two tensors is summed over). Returns the
result array and its open labels in result order. Julia arrays are column-major, so the
labels go to the row-major C side reversed, and the result comes back reversed."""
function device_contract_network(ctx::DeviceContext, tensors::AbstractVector,
        labels::AbstractVector{<:AbstractVector{<:Integer}}; order_alg::Symbol = :cheapest)
    nl = maximum(Iterators.flatten(labels); init = -1) + 1
    dims = zeros(Cint, max(nl, 1))
    for (t, l) in zip(tensors, labels), (d, x) in zip(size(t), l)
        dims[x + 1] = d
    end
    ranks = Cint[length(l) for l in labels]
    labs = Cint[x for l in labels for x in reverse(l)]
    data = isempty(tensors) ? zeros(Float64, 1) :
        reduce(vcat, [vec(Array{Float64}(t)) for t in tensors])
    counts = zeros(Int, max(nl, 1))
    foreach(x -> counts[x + 1] += 1, Iterators.flatten(labels))
    cap = prod(Int[dims[i] for i in 1:nl if counts[i] == 1]; init = 1)
    orank = Ref{Cint}(0)
    olabs = zeros(Cint, 64)
    out = zeros(Float64, cap)
    GC.@preserve ranks labs dims data olabs out check(ccall((:bp_contract_network, libbp[]),
        Cint, (Ptr{Cvoid}, Cint, Ptr{Cint}, Ptr{Cint}, Cint, Ptr{Cint}, Ptr{Float64},
            Ptr{Cint}, Cint, Clonglong, Ref{Cint}, Ptr{Cint}, Ptr{Float64}, Clonglong),
        ctx.handle, length(tensors), ranks, labs, nl, dims, data, C_NULL,
        BP_CN_ORDER[order_alg], 0, orank, olabs, out, cap), "bp_contract_network")
    ol = reverse(Int.(olabs[1:orank[]]))
    return reshape(out, Tuple(Int(dims[x + 1]) for x in ol)), ol
end

"""Autotune the kernel choices for this context's shapes; `mode` 0 applies the cache at
`cache` only, 1 tunes what it lacks (and rewrites it), 2 retunes everything."""
function autotune!(ctx::DeviceContext; cache::AbstractString = "", mode::Integer = 1)
    n = Ref{Cint}(0)
    GC.@preserve cache begin
        path = isempty(cache) ? Ptr{UInt8}(C_NULL) : pointer(cache)
        check(ccall((:bp_autotune, libbp[]), Cint, (Ptr{Cvoid}, Ptr{UInt8}, Cint, Ref{Cint}),
            ctx.handle, path, mode, n), "bp_autotune")
    end
    return Int(n[])
end

"""
    prepare_run!(ctx, maxiter)

Capture the sweep graphs a later `maxiter`-sweep run on `ctx` replays, without running a
sweep (bp_prepare_run), so that loops timing or repeating runs start them at once.
"""
function prepare_run!(ctx::DeviceContext, maxiter::Integer)
    check(ccall((:bp_prepare_run, libbp[]), Cint, (Ptr{Cvoid}, Cint), ctx.handle, maxiter),
        "bp_prepare_run")
    return ctx
end

# ChainRulesCore rrules: the pullback of a converged-BP quantity with respect to the flat
# site buffer is its adjoint gradient (implicit differentiation of the fixed point).  Each
# rrule converges BP once and takes the value and the gradient from one library call.

function converge!(ctx::DeviceContext, sites::AbstractVector{Float64}; maxiter = 2000,
        tol = 1.0e-12)
    GC.@preserve sites check(ccall((:bp_upload_sites, libbp[]), Cint,
        (Ptr{Cvoid}, Ptr{Float64}), ctx.handle, sites), "bp_upload_sites")
    check(ccall((:bp_init_messages, libbp[]), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}),
        ctx.handle, Cint(0), C_NULL), "bp_init_messages")
    check(ccall((:bp_run, libbp[]), Cint,
            (Ptr{Cvoid}, Cint, Cdouble, Cint, Cint, Ptr{Cint}, Ptr{Float64}, Ptr{Cint}),
            ctx.handle, maxiter, tol, BP_CRIT_FIDELITY, 1, C_NULL, C_NULL, C_NULL), "bp_run")
    return ctx
end

function tfim_energy(ctx::DeviceContext, sites::AbstractVector{Float64}; J = 1.0, h = 3.0)
    converge!(ctx, sites)
    E = Ref{Float64}(0.0)
    it = Ref{Cint}(0)
    res = Ref{Float64}(0.0)
    check(ccall((:bp_tfim_energy, libbp[]), Cint,
        (Ptr{Cvoid}, Cdouble, Cdouble, Ref{Float64}, Ptr{Float64}, Cdouble, Cint, Ref{Cint},
            Ref{Float64}), ctx.handle, J, h, E, C_NULL, 1.0e-8, 1000, it, res),
        "bp_tfim_energy")
    return E[]
end

"""BP log-norm (bethe_free_energy) of the network with site buffer `sites`."""
function lognorm(ctx::DeviceContext, sites::AbstractVector{Float64})
    converge!(ctx, sites)
    return device_lognorm(ctx)
end

"""Local energy sum_v <O_v> + sum_e <A_e B_e> of the network with site buffer `sites`."""
function local_energy(ctx::DeviceContext, sites::AbstractVector{Float64}, O, A, B)
    converge!(ctx, sites)
    ro, ra, rb = _rowmajor(O), _rowmajor(A), _rowmajor(B)
    E = Ref{Float64}(0.0)
    it = Ref{Cint}(0)
    res = Ref{Float64}(0.0)
    GC.@preserve ro ra rb check(ccall((:bp_local_energy, libbp[]), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Cint, Ref{Float64},
            Ptr{Float64}, Cdouble, Cint, Ref{Cint}, Ref{Float64}), ctx.handle, ro, ra, rb,
        size(O, 1), E, C_NULL, 1.0e-8, 1000, it, res), "bp_local_energy")
    return E[]
end

using ChainRulesCore: ChainRulesCore, NoTangent

function ChainRulesCore.rrule(::typeof(tfim_energy), ctx::DeviceContext,
        sites::AbstractVector{Float64}; J = 1.0, h = 3.0)
    converge!(ctx, sites)
    E, g = device_tfim_energy_and_grad(ctx, length(sites); J, h)  # value and gradient at once
    pullback(ΔE) = (NoTangent(), NoTangent(), ΔE .* g)
    return E, pullback
end

function ChainRulesCore.rrule(::typeof(lognorm), ctx::DeviceContext,
        sites::AbstractVector{Float64})
    converge!(ctx, sites)
    L = device_lognorm(ctx)
    g = device_grad_lognorm(ctx, length(sites))
    pullback(ΔL) = (NoTangent(), NoTangent(), ΔL .* g)
    return L, pullback
end

function ChainRulesCore.rrule(::typeof(local_energy), ctx::DeviceContext,
        sites::AbstractVector{Float64}, O, A, B)
    converge!(ctx, sites)
    E, g = device_local_energy_and_grad(ctx, length(sites), O, A, B)
    pullback(ΔE) = (NoTangent(), NoTangent(), ΔE .* g, NoTangent(), NoTangent(), NoTangent())
    return E, pullback
end

end # module
