# CATNHip — Julia binding of libcatn_hip (include/catn_hip.h) for CATN.jl users.
#
# Replaces the reference's GPU dispatch (TensorNetwork{T,AT} with AT = CuArray,
# src/tensor_network.jl:1, src/adapt.jl:6-24) for the contraction path: networks
# live on the MI355X and contraction!, free_energy, magnetization and correlation
# run through the library. ising_lnZ carries a ChainRulesCore rrule whose
# pullback is the library's hand-written reverse sweep (catn_gradient_batch +
# catn_ising_gradient), so Zygote / ChainRules users get d lnZ / d(w, h, beta)
# from one backward pass.
#
# No Julia toolchain exists on the build image, so this file is not executed
# by the test suite; every ccall signature mirrors include/catn_hip.h, and the
# Python mirror (catn.jl_amd/__init__.py) exercises the same entry points.
module CATNHip

using ChainRulesCore

export ising_network, tensor_network, braket_network, contraction!, free_energy,
       magnetization, correlation, ising_lnZ, ising_lnZ_batch, result_tensor, set_launch!,
       set_batch_streams!, prepare_gradient!, tensor_lnZ

const lib = get(ENV, "CATN_HIP_LIB", joinpath(@__DIR__, "..", "catn.jl_amd", "lib", "libcatn_hip.so"))

# layout of catn_params, field for field (include/catn_hip.h)
struct Params
    Dmax::Int32
    chi::Int32
    cutoff::Float64
    norm_method::Int32
    select::Int32
    reverse::Int32
    svdopt::Int32
    swapopt::Int32
    compress::Int32
    cut_bond::Int32
    seed::Int32
    arena_elems::Int64
    scratch_elems::Int64
    rsvd_both::Int32       # swap!'s randomized-SVD switch (src/mps_node.jl:154)
    rsvd_any::Int32
    precision::Int32       # 0 Float64; 1 Float32 storage, 3-pass BF16; 2 Float32, 6-pass
    loops::Int32           # braket_network only: 1 accepts cyclic bond graphs
    braket_dmax::Int32     # braket_network only: merged-bond Dmax (< 0: none)
    eltype::Int32          # 0 real, 1 complex (ComplexF64; ComplexF32 with precision 1)
end

lasterr() = unsafe_string(ccall((:catn_last_error, lib), Cstring, ()))
check(rc) = rc == 0 || error("libcatn_hip: " * lasterr())

const CTX = Ref{Ptr{Cvoid}}(C_NULL)
function ctx()
    if CTX[] == C_NULL
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:catn_create, lib), Cint, (Ref{Ptr{Cvoid}}, Cint), r, 0))
        CTX[] = r[]
    end
    CTX[]
end

function params(; Dmax=32, chi=32, cutoff=1e-15, norm_method=1, select=1, reverse=true,
                svdopt=true, swapopt=true, compress=false, cut_bond=false, seed=1,
                arena_elems=0, scratch_elems=0, rsvd_both=7000, rsvd_any=20000,
                loops=false, braket_dmax=-1, T::Type=Float64)
    prec = T === Float64 || T === ComplexF64 ? 0 : T === Float32 || T === ComplexF32 ? 1 :
           error("element type $T is not supported")
    elt = T <: Complex ? 1 : 0
    Params(Dmax, chi, cutoff, norm_method, select, reverse, svdopt, swapopt, compress,
           cut_bond, seed, arena_elems, scratch_elems, rsvd_both, rsvd_any, prec, loops,
           braket_dmax, elt)
end

# stands in for TensorNetwork{T,AT}: the library owns the device state
mutable struct HipNetwork
    ptr::Ptr{Cvoid}
    n::Int
    beta::Float64
    nedges::Int
    T::Type                # element type: Float64, Float32 or ComplexF64
end
function wrap(ptr, n, beta, nedges, T=Float64)
    tn = HipNetwork(ptr, n, beta, nedges, T)
    finalizer(t -> ccall((:catn_net_free, lib), Cint, (Ptr{Cvoid},), t.ptr), tn)
end
# a phase output of the library: 1 double, or (re, im) for complex networks
phase(buf, T) = T <: Complex ? complex(buf[1], buf[2]) : buf[1]

# ising_network (src/ising.jl:20-227)
function ising_network(n, edges, w, h, β; pos1=0, val1=0, pos2=0, val2=0, kw...)
    e = Int32[x for (i, j) in edges for x in (i, j)]
    p = Ref(params(; kw...))
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:catn_ising_network, lib), Cint,
                (Ptr{Cvoid}, Int32, Int32, Ptr{Int32}, Ptr{Float64}, Ptr{Float64}, Float64,
                 Int32, Int32, Int32, Int32, Ref{Params}, Ref{Ptr{Cvoid}}),
                ctx(), n, length(edges), e, Float64.(collect(w)), Float64.(collect(h)), β,
                pos1, val1, pos2, val2, p, out))
    wrap(out[], n, Float64(β), length(edges))
end

# labels -> Int64 ids in order of first appearance (open labels get the
# device's sentinels -1, -2, ... in that order)
function _flatten(tensors, ixs)
    ids = Dict{Any,Int64}()
    labels = Int64[]
    for ix in ixs, l in ix
        push!(labels, get!(ids, l, length(ids)))
    end
    nds = Int32[ndims(t) for t in tensors]
    dims = Int64[d for t in tensors for d in size(t)]
    # complex tensors go over as interleaved (re, im) doubles: ComplexF64's
    # layout; Float32 / ComplexF32 inputs select FP32 storage on the device
    cx = any(t -> eltype(t) <: Complex, tensors)
    f32 = !isempty(tensors) && all(t -> real(eltype(t)) === Float32, tensors)
    W = cx ? ComplexF64 : Float64
    data = reduce(vcat, [vec(W.(t)) for t in tensors]; init=W[])
    T = cx ? (f32 ? ComplexF32 : ComplexF64) : (f32 ? Float32 : Float64)
    nds, dims, reinterpret(Float64, data) |> collect, labels, T
end

# TensorNetwork(tensors, ixs; kw...) (src/tensor_network.jl:41-140)
function tensor_network(tensors, ixs; kw...)
    nd, dm, dt, lb, T = _flatten(tensors, ixs)
    p = Ref(params(; T=T, kw...))
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:catn_tensor_network, lib), Cint,
                (Ptr{Cvoid}, Int32, Ptr{Int32}, Ptr{Int64}, Ptr{Float64}, Ptr{Int64},
                 Ref{Params}, Ref{Ptr{Cvoid}}), ctx(), length(tensors), nd, dm, dt, lb, p, out))
    wrap(out[], length(tensors), 1.0, 0, T)
end

# braket_network(tensors, ixs; chi, cutoff, ...) (src/braket.jl:369-496), real or
# complex tensors
function braket_network(tensors, ixs; chi=64, loops=false, Dmax=-1, kw...)
    nd, dm, dt, lb, T = _flatten(tensors, ixs)
    p = Ref(params(; chi=chi, loops=loops, braket_dmax=Dmax, T=T, kw...))
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:catn_braket_network, lib), Cint,
                (Ptr{Cvoid}, Int32, Ptr{Int32}, Ptr{Int64}, Ptr{Float64}, Ptr{Int64},
                 Ref{Params}, Ref{Ptr{Cvoid}}), ctx(), length(tensors), nd, dm, dt, lb, p, out))
    wrap(out[], length(tensors), 1.0, 0, T)
end

# contraction!(tn) (src/tensor_network.jl:556) -> (lnZ, error, psi); psi::T
function contraction!(tn::HipNetwork)
    lz, err, psi = Ref(0.0), Ref(0.0), zeros(2)
    check(ccall((:catn_contraction, lib), Cint,
                (Ptr{Cvoid}, Ref{Float64}, Ref{Float64}, Ptr{Float64}), tn.ptr, lz, err, psi))
    (lz[], err[], phase(psi, tn.T))
end

# batched contraction! of many networks in one persistent launch per segment
# (one element type per batch)
function contraction!(nets::Vector{HipNetwork})
    k = length(nets)
    T = isempty(nets) ? Float64 : nets[1].T
    w = T <: Complex ? 2 : 1
    lz, err, psi = zeros(k), zeros(k), zeros(w * k)
    check(ccall((:catn_contraction_batch, lib), Cint,
                (Ptr{Cvoid}, Int32, Ptr{Ptr{Cvoid}}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                ctx(), k, [t.ptr for t in nets], lz, err, psi))
    lz, err, (w == 2 ? reinterpret(ComplexF64, psi) |> collect : psi)
end

# free_energy (src/ising.jl:239-244) -> (lnZ/n, -lnZ/(n β))
function free_energy(tn::HipNetwork)
    lnZ, _, _ = contraction!(tn)
    (lnZ / tn.n, -lnZ / (tn.n * tn.beta))
end

# result_tensor (src/tensor_network.jl:517-523), after contraction!
function result_tensor(tn::HipNetwork)
    rank = Ref{Int32}(0); dims = zeros(Int32, 64); labs = zeros(Int32, 64)
    check(ccall((:catn_result_tensor, lib), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Int64, Ref{Int32}, Ptr{Int32}, Ptr{Int32}, Int32),
                tn.ptr, C_NULL, 0, rank, dims, labs, 64))
    T = tn.T <: Complex ? ComplexF64 : Float64
    rank[] == 0 && return fill(one(T))
    out = zeros(T, Tuple(Int.(dims[1:rank[]])))
    buf = reinterpret(Float64, vec(out))     # ComplexF64 as (re, im) pairs
    check(ccall((:catn_result_tensor, lib), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Int64, Ref{Int32}, Ptr{Int32}, Ptr{Int32}, Int32),
                tn.ptr, buf, length(buf), rank, dims, labs, 64))
    out
end

# ---------------------------------------------------------------------------
# chosen form: lnZ and its gradient from one recorded forward and one reverse
# sweep per batch (catn_gradient_batch), read back per network as
# d lnZ / d w (one per edge), d lnZ / d h (one per spin), d lnZ / d beta
function lnz_and_grad!(nets::Vector{HipNetwork})
    k = length(nets)
    lz, err, psi = zeros(k), zeros(k), zeros(k)
    check(ccall((:catn_gradient_batch, lib), Cint,
                (Ptr{Cvoid}, Int32, Ptr{Ptr{Cvoid}}, Int32, Ptr{Float64}, Ptr{Float64},
                 Ptr{Float64}), ctx(), k, [t.ptr for t in nets], 0, lz, err, psi))
    map(enumerate(nets)) do (q, t)
        dw, dh, db = zeros(t.nedges), zeros(t.n), Ref(0.0)
        check(ccall((:catn_ising_gradient, lib), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ref{Float64}), t.ptr, dw, dh, db))
        (lz[q], dw, dh, db[])
    end
end

# ---------------------------------------------------------------------------
# Lifted gradient. The reverse sweep differentiates the executed schedule (the
# kept count of every split held fixed). A frustrated cycle of +-J couplings can
# make a split exactly rank-deficient at w: the cutoff drops a zero singular
# value that any perturbation of w restores, and the exact lnZ then has a
# derivative the schedule at w does not carry. Each instance is run at
# w +- eta for a fixed generic eta (splitmix64 of the index, the sequence of the
# Python mirror's lift_direction), which lifts such zeros above the cutoff; the
# average is the exact gradient to O(|eta|^2). FP32 storage needs a larger lift
# to clear its noise floor and adds a pair at 2 eta (Richardson step).
const LIFT = Dict(0 => 1e-6, 2 => 2e-3, 1 => 1e-2)     # by precision level

function splitmix(seed::UInt64, i::Integer)
    z = seed + UInt64(i) * 0x9E3779B97F4A7C15         # UInt64 arithmetic wraps
    z = (z ⊻ (z >> 30)) * 0xBF58476D1CE4E5B9
    z = (z ⊻ (z >> 27)) * 0x94D049BB133111EB
    z ⊻ (z >> 31)
end
lift_direction(ne, n; seed=0x5EED) =
    [Float64(splitmix(UInt64(seed), i) >> 11) * 2.0^-52 - 1.0 for i in 1:(ne + n)]

# [(lnZ, dw, dh, dβ) per instance] for instances (n, edges, w, h, β); lift =
# nothing takes the precision's default, 0 the executed schedule's derivative
function lifted_grads(instances; lift=nothing, kw...)
    prec = Int(params(; kw...).precision)
    s0 = lift === nothing ? LIFT[prec] : Float64(lift)
    s0 >= 0 || error("lift must be >= 0, got $lift")
    s0 == 0 && return lnz_and_grad!([ising_network(i...; kw...) for i in instances])
    lifts = prec == 0 ? [s0] : [s0, 2s0]
    nets, etas = HipNetwork[], Tuple{Vector{Float64},Vector{Float64}}[]
    for (n, edges, w, h, β) in instances
        w, h = Float64.(collect(w)), Float64.(collect(h))
        g = lift_direction(length(edges), n)
        sc = max(1.0, maximum(abs, w; init=0.0), maximum(abs, h; init=0.0))
        for s in lifts
            ew, eh = (s * sc) .* g[1:length(edges)], (s * sc) .* g[length(edges)+1:end]
            push!(nets, ising_network(n, edges, w .+ ew, h .+ eh, β; kw...))
            push!(nets, ising_network(n, edges, w .- ew, h .- eh, β; kw...))
            push!(etas, (ew, eh))
        end
    end
    res = lnz_and_grad!(nets)
    out = Tuple{Float64,Vector{Float64},Vector{Float64},Float64}[]
    q = 0
    for _ in instances
        pairs = []
        for _ in lifts
            (lp, wp, hp, bp), (lm, wm, hm, bm) = res[q + 1], res[q + 2]
            ew, eh = etas[q ÷ 2 + 1]
            q += 2
            lz = (lp + lm) / 2 - (sum(ew .* (wp .- wm)) + sum(eh .* (hp .- hm))) / 4
            push!(pairs, (lz, (wp .+ wm) ./ 2, (hp .+ hm) ./ 2, (bp + bm) / 2))
        end
        if length(pairs) == 1
            push!(out, pairs[1])
        else
            a, b = pairs
            push!(out, (a[1], (4 .* a[2] .- b[2]) ./ 3, (4 .* a[3] .- b[3]) ./ 3, (4a[4] - b[4]) / 3))
        end
    end
    out
end

"""
    ising_lnZ(n, edges, w, h, β; lift=nothing, kw...) -> lnZ

lnZ of ising_network(n, edges, w, h, β; kw...) by contraction! on the device.
Differentiable: the rrule's pullback is the library's reverse sweep, giving
d lnZ/d w = β<s_i s_j>, d lnZ/d h = β<s_i> and d lnZ/d β = <E>, lifted (see
lifted_grads) so that exactly rank-deficient splits of ±J instances get the
exact derivative. The library refuses a gradient run whose schedule reaches a
Dmax cut on the gfx950 device (an error is raised); use Dmax = -1 or the
pinning route.
"""
ising_lnZ(n, edges, w, h, β; lift=nothing, kw...) =
    contraction!(ising_network(n, edges, w, h, β; kw...))[1]

function ChainRulesCore.rrule(::typeof(ising_lnZ), n, edges, w, h, β; lift=nothing, kw...)
    lz, dw, dh, db = lifted_grads([(n, edges, w, h, β)]; lift=lift, kw...)[1]
    function ising_lnZ_pullback(Δ)
        δ = unthunk(Δ)
        (NoTangent(), NoTangent(), NoTangent(), δ .* dw, δ .* dh, δ * db)
    end
    lz, ising_lnZ_pullback
end

"""
    ising_lnZ_batch(instances; lift=nothing, kw...) -> Vector of lnZ

Batched form: `instances` is a vector of (n, edges, w, h, β); every instance's
forward and reverse sweep (both lifted runs) run in the same launches.
"""
ising_lnZ_batch(instances; lift=nothing, kw...) =
    first.(contraction!([ising_network(i...; kw...) for i in instances]))

function ChainRulesCore.rrule(::typeof(ising_lnZ_batch), instances; lift=nothing, kw...)
    res = lifted_grads(instances; lift=lift, kw...)
    function ising_lnZ_batch_pullback(Δ)
        δ = unthunk(Δ)
        grads = [Tangent{Tuple}(NoTangent(), NoTangent(), δ[q] .* r[2], δ[q] .* r[3], δ[q] * r[4])
                 for (q, r) in enumerate(res)]
        (NoTangent(), grads)
    end
    [r[1] for r in res], ising_lnZ_batch_pullback
end

"""
    tensor_lnZ(tensors, ixs; kw...) -> ln|Z|

ln|Z| of a real tensor network (`TensorNetwork(tensors, ixs)`, then
`contraction!`). Its rrule's pullback is the reverse sweep through the
contraction and the raw2mps splits (catn_gradient_batch +
catn_net_tensor_adjoint): d ln|Z| / d T_k = environment_k / Z per tensor.
"""
tensor_lnZ(tensors, ixs; kw...) = first(contraction!(tensor_network(tensors, ixs; kw...)))

function tensor_lnZ_and_grads(tensors, ixs; kw...)
    tn = tensor_network(tensors, ixs; kw...)
    lz, err, psi = zeros(1), zeros(1), zeros(1)
    check(ccall((:catn_gradient_batch, lib), Cint,
                (Ptr{Cvoid}, Int32, Ptr{Ptr{Cvoid}}, Int32, Ptr{Float64}, Ptr{Float64},
                 Ptr{Float64}), ctx(), 1, [tn.ptr], 0, lz, err, psi))
    grads = map(enumerate(tensors)) do (k, t)
        g = zeros(Float64, size(t))
        check(ccall((:catn_net_tensor_adjoint, lib), Cint,
                    (Ptr{Cvoid}, Int32, Ptr{Float64}, Int64), tn.ptr, k - 1, g, length(g)))
        convert(Array{eltype(t)}, g)
    end
    lz[1], grads
end

function ChainRulesCore.rrule(::typeof(tensor_lnZ), tensors, ixs; kw...)
    lz, grads = tensor_lnZ_and_grads(tensors, ixs; kw...)
    tensor_lnZ_pullback(Δ) = (NoTangent(), [unthunk(Δ) .* g for g in grads], NoTangent())
    lz, tensor_lnZ_pullback
end

# magnetization / correlation (src/ising.jl:263-339). method = :pinning runs the
# reference's 1 + 2n / 1 + 4|E| contractions in one batch; method = :adjoint uses
# one reverse sweep (exact isolated-spin values, tanh(β h)).
function magnetization(n, edges, w, h, β; method=:pinning, lift=nothing, kw...)
    if method === :adjoint
        _, _, dh, _ = lifted_grads([(n, edges, w, h, β)]; lift=lift, kw...)[1]
        return dh ./ β
    end
    nets = [ising_network(n, edges, w, h, β; kw...)]
    for k in 1:n, v in (1, 2)
        push!(nets, ising_network(n, edges, w, h, β; pos1=k, val1=v, kw...))
    end
    lz, _, _ = contraction!(nets)
    # network 2k pins spin k to -1 (val1 = 1 zeroes the +1 value), 2k + 1 to +1
    [exp(lz[2k + 1] - lz[1]) - exp(lz[2k] - lz[1]) for k in 1:n]
end

function correlation(n, edges, w, h, β; method=:pinning, lift=nothing, kw...)
    if method === :adjoint
        _, dw, _, _ = lifted_grads([(n, edges, w, h, β)]; lift=lift, kw...)[1]
        return dw ./ β
    end
    nets = [ising_network(n, edges, w, h, β; kw...)]
    for (a, b) in edges, v1 in (1, 2), v2 in (1, 2)
        push!(nets, ising_network(n, edges, w, h, β; pos1=a, val1=v1, pos2=b, val2=v2, kw...))
    end
    lz, _, _ = contraction!(nets)
    map(1:length(edges)) do e
        p = exp.(lz[2 + 4(e - 1):1 + 4e] .- lz[1])
        p[1] - p[2] - p[3] + p[4]
    end
end

# launch shape: lanes per workgroup (64..1024), LDS doubles per workgroup, Householder panel
# width, Jacobi lanes per pair (0 = automatic / default)
set_launch!(threads=256, lds=0, panel=0, jacobi_lanes=0) =
    check(ccall((:catn_set_launch, lib), Cint, (Ptr{Cvoid}, Int32, Int32, Int32, Int32), ctx(),
                threads, lds, panel, jacobi_lanes))

# allocate a network's reverse-sweep memory (tape, adjoint arena) ahead of
# lnz_and_grad! (tape_cap 0: the library default)
prepare_gradient!(tn::HipNetwork, tape_cap=0) =
    check(ccall((:catn_net_prepare_gradient, lib), Cint, (Ptr{Cvoid}, Int32), tn.ptr, tape_cap))

# concurrent groups (streams) of a forward batch of distinct networks, 1..16 (default 4)
set_batch_streams!(streams=4) =
    check(ccall((:catn_set_batch_streams, lib), Cint, (Ptr{Cvoid}, Int32), ctx(), streams))

end # module
