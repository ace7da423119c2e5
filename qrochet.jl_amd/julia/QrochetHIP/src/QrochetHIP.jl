"""
    QrochetHIP

Device-resident drop-in for Qrochet's MPS update path.  `HIPChain(ψ::Chain)`
uploads an open `Chain` (State socket) to one MI355X; `evolve!`, `simple_update!`,
`canonize!`, `canonize_site!`, `mixed_canonize!`, `normalize!`, `truncate!`,
`contract!(…, :between, …)`, `expect`, `overlap` and `norm` keep their Qrochet
signatures and run in libqrochet_hip.so through `ccall`; `Chain(h::HIPChain)`
downloads the state again (the `adapt` round trip of ext/QrochetAdaptExt.jl).
"""
module QrochetHIP

using Qrochet
using Qrochet: Chain, Dense, Site, Operator, MPS, id, inputs, outputs, socket, nsites
using Qrochet: MissingSchmidtCoefficientsException
using Tenet
using LinearAlgebra

const libqh = get(ENV, "QROCHET_HIP_LIB", joinpath(@__DIR__, "..", "..", "..", "libqrochet_hip.so"))

const QH_OK = 0
const QH_ERR_ARGUMENT = -1
const QH_ERR_MISSING_SCHMIDT = -2
const QH_ERR_DIMENSION = -3
const QH_ERR_ASSERTION = -4
const QH_CANONICAL = UInt32(1)
const QH_RENORMALIZE = UInt32(2)
const QH_PREC = Dict(:fp64 => UInt32(0), :fp32 => UInt32(16), :bf16x3 => UInt32(32))
const QH_DIR = Dict(:left => Cint(0), :right => Cint(1))
const QH_METHOD = Dict(:qr => Cint(0), :svd => Cint(1))

mutable struct Context
    ptr::Ptr{Cvoid}
    function Context(device::Integer = 0)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        rc = ccall((:qh_ctx_create, libqh), Cint, (Cint, Ref{Ptr{Cvoid}}), device, r)
        rc == QH_OK || error("qh_ctx_create failed: ", unsafe_string(ccall((:qh_last_error, libqh), Cstring, (Ptr{Cvoid},), C_NULL)))
        ctx = new(r[])
        finalizer(c -> ccall((:qh_ctx_destroy, libqh), Cint, (Ptr{Cvoid},), c.ptr), ctx)
        ctx
    end
end

const DEFAULT = Ref{Union{Nothing,Context}}(nothing)
default_context() = (DEFAULT[] === nothing && (DEFAULT[] = Context(0)); DEFAULT[])

function check(ctx::Context, rc, bond = nothing)
    rc == QH_OK && return nothing
    msg = unsafe_string(ccall((:qh_last_error, libqh), Cstring, (Ptr{Cvoid},), ctx.ptr))
    rc == QH_ERR_ARGUMENT && throw(ArgumentError(msg))
    rc == QH_ERR_MISSING_SCHMIDT && throw(MissingSchmidtCoefficientsException(bond))
    rc == QH_ERR_DIMENSION && throw(DimensionMismatch(msg))
    rc == QH_ERR_ASSERTION && throw(AssertionError(msg))
    error(msg)
end

"""
    HIPChain <: Qrochet.Ansatz

Open-boundary MPS whose tensors live in device memory.
"""
mutable struct HIPChain <: Qrochet.Ansatz
    ctx::Context
    ptr::Ptr{Cvoid}
    n::Int
    d::Int
end

function wrap(ctx, ptr)
    n = Ref{Cint}(0); d = Ref{Cint}(0)
    ccall((:qh_mps_nsites, libqh), Cint, (Ptr{Cvoid}, Ref{Cint}, Ref{Cint}), ptr, n, d)
    h = HIPChain(ctx, ptr, n[], d[])
    finalizer(x -> ccall((:qh_mps_free, libqh), Cint, (Ptr{Cvoid},), x.ptr), h)
    h
end

# site tensor of ψ as an (o, l, r) ComplexF64 array (boundary tensors keep their 2 dims)
function olr_array(ψ::Chain, i::Int)
    t = Qrochet.tensors(ψ; at = Site(i))
    o = Qrochet.inds(ψ; at = Site(i))
    l = i == 1 ? nothing : Qrochet.leftindex(ψ, Site(i))
    r = i == nsites(ψ) ? nothing : Qrochet.rightindex(ψ, Site(i))
    order = filter(!isnothing, [o, l, r])
    ComplexF64.(permutedims(parent(t), map(x -> findfirst(==(x), inds(t)), order)))
end

"""
    HIPChain(ψ::Chain; context)

Upload (the `adapt` of ext/QrochetAdaptExt.jl:9).  Λ vectors present on bonds are uploaded too.
"""
function HIPChain(ψ::Chain; context::Context = default_context())
    n = nsites(ψ)
    arrays = [olr_array(ψ, i) for i in 1:n]
    d = size(arrays[1], 1)
    dims = Cint[size(arrays[i], ndims(arrays[i])) for i in 1:n-1]
    lams = Vector{Union{Nothing,Vector{Float64}}}(undef, n - 1)
    for b in 1:n-1
        λ = try
            Qrochet.tensors(ψ; between = (Site(b), Site(b + 1)))
        catch
            nothing
        end
        lams[b] = (λ isa Tenet.Tensor && ndims(λ) == 1) ? Float64.(real.(parent(λ))) : nothing
    end
    ptrs = [pointer(a) for a in arrays]
    lptrs = [l === nothing ? Ptr{Float64}(C_NULL) : pointer(l) for l in lams]
    out = Ref{Ptr{Cvoid}}(C_NULL)
    GC.@preserve arrays lams begin
        rc = ccall((:qh_mps_create, libqh), Cint,
                   (Ptr{Cvoid}, Cint, Cint, Ptr{Cint}, Ptr{Ptr{ComplexF64}}, Ptr{Ptr{Float64}}, Ref{Ptr{Cvoid}}),
                   context.ptr, n, d, dims, ptrs, lptrs, out)
    end
    check(context, rc)
    wrap(context, out[])
end

bond_dims(h::HIPChain) = (v = Vector{Cint}(undef, h.n - 1); ccall((:qh_mps_bond_dims, libqh), Cint, (Ptr{Cvoid}, Ptr{Cint}, Ptr{Cint}), h.ptr, v, C_NULL); Int.(v))

"""
    Chain(h::HIPChain)

Download to a host `Chain` (Λ vectors become separate tensors on their bonds).
"""
function Qrochet.Chain(h::HIPChain)
    dims = [1; bond_dims(h); 1]
    arrays = map(1:h.n) do i
        buf = Array{ComplexF64}(undef, h.d, dims[i], dims[i+1])
        check(h.ctx, ccall((:qh_mps_get_site, libqh), Cint, (Ptr{Cvoid}, Cint, Ptr{ComplexF64}), h.ptr, i, buf))
        i == 1 ? buf[:, 1, :] : i == h.n ? buf[:, :, 1] : buf
    end
    ψ = MPS(arrays)
    # re-attach Λ vectors as hyperindex tensors
    for b in 1:h.n-1
        λ = Vector{Float64}(undef, dims[b+1])
        rc = ccall((:qh_mps_get_lambda, libqh), Cint, (Ptr{Cvoid}, Cint, Ptr{Float64}), h.ptr, b, λ)
        rc == QH_ERR_MISSING_SCHMIDT && continue
        check(h.ctx, rc)
        push!(TensorNetwork(ψ), Tenet.Tensor(λ, [Qrochet.rightindex(ψ, Site(b))]))
    end
    ψ
end

"""
    rand(HIPChain, Open, State; n, χ, p = 2, eltype = Float64, seed = 0, context)

`rand(Chain, Open, State; n, χ, p, eltype)` (src/Ansatz/Chain.jl:223-256) sampled on the
device (`qh_mps_rand`): bond b of dimension min(χ, p^b, p^(n-b)), right-canonical from
site 2, norm 1.  `seed` replaces Julia's RNG.
"""
function Base.rand(::Type{HIPChain}, ::Type{Qrochet.Open}, ::Type{Qrochet.State}; n::Integer, χ::Integer, p::Integer = 2,
                   eltype::Type = Float64, seed::Integer = 0, context::Context = default_context())
    out = Ref{Ptr{Cvoid}}(C_NULL)
    check(context, ccall((:qh_mps_rand, libqh), Cint, (Ptr{Cvoid}, Cint, Cint, Cint, Cint, Culonglong, Ref{Ptr{Cvoid}}),
                         context.ptr, n, χ, p, eltype <: Complex ? 1 : 0, seed % Culonglong, out))
    wrap(context, out[])
end

"""
    autotune!(context; max_q = 512, batch = 32)

Pick the batched SVD's schedule (Jacobi block width, QR passes) per matrix size class by
measurement (`qh_ctx_autotune`); returns (q, block, passes, best_ms, previous_ms) per class.
"""
function autotune!(context::Context = default_context(); max_q::Integer = 512, batch::Integer = 32)
    rep = zeros(Float64, 40)
    n = Ref{Cint}(0)
    check(context, ccall((:qh_ctx_autotune, libqh), Cint, (Ptr{Cvoid}, Cint, Cint, Ptr{Float64}, Ref{Cint}),
                         context.ptr, max_q, batch, rep, n))
    [(q = Int(rep[5k+1]), block = Int(rep[5k+2]), passes = Int(rep[5k+3]), best_ms = rep[5k+4], previous_ms = rep[5k+5])
     for k in 0:n[]-1]
end

"Replay each Jacobi sweep's rounds as one captured HIP graph (`true`, default) or launch them one by one."
set_graphs!(context::Context, enable::Bool) =
    check(context, ccall((:qh_ctx_set_graphs, libqh), Cint, (Ptr{Cvoid}, Cint), context.ptr, enable ? 1 : 0))

"First QR pass of the SVD's preconditioner on the columns ordered by decreasing norm (`true`, default) or as they are."
set_pivoting!(context::Context, enable::Bool) =
    check(context, ccall((:qh_ctx_set_pivoting, libqh), Cint, (Ptr{Cvoid}, Cint), context.ptr, enable ? 1 : 0))

"`overlap(a, b)` without device allocation (`qh_overlap_noalloc`; open chains up to about χ = 1024)."
function overlap_noalloc(a::HIPChain, b::HIPChain)
    out = zeros(Float64, 2)
    check(a.ctx, ccall((:qh_overlap_noalloc, libqh), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}),
                       a.ctx.ptr, a.ptr, b.ptr, out))
    complex(out[1], out[2])
end

Base.copy(h::HIPChain) = (r = Ref{Ptr{Cvoid}}(C_NULL); check(h.ctx, ccall((:qh_mps_clone, libqh), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{Ptr{Cvoid}}), h.ctx.ptr, h.ptr, r)); wrap(h.ctx, r[]))
Base.deepcopy(h::HIPChain) = copy(h)

# gate array of a Dense operator, permuted to (outs..., ins...) with lanes ascending
function gate_buffer(gate::Dense)
    sites = collect(keys(Quantum(gate).sites))
    t = only(Tenet.tensors(TensorNetwork(gate)))
    lanes = sort!(unique(id.(sites)))
    order = vcat([Quantum(gate).sites[Site(l)] for l in lanes], [Quantum(gate).sites[Site(l; dual = true)] for l in lanes])
    ComplexF64.(permutedims(parent(t), map(x -> findfirst(==(x), inds(t)), order)))
end

"""
    evolve!(h::HIPChain, gate::Dense; threshold, maxdim, iscanonical, renormalize, precision = :fp64)

`evolve!` of src/Ansatz/Chain.jl:543-584 on the device.
"""
function Qrochet.evolve!(h::HIPChain, gate::Dense; threshold = nothing, maxdim = nothing, iscanonical = false,
                         renormalize = false, precision::Symbol = :fp64)
    socket(gate) isa Operator || throw(ArgumentError("Gate must be an operator, but got $(socket(gate))"))
    issetequal(adjoint.(inputs(gate)), outputs(gate)) ||
        throw(ArgumentError("Gate inputs ($(inputs(gate))) and outputs ($(outputs(gate))) must be the same"))
    lanes = sort!(map(id, outputs(gate)))
    all(l -> 1 <= l <= h.n, lanes) || throw(ArgumentError("Gate inputs must be a subset of the TN sites"))
    g = gate_buffer(gate)
    if length(lanes) == 1
        check(h.ctx, ccall((:qh_evolve1, libqh), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{ComplexF64}), h.ctx.ptr, h.ptr, lanes[1], g))
    elseif length(lanes) == 2
        lanes[2] == lanes[1] + 1 || throw(ArgumentError("Gate lanes must be contiguous"))
        flags = QH_PREC[precision] | (iscanonical ? QH_CANONICAL : UInt32(0)) | (renormalize ? QH_RENORMALIZE : UInt32(0))
        thr = threshold === nothing ? NaN : Float64(threshold)
        md = maxdim === nothing ? Cint(-1) : Cint(maxdim)
        check(h.ctx, ccall((:qh_evolve2, libqh), Cint,
                           (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{ComplexF64}, Cdouble, Cint, Cuint, Ptr{Cint}, Ptr{Cdouble}),
                           h.ctx.ptr, h.ptr, lanes[1], g, thr, md, flags, C_NULL, C_NULL), (Site(lanes[1]), Site(lanes[2])))
    else
        throw(ArgumentError("Invalid number of lanes $(length(lanes)), maximum is 2"))
    end
    h
end

simple_update!(h::HIPChain, gate::Dense; kwargs...) = Qrochet.evolve!(h, gate; iscanonical = true, kwargs...)

"""
    evolve_layer!(h::HIPChain, gates::Vector{<:Dense}; threshold, maxdim, iscanonical, renormalize, precision = :fp64)

A layer of gates on disjoint lanes: one-site gates with `evolve!`, all two-site
gates in one batched call (`qh_evolve2_batch`).  Same result as calling `evolve!`
gate by gate.
"""
function evolve_layer!(h::HIPChain, gates::AbstractVector{<:Dense}; threshold = nothing, maxdim = nothing,
                       iscanonical = false, renormalize = false, precision::Symbol = :fp64)
    two = Dense[]
    for g in gates
        length(outputs(g)) == 1 ? Qrochet.evolve!(h, g; precision) : push!(two, g)
    end
    isempty(two) && return h
    count = length(two)
    bonds = Cint[]
    buf = ComplexF64[]
    for g in two
        lanes = sort!(map(id, outputs(g)))
        (length(lanes) == 2 && lanes[2] == lanes[1] + 1) || throw(ArgumentError("Gate lanes must be contiguous"))
        push!(bonds, lanes[1])
        append!(buf, vec(gate_buffer(g)))
    end
    flags = QH_PREC[precision] | (iscanonical ? QH_CANONICAL : UInt32(0)) | (renormalize ? QH_RENORMALIZE : UInt32(0))
    thr = fill(threshold === nothing ? NaN : Float64(threshold), count)
    md = fill(maxdim === nothing ? Cint(-1) : Cint(maxdim), count)
    handles = fill(h.ptr, count)
    check(h.ctx, ccall((:qh_evolve2_batch, libqh), Cint,
                       (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Cint}, Ptr{ComplexF64}, Ptr{Cdouble}, Ptr{Cint}, Cuint,
                        Ptr{Cint}, Ptr{Cdouble}),
                       h.ctx.ptr, count, handles, bonds, buf, thr, md, flags, C_NULL, C_NULL))
    h
end

"""
    layers(gates, lanes_of)

Greedy as-soon-as-possible schedule of `gates` into layers of disjoint lanes that
keeps the order of the gates on every lane (`lanes_of(g)` gives a gate's lanes).
"""
function layers(gates, lanes_of)
    free = Dict{Int,Int}()
    out = Vector{Vector{eltype(gates)}}()
    for g in gates
        ls = lanes_of(g)
        k = maximum(l -> get(free, l, 1), ls)
        k > length(out) && push!(out, eltype(gates)[])
        push!(out[k], g)
        foreach(l -> free[l] = k + 1, ls)
    end
    out
end

function Qrochet.truncate!(h::HIPChain, bond; threshold = nothing, maxdim = nothing)
    b1, b2 = id.(bond)
    b2 == b1 + 1 || throw(ArgumentError("Invalid bond $bond"))
    thr = threshold === nothing ? NaN : Float64(threshold)
    md = maxdim === nothing ? Cint(-1) : Cint(maxdim)
    check(h.ctx, ccall((:qh_truncate, libqh), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cdouble, Cint), h.ctx.ptr, h.ptr, b1, thr, md), bond)
    h
end

Qrochet.canonize!(h::HIPChain) = (check(h.ctx, ccall((:qh_canonize, libqh), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), h.ctx.ptr, h.ptr)); h)
Qrochet.canonize(h::HIPChain) = Qrochet.canonize!(copy(h))

function Qrochet.canonize_site!(h::HIPChain, site::Site; direction::Symbol, method = :qr)
    check(h.ctx, ccall((:qh_canonize_site, libqh), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint),
                       h.ctx.ptr, h.ptr, id(site), get(QH_DIR, direction, Cint(-1)), get(QH_METHOD, method, Cint(-1))))
    h
end

Qrochet.mixed_canonize!(h::HIPChain, center::Site) =
    (check(h.ctx, ccall((:qh_mixed_canonize, libqh), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint), h.ctx.ptr, h.ptr, id(center))); h)

LinearAlgebra.normalize!(h::HIPChain, root::Site) =
    (check(h.ctx, ccall((:qh_normalize, libqh), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint), h.ctx.ptr, h.ptr, id(root))); h)

function Tenet.contract!(h::HIPChain, ::Val{:between}, site1::Site, site2::Site; direction::Symbol = :left, delete_Λ = true)
    check(h.ctx, ccall((:qh_contract_between, libqh), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cint, Cint),
                       h.ctx.ptr, h.ptr, id(site1), id(site2), get(QH_DIR, direction, Cint(-1)), delete_Λ))
    h
end
Tenet.contract!(h::HIPChain, q::Symbol, args...; kwargs...) = contract!(h, Val(q), args...; kwargs...)

struct Obs
    nsites::Cint
    site::Cint
    op::Ptr{ComplexF64}
end

function Qrochet.expect(h::HIPChain, observables)
    bufs = [gate_buffer(o) for o in observables]
    obs = [Obs(Cint(length(unique(id.(outputs(o))))), Cint(minimum(id.(outputs(o)))), pointer(b)) for (o, b) in zip(observables, bufs)]
    out = Vector{ComplexF64}(undef, 1)
    GC.@preserve bufs check(h.ctx, ccall((:qh_expect, libqh), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Obs}, Ptr{ComplexF64}),
                                         h.ctx.ptr, h.ptr, length(obs), obs, out))
    out[1]
end

function Qrochet.overlap(a::HIPChain, b::HIPChain)
    out = Vector{ComplexF64}(undef, 1)
    check(a.ctx, ccall((:qh_overlap, libqh), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{ComplexF64}), a.ctx.ptr, a.ptr, b.ptr, out))
    out[1]
end

function LinearAlgebra.norm(h::HIPChain, p::Real = 2)
    p == 2 || throw(ArgumentError("only L2-norm is implemented yet"))
    out = Ref{Cdouble}(0)
    check(h.ctx, ccall((:qh_norm, libqh), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{Cdouble}), h.ctx.ptr, h.ptr, out))
    out[]
end

for f in (:isleftcanonical, :isrightcanonical)
    cname = f === :isleftcanonical ? :qh_is_left_canonical : :qh_is_right_canonical
    @eval function Qrochet.$f(h::HIPChain, site; atol::Real = 1e-12)
        res = Ref{Cint}(0)
        check(h.ctx, ccall(($(QuoteNode(cname)), libqh), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cdouble, Ref{Cint}, Ptr{Cdouble}),
                           h.ctx.ptr, h.ptr, id(site), atol, res, C_NULL))
        res[] == 1
    end
end

export HIPChain, simple_update!, evolve_layer!, autotune!, set_graphs!, set_pivoting!, overlap_noalloc

end
