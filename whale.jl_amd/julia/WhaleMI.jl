# WhaleMI: Whale.jl's likelihood path on an MI355X through libwhale_amd.so.
#
# The Julia side of the drop-in: it flattens Whale's own WhaleModel and CCDs
# (src/model.jl:96-145, src/ccd.jl:59-139) into the arrays of the C ABI
# (include/whale_amd.h) and calls it with ccall.  Node ids, wm.order, slice
# counts and WGD ids come from the reference's WhaleModel unchanged, so the
# numbering rules stay in Whale.jl.
#
#   using Whale, WhaleMI
#   wm  = WhaleModel(DLWGD(λ=zeros(n), μ=zeros(n), q=[0.2, 0.1], η=0.9), tree, 0.05)
#   ccd = read_ale(dir, wm)
#   g   = WhaleMI.device(wm, ccd)           # device-resident CCDs
#   WhaleMI.logpdf(g, wm)                    # logpdf(wm, ccd), src/core.jl:58-64
#   v, ∇ = WhaleMI.gradient(g, wm)           # ∂/∂[λ; μ; q; η] (the rates' own vector)
#   Zygote.gradient(θ -> WhaleMI.loglik(g, wm, θ), θ)   # through the rrule below
#   WhaleMI.mixture(g, MixtureModel([wm(θ1), wm(θ2)], [0.7, 0.3]); gradient=true)
#
# This file is not run in this repository's tests (its image has no Julia);
# the C ABI it binds is what the tests exercise through ctypes.
module WhaleMI

using Whale, NewickTree
using Distributions: MixtureModel, probs
using ChainRulesCore

const libwhale = joinpath(@__DIR__, "..", "libwhale_amd.so")
const FP64_PLAIN, FP32_SCALED, FP64_SCALED = Int32(0), Int32(1), Int32(2)
const KEEP_ELL = Int32(0x100)

function check(rc)
    rc == 0 && return nothing
    msg = unsafe_string(ccall((:whale_last_error, libwhale), Cstring, ()))
    error("libwhale_amd: $msg")
end

mutable struct Ctx
    h::Ptr{Cvoid}
end
function Ctx(dev::Integer=0)
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:whale_ctx_create, libwhale), Cint, (Cint, Ptr{Ptr{Cvoid}}), dev, r))
    finalizer(c -> ccall((:whale_ctx_destroy, libwhale), Cint, (Ptr{Cvoid},), c.h), Ctx(r[]))
end

"""Device-resident model structure and families of one WhaleModel layout."""
mutable struct GPUWhale
    ctx::Ctx
    model::Ptr{Cvoid}
    batch::Ptr{Cvoid}
    nnodes::Int
    nwgd::Int
    nfam::Int
    mix::Vector{Ptr{Cvoid}}   # device models of mixture components (same structure)
end

condcode(c) = c isa Whale.RootCondition ? Int32(0) : c isa Whale.NonExtinctCondition ? Int32(1) : Int32(2)

# WhaleModel -> whale_model (the tree is already discretised, src/model.jl:16-22)
function device_model(ctx::Ctx, wm::WhaleModel)
    n = length(wm.order)
    node(i) = wm[i]
    par  = Int32[isroot(node(i)) ? -1 : id(parent(node(i))) - 1 for i=1:n]
    kind = Int32[isroot(node(i)) ? 3 : Whale.iswgd(node(i)) ? 2 : isleaf(node(i)) ? 0 : 1 for i=1:n]
    ch(i, k) = length(children(node(i))) >= k ? Int32(id(children(node(i))[k]) - 1) : Int32(-1)
    c0 = Int32[ch(i, 1) for i=1:n]
    c1 = Int32[ch(i, 2) for i=1:n]
    k  = Int32[length(node(i)) - 1 for i=1:n]            # slices k_e, rows k_e + 1
    Δ  = Float64[node(i)[end, 1] for i=1:n]              # slice length (column 1 of the slices)
    lp = Float64[node(i).data.leafℙ for i=1:n]
    wi = Int32[Whale.iswgd(node(i)) ? Whale.wgdid(node(i)) - 1 : -1 for i=1:n]
    sp = Int32[isleaf(node(i)) ? first(node(i).data.clade) - 1 : -1 for i=1:n]
    ord = Int32[id(x) - 1 for x in wm.order]
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:whale_model_create, libwhale), Cint,
        (Ptr{Cvoid}, Int32, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Int32}, Ptr{Float64},
         Ptr{Float64}, Ptr{Int32}, Ptr{Int32}, Int32, Ptr{Int32}, Int32, Ptr{Ptr{Cvoid}}),
        ctx.h, n, par, kind, c0, c1, k, Δ, lp, wi, sp, n, ord, condcode(wm.condition), r))
    r[]
end

# Vector{CCD} -> whale_batch (clades are already in the reference's order, src/ccd.jl:107-108)
function device_batch(model::Ptr{Cvoid}, ccds::AbstractVector{<:CCD})
    off = Int64[0]; sz = Int32[]; sp = Int32[]; soff = Int64[0]; g1 = Int32[]; g2 = Int32[]; p = Float64[]
    for x in ccds
        for γ in x.clades
            push!(sz, length(γ.leaves))
            push!(sp, length(γ.leaves) == 1 ? first(γ.species) - 1 : -1)
            for t in γ.splits
                push!(g1, t.γ1 - 1); push!(g2, t.γ2 - 1); push!(p, t.p)
            end
            push!(soff, length(g1))
        end
        push!(off, length(sz))
    end
    r = Ref{Ptr{Cvoid}}(C_NULL)
    check(ccall((:whale_batch_create, libwhale), Cint,
        (Ptr{Cvoid}, Int32, Ptr{Int64}, Ptr{Int32}, Ptr{Int32}, Ptr{Int64}, Ptr{Int32}, Ptr{Int32},
         Ptr{Float64}, Ptr{Ptr{Cvoid}}), model, length(ccds), off, sz, sp, soff, g1, g2, p, r))
    r[]
end

"""device(wm, ccds): the model structure and the families on the device."""
function device(wm::WhaleModel, ccds::AbstractVector{<:CCD}; dev::Integer=0)
    ctx = Ctx(dev)
    m = device_model(ctx, wm)
    b = device_batch(m, ccds)
    g = GPUWhale(ctx, m, b, length(wm.order), length(wm.rates.q), length(ccds), Ptr{Cvoid}[])
    finalizer(g) do x
        ccall((:whale_batch_destroy, libwhale), Cint, (Ptr{Cvoid},), x.batch)
        for h in x.mix
            ccall((:whale_model_destroy, libwhale), Cint, (Ptr{Cvoid},), h)
        end
        ccall((:whale_model_destroy, libwhale), Cint, (Ptr{Cvoid},), x.model)
    end
end

# (m::WhaleModel)(θ): getθ per node on the natural scale (src/rmodels.jl:31-64)
set_params!(g::GPUWhale, wm::WhaleModel) = set_params!(g.model, g, wm)
function set_params!(h::Ptr{Cvoid}, g::GPUWhale, wm::WhaleModel)
    n = g.nnodes
    θ = [Whale.getθ(wm.rates, wm[i]) for i=1:n]
    λ = Float64[t.λ for t in θ]
    μ = Float64[t.μ for t in θ]
    p = Float64[haskey(t, :p) ? t.p : 0.0 for t in θ]
    q = Float64.(wm.rates.q)
    check(ccall((:whale_model_set_params, libwhale), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Int32, Float64, Ptr{Float64}),
        h, λ, μ, isempty(q) ? [0.0] : q, length(q), Float64(wm.rates.η), p))
end

"""logpdf(wm, ccds) (src/core.jl:58-64): the conditioned batch log-likelihood,
and the unconditioned per-family values."""
function logpdf(g::GPUWhale, wm::WhaleModel; mode=FP64_PLAIN, keep=false)
    set_params!(g, wm)
    fam = zeros(g.nfam); tot = Ref(0.0)
    check(ccall((:whale_logpdf, libwhale), Cint, (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Ref{Float64}),
                g.batch, Int32(mode) | (keep ? KEEP_ELL : Int32(0)), Int32(0), fam, tot))
    tot[], fam
end

"""logpdf!(wm, ccds) (src/core.jl:46-56): as logpdf, keeping ℓ on the device
for backtrack."""
logpdf!(g::GPUWhale, wm::WhaleModel; mode=FP64_PLAIN) = logpdf(g, wm; mode=mode, keep=true)

# per-node natural-scale derivatives -> the rates' own parameter vector
function chain(wm::WhaleModel, dλ, dμ, dq, dη)
    r = wm.rates
    if r isa ConstantDLWGD
        return vcat(sum(dλ[dλ .!= 0]), sum(dμ[dμ .!= 0]), dq[1:length(r.q)], dη)
    end
    nl = length(r.λ)
    gλ = zeros(nl); gμ = zeros(nl)
    for i in 1:length(wm.order)
        n = wm[i]
        j = Whale.iswgd(n) ? id(Whale.nonwgdchild(n)) : id(n)
        j > nl && continue                      # the root when λ has no entry for it
        θ = Whale.getθ(r, n)
        dλ[i] != 0 && (gλ[j] += θ.λ * dλ[i])    # DLWGD is on log scale: ∂/∂log λ = λ ∂/∂λ
        dμ[i] != 0 && (gμ[j] += θ.μ * dμ[i])
    end
    vcat(gλ, gμ, dq[1:length(r.q)], dη)
end

"""(value, gradient) of logpdf(wm, ccds) with respect to the rates' vector
([λ; μ; q; η], λ and μ on log scale for DLWGD), by the device's reverse sweep."""
function gradient(g::GPUWhale, wm::WhaleModel; mode=FP32_SCALED)
    set_params!(g, wm)
    tot = Ref(0.0); dλ = zeros(g.nnodes); dμ = zeros(g.nnodes); dq = zeros(max(1, g.nwgd)); dη = Ref(0.0)
    check(ccall((:whale_logpdf_grad, libwhale), Cint,
        (Ptr{Cvoid}, Int32, Int32, Ptr{Float64}, Ref{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{Float64}),
        g.batch, Int32(mode), Int32(0), C_NULL, tot, dλ, dμ, dq, dη))
    tot[], chain(wm, dλ, dμ, dq, dη[])
end

unpack(wm::WhaleModel, θ) = wm.rates isa ConstantDLWGD ?
    (λ=θ[1], μ=θ[2], q=θ[3:end-1], η=θ[end]) :
    (n = length(wm.rates.λ); nq = length(wm.rates.q);
     (λ=θ[1:n], μ=θ[n+1:2n], q=θ[2n+1:2n+nq], η=θ[end]))

"""loglik(g, wm, θ): logpdf of wm(θ) for a parameter vector laid out as the
rates' own ([λ; μ; q; η]); differentiable through the rrule below."""
loglik(g::GPUWhale, wm::WhaleModel, θ::AbstractVector; mode=FP32_SCALED) =
    first(logpdf(g, wm(unpack(wm, θ)); mode=mode))

function ChainRulesCore.rrule(::typeof(loglik), g::GPUWhale, wm::WhaleModel, θ::AbstractVector; mode=FP32_SCALED)
    v, ∇ = gradient(g, wm(unpack(wm, θ)); mode=mode)
    pullback(ȳ) = (NoTangent(), NoTangent(), NoTangent(), ȳ .* ∇)
    v, pullback
end

"""mixture(g, mm; gradient): logpdf(mm::MixtureModel, ccds) (src/core.jl:66-76)
for components sharing g's tree layout, every size-class launch working for
all components (whale_mixture_eval).  Returns the value and the
responsibilities (families × components); with `gradient`, also each
component's gradient in its rates' vector and d/dprior."""
function mixture(g::GPUWhale, mm::MixtureModel; mode=FP32_SCALED, gradient::Bool=false)
    comps = mm.components
    K = length(comps)
    while length(g.mix) < K
        push!(g.mix, device_model(g.ctx, comps[1]))
    end
    for j in 1:K
        set_params!(g.mix[j], g, comps[j])
    end
    prior = Float64.(probs(mm))
    n, nw = g.nnodes, g.nwgd
    ng = 2n + nw + 1
    tot = Ref(0.0)
    resp = zeros(K, g.nfam)          # family-major in C: r[i, j] at j + K(i-1)
    grad = zeros(ng, K)
    dprior = zeros(K)
    check(ccall((:whale_mixture_eval, libwhale), Cint,
        (Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Int32, Ptr{Float64}, Int32, Int32, Ref{Float64}, Ptr{Float64},
         Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
        g.batch, g.mix, K, prior, Int32(mode), Int32(gradient), tot, C_NULL, resp,
        gradient ? grad : C_NULL, dprior))
    gradient || return tot[], permutedims(resp)
    ∇ = [chain(comps[j], grad[1:n, j], grad[n+1:2n, j], grad[2n+1:2n+nw, j], grad[end, j]) for j in 1:K]
    tot[], permutedims(resp), ∇, dprior
end

"""backtrack(g, n; seed): n reconciled trees per family from the ℓ the last
logpdf! kept (src/track.jl:188-414), as node tables: per node (clade or 0 for
a loss, species node id, slice t, parent index or 0)."""
function backtrack(g::GPUWhale, n::Integer; seed::Integer=0, cap::Integer=4096)
    counts = zeros(Int32, g.nfam * n)
    nodes = zeros(Int32, 4 * cap * g.nfam * n)
    check(ccall((:whale_backtrack, libwhale), Cint,
        (Ptr{Cvoid}, Int32, Int32, Int32, UInt64, Int32, Ptr{Int32}, Ptr{Int32}),
        g.batch, 0, g.nfam, n, UInt64(seed), cap, counts, nodes))
    any(<(0), counts) && error("a reconciled tree has more than $cap nodes")
    [[begin
          k = (f - 1) * n + s
          base = (k - 1) * cap * 4
          [(nodes[base+4j+1] + 1, nodes[base+4j+2] + 1, nodes[base+4j+3], nodes[base+4j+4] + 1)
           for j=0:counts[k]-1]
      end for s=1:n] for f=1:g.nfam]
end

end  # module
