"""
    JuESHip

Julia binding of `libjues_hip.so` (the C ABI in `include/jues.h`).  It provides
the JuES entry points of the correlation hot path with the same names and
arguments as the reference modules (`Transformation.tei_transform`,
`IntegralTransformation.get_eri`, `MollerPlesset.do_rmp2`,
`CoupledCluster.{RCCD,mRCCD}.do_rccd`, `CoupledCluster.{RCCSD,mRCCSD}.do_rccsd`)
and a ChainRulesCore `rrule` for the RMP2 energy as a function of the MO
coefficients and the AO integrals.

Julia arrays are column-major, which is the memory order the library expects,
so arrays are passed without copies (`GC.@preserve` keeps them alive during the
call).  A non-zero status becomes a Julia `error`, as the reference's `error()`
calls do.

The library location comes from `ENV["JUES_HIP_LIB"]`, else the in-tree build
`jues.jl_amd/lib/libjues_hip.so`.
"""
module JuESHip

using ChainRulesCore

const libjues = get(ENV, "JUES_HIP_LIB", joinpath(@__DIR__, "..", "lib", "libjues_hip.so"))

const JUES_FP64 = Cint(0)
const JUES_FP32_MIXED = Cint(1)
const JUES_FP16X3_MIXED = Cint(2)

struct CCOpts
    maxit::Int32
    e_conv::Float64
    r_conv::Float64
    diis::Int32
    diis_max::Int32
    guess::Int32
    ccsd_eqs::Int32
    fcn::Int32
    fp32_r_conv::Float64
    refine_maxit::Int32
    verbose::Int32
    coupling::Float64
    continuation::Int32
    ladder::Int32
    refine::Int32
end

function CCOpts(; maxit=100, cc_max_iter=maxit, cc_e_conv=1e-10, cc_max_rms=1e-9, diis=true,
                diis_max=5, guess=:mp2, ccsd_eqs=:mRCCSD, fcn=0, fp32_r_conv=1e-5,
                refine_maxit=60, doprint=false, coupling=1.0, continuation=1, ladder=:auto,
                refine=:restart)
    g = Dict(:zero => 0, :mp2 => 1, :ovov => 2, :user => 3)[guess]
    e = Dict(:mRCCSD => 0, :RCCSD => 1)[ccsd_eqs]
    l = Dict(:auto => 0, :stored => 1, :recompute => 2)[ladder]
    r = Dict(:restart => 0, :defect => 1)[refine]
    CCOpts(cc_max_iter, cc_e_conv, cc_max_rms, diis, diis_max, g, e, fcn, fp32_r_conv,
           refine_maxit, doprint, coupling, continuation, l, r)
end

precision_code(p::Symbol) = p === :fp64 ? JUES_FP64 : p === :fp32_mixed ? JUES_FP32_MIXED :
                             p === :fp16x3_mixed ? JUES_FP16X3_MIXED :
                             error("precision must be :fp64, :fp32_mixed or :fp16x3_mixed")

# ---------------------------------------------------------------------------
# context (one per Julia thread, created lazily)
# ---------------------------------------------------------------------------
mutable struct Context
    handle::Ptr{Cvoid}
    ao_key::Any
end

const CONTEXTS = Dict{Int,Context}()

function context(device::Integer=0)
    get!(CONTEXTS, Threads.threadid()) do
        h = Ref{Ptr{Cvoid}}(C_NULL)
        st = ccall((:jues_create, libjues), Cint, (Ptr{Ptr{Cvoid}}, Cint), h, device)
        st == 0 || error("jues_create failed (status $st)")
        ctx = Context(h[], nothing)
        finalizer(c -> ccall((:jues_destroy, libjues), Cvoid, (Ptr{Cvoid},), c.handle), ctx)
        ctx
    end
end

function check(ctx::Context, st::Cint; allow_not_converged=false)
    st == 0 && return
    (allow_not_converged && st == 1) && return
    msg = unsafe_string(ccall((:jues_last_error, libjues), Cstring, (Ptr{Cvoid},), ctx.handle))
    error(msg)
end

"""AO integrals carried by a Wfn (`ao_eri` dense n^4) or a factor array B[mu,nu,P]."""
struct FactoredERI
    B::Array{Float64,3}
end

function ensure_ao!(ctx::Context, gao::Array{Float64,4})
    if ctx.ao_key !== gao
        n = size(gao, 1)
        GC.@preserve gao check(ctx, ccall((:jues_set_ao_eri, libjues), Cint,
                                         (Ptr{Cvoid}, Ptr{Float64}, Cint), ctx.handle, gao, n))
        ctx.ao_key = gao
    end
    size(gao, 1)
end

function ensure_ao!(ctx::Context, f::FactoredERI)
    if ctx.ao_key !== f
        n, _, naux = size(f.B)
        GC.@preserve f check(ctx, ccall((:jues_set_ao_factors, libjues), Cint,
                                       (Ptr{Cvoid}, Ptr{Float64}, Cint, Cint), ctx.handle, f.B, n, naux))
        ctx.ao_key = f
    end
    size(f.B, 1)
end

"""The pair `setup_df` returns (Backend/DF.jl:30-51): pqP[mu,nu,P] and Jpqh = (P|Q)^(-1/2).
Uploading it forms b = pqP*Jpqh on the device (make_bμν, DF.jl:52-58)."""
struct DFIntegrals
    pqP::Array{Float64,3}
    Jpqh::Matrix{Float64}
end

function ensure_ao!(ctx::Context, d::DFIntegrals)
    if ctx.ao_key !== d
        n, _, naux = size(d.pqP)
        size(d.Jpqh) == (naux, naux) || error("Jpqh must be naux x naux")
        GC.@preserve d check(ctx, ccall((:jues_set_df_integrals, libjues), Cint,
                                       (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Cint),
                                       ctx.handle, d.pqP, d.Jpqh, n, naux))
        ctx.ao_key = d
    end
    size(d.pqP, 1)
end

"make_bμν (DF.jl:52-58) on the device."
function make_bμν(pqP::Array{Float64,3}, Jpqh::Matrix{Float64})
    ctx = context()
    d = DFIntegrals(pqP, Jpqh)
    n = ensure_ao!(ctx, d)
    naux = size(pqP, 3)
    B = zeros(n, n, naux)
    GC.@preserve B check(ctx, ccall((:jues_get_ao_factors, libjues), Cint,
                                    (Ptr{Cvoid}, Ptr{Float64}, Cint, Cint), ctx.handle, B, n, naux))
    return B
end

# ---------------------------------------------------------------------------
# Transformation.tei_transform   (Backend/Transformation.jl:15-93)
# ---------------------------------------------------------------------------
module Transformation
import ..JuESHip: context, ensure_ao!, check, libjues, precision_code, FactoredERI
export tei_transform

function tei_transform(gao, C1::Matrix{Float64}, C2::Matrix{Float64}, C3::Matrix{Float64},
                       C4::Matrix{Float64}, name::String="default"; precision=:fp64)
    ctx = context()
    n = ensure_ao!(ctx, gao)
    d = (size(C1, 2), size(C2, 2), size(C3, 2), size(C4, 2))
    out = Array{Float64,4}(undef, d...)
    GC.@preserve C1 C2 C3 C4 out check(ctx, ccall((:jues_tei_transform, libjues), Cint,
        (Ptr{Cvoid}, Cint, Ptr{Float64}, Cint, Ptr{Float64}, Cint, Ptr{Float64}, Cint, Ptr{Float64}, Cint,
         Cint, Ptr{Float64}),
        ctx.handle, precision_code(precision), C1, d[1], C2, d[2], C3, d[3], C4, d[4], n, out))
    return out
end
tei_transform(gao, C::Matrix{Float64}, name::String="default"; kw...) =
    tei_transform(gao, C, C, C, C, name; kw...)
end

# ---------------------------------------------------------------------------
# IntegralTransformation.get_eri   (Backend/IntegralTransformation.jl:38-101)
# ---------------------------------------------------------------------------
module IntegralTransformation
import ..JuESHip: context, ensure_ao!, check, libjues, precision_code
export get_eri, get_fock

function get_eri(wfn, eri_string::String; notation::String="phys", fcn::Int=0, precision=:fp64)
    sizeof(eri_string) == 4 || error("Invalid string given to JuES.IntegralTransformation.get_eri: $eri_string")
    ctx = context()
    n = ensure_ao!(ctx, wfn.ao_eri)
    # IntegralTransformation.jl:52-68: upper case alpha (Cao, Cav), lower case beta (Cbo, Cbv)
    size_of = Dict('O' => wfn.nalpha - fcn, 'o' => wfn.nbeta - fcn, 'V' => wfn.nvira, 'v' => wfn.nvirb)
    all(c -> haskey(size_of, c), eri_string) ||
        error("Invalid string given to JuES.IntegralTransformation.get_eri: $eri_string")
    out = Array{Float64,4}(undef, Tuple(size_of[c] for c in eri_string)...)
    Cao, Cav, Cbo, Cbv = wfn.Cao, wfn.Cav, wfn.Cbo, wfn.Cbv
    GC.@preserve Cao Cav Cbo Cbv out check(ctx, ccall((:jues_get_eri, libjues), Cint,
        (Ptr{Cvoid}, Cint, Cstring, Cstring, Ptr{Float64}, Cint, Ptr{Float64}, Cint, Ptr{Float64}, Cint,
         Ptr{Float64}, Cint, Cint, Cint, Ptr{Float64}),
        ctx.handle, precision_code(precision), eri_string, notation, Cao, wfn.nalpha, Cav, wfn.nvira,
        Cbo, wfn.nbeta, Cbv, wfn.nvirb, n, fcn, out))
    return out
end

"IntegralTransformation.jl:119-141: MO Fock matrix of the spin's occupied density (needs `wfn.hao`)."
function get_fock(wfn; spin="alpha")
    sp = lowercase(spin)
    if sp in ("alpha", "up", "a")
        C, Co = wfn.Ca, wfn.Cao
    elseif sp in ("beta", "down", "b")
        C, Co = wfn.Cb, wfn.Cbo
    else
        error("Invalid Spin option given to JuES.IntegralTransformation.get_fock: $spin")
    end
    ctx = context()
    n = ensure_ao!(ctx, wfn.ao_eri)
    h = wfn.hao
    m = size(C, 2)
    f = zeros(m, m)
    GC.@preserve h C Co f check(ctx, ccall((:jues_get_fock, libjues), Cint,
        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Cint, Ptr{Float64}, Cint, Cint, Ptr{Float64}),
        ctx.handle, h, C, m, Co, size(Co, 2), n, f))
    return f
end
end

# ---------------------------------------------------------------------------
# MollerPlesset.do_rmp2   (MollerPlesset/RMP2.jl:11-45) and its adjoint
# ---------------------------------------------------------------------------
module MollerPlesset
import ..JuESHip: context, ensure_ao!, check, libjues, precision_code, FactoredERI, DFIntegrals
using ChainRulesCore
export do_rmp2, do_rmp2_noalloc, do_df_rmp2, do_ump2, rmp2_energy

function do_rmp2(refWfn; precision=:fp64, kwargs...)
    ctx = context()
    n = ensure_ao!(ctx, refWfn.ao_eri)
    Co, Cv, eps = refWfn.Cao, refWfn.Cav, refWfn.epsa
    E = Ref{Float64}(0.0)
    GC.@preserve Co Cv eps check(ctx, ccall((:jues_rmp2, libjues), Cint,
        (Ptr{Cvoid}, Cint, Ptr{Float64}, Cint, Ptr{Float64}, Cint, Cint, Ptr{Float64}, Ref{Float64}),
        ctx.handle, precision_code(precision), Co, refWfn.nalpha, Cv, refWfn.nvira, n, eps, E))
    return E[]
end

"""
    do_rmp2_noalloc(refWfn)

FP64 `do_rmp2` through `jues_rmp2_noalloc`, which allocates no device memory during the
call (for callers under HIP graph capture or harnesses that forbid allocation): every
intermediate lives in the library's static per-GPU scratch array (JUES_NOALLOC_SCRATCH_BYTES).
Needs factored AO integrals; their upload, on first use, happens before the call.
"""
function do_rmp2_noalloc(refWfn)
    refWfn.ao_eri isa Union{DFIntegrals,FactoredERI} || error("do_rmp2_noalloc needs DFIntegrals or FactoredERI")
    ctx = context()
    n = ensure_ao!(ctx, refWfn.ao_eri)
    Co, Cv, eps = refWfn.Cao, refWfn.Cav, refWfn.epsa
    E = Ref{Float64}(0.0)
    GC.@preserve Co Cv eps check(ctx, ccall((:jues_rmp2_noalloc, libjues), Cint,
        (Ptr{Cvoid}, Cint, Ptr{Float64}, Cint, Ptr{Float64}, Cint, Cint, Ptr{Float64}, Ref{Float64}),
        ctx.handle, precision_code(:fp64), Co, refWfn.nalpha, Cv, refWfn.nvira, n, eps, E))
    return E[]
end

"UMP2.jl:12-79 on the device (alpha-alpha, alpha-beta, beta-beta blocks)."
function do_ump2(refWfn; precision=:fp64)
    ctx = context()
    n = ensure_ao!(ctx, refWfn.ao_eri)
    Cao, Cav, Cbo, Cbv = refWfn.Cao, refWfn.Cav, refWfn.Cbo, refWfn.Cbv
    ea, eb = refWfn.epsa, refWfn.epsb
    E = Ref{Float64}(0.0)
    GC.@preserve Cao Cav Cbo Cbv ea eb check(ctx, ccall((:jues_ump2, libjues), Cint,
        (Ptr{Cvoid}, Cint, Ptr{Float64}, Cint, Ptr{Float64}, Cint, Ptr{Float64}, Cint, Ptr{Float64}, Cint, Cint,
         Ptr{Float64}, Ptr{Float64}, Ref{Float64}),
        ctx.handle, precision_code(precision), Cao, size(Cao, 2), Cav, size(Cav, 2), Cbo, size(Cbo, 2), Cbv,
        size(Cbv, 2), n, ea, eb, E))
    return E[]
end

"DF-RMP2.jl:1-46: the factor route of do_rmp2 on `ao_eri::Union{DFIntegrals,FactoredERI}`."
function do_df_rmp2(refWfn; precision=:fp64)
    refWfn.ao_eri isa Union{DFIntegrals,FactoredERI} || error("do_df_rmp2 needs DFIntegrals or FactoredERI")
    return do_rmp2(refWfn; precision=precision)
end

"""
    rmp2_energy(C, gao, eps, nocc)

RMP2 correlation energy as a differentiable function of the full MO matrix `C`
(n x n, occupied columns first), the dense AO integrals `gao` (n^4) and the
orbital energies `eps` (n, occupied first).
"""
function rmp2_energy(C::Matrix{Float64}, gao::Array{Float64,4}, eps::Vector{Float64}, nocc::Int;
                     precision=:fp64)
    E, _, _, _ = rmp2_grad(C, gao, eps, nocc; want_dC=false, want_dg=false, want_deps=false,
                           precision=precision)
    return E
end

"(E, dE/dC, dE/dg, dE/deps) from the library's hand-written adjoint; unwanted parts are `nothing`."
function rmp2_grad(C, gao, eps, nocc; want_dC=true, want_dg=true, want_deps=true, precision=:fp64)
    ctx = context()
    n = ensure_ao!(ctx, gao)
    E = Ref{Float64}(0.0)
    dC = want_dC ? zeros(n, n) : nothing
    dg = want_dg ? zeros(n, n, n, n) : nothing
    de = want_deps ? zeros(n) : nothing
    pC = dC === nothing ? Ptr{Float64}(C_NULL) : pointer(dC)
    pg = dg === nothing ? Ptr{Float64}(C_NULL) : pointer(dg)
    pe = de === nothing ? Ptr{Float64}(C_NULL) : pointer(de)
    GC.@preserve C eps dC dg de check(ctx, ccall((:jues_rmp2_grad, libjues), Cint,
        (Ptr{Cvoid}, Cint, Ptr{Float64}, Cint, Cint, Ptr{Float64}, Ref{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}),
        ctx.handle, precision_code(precision), C, nocc, n, eps, E, pC, pg, pe))
    return E[], dC, dg, de
end

# reverse rule: the library's hand-written adjoint (recomputes the half-transformed integrals);
# eps enters every denominator, so its tangent is dE/deps, not zero
function ChainRulesCore.rrule(::typeof(rmp2_energy), C::Matrix{Float64}, gao::Array{Float64,4},
                              eps::Vector{Float64}, nocc::Int; precision=:fp64)
    E, dC, dg, de = rmp2_grad(C, gao, eps, nocc; precision=precision)
    function rmp2_energy_pullback(Ebar)
        eb = unthunk(Ebar)
        return NoTangent(), eb .* dC, eb .* dg, eb .* de, NoTangent()
    end
    return E, rmp2_energy_pullback
end
end

# ---------------------------------------------------------------------------
# CoupledCluster.{RCCD,mRCCD,RCCSD,mRCCSD}
# ---------------------------------------------------------------------------
module CoupledCluster
import ..JuESHip: context, ensure_ao!, check, libjues, precision_code, CCOpts, FactoredERI, DFIntegrals

function _cc(method, refWfn; precision=:fp64, return_T=false, check_convergence=true,
             default_guess=:mp2, kwargs...)
    kw = Dict{Symbol,Any}(kwargs)
    haskey(kw, :guess) || (kw[:guess] = default_guess)
    opts = Ref(CCOpts(; kw...))
    ctx = context()
    n = ensure_ao!(ctx, refWfn.ao_eri)
    o, v = refWfn.nalpha, refWfn.nvira
    oa = o - opts[].fcn
    Co, Cv, eps = refWfn.Cao, refWfn.Cav, refWfn.epsa
    T2 = zeros(oa, oa, v, v)
    T1 = zeros(oa, v)
    E = Ref{Float64}(0.0)
    st = GC.@preserve Co Cv eps T1 T2 begin
        if method == 0
            ccall((:jues_rccd, libjues), Cint,
                  (Ptr{Cvoid}, Cint, Ref{CCOpts}, Ptr{Float64}, Cint, Ptr{Float64}, Cint, Cint, Ptr{Float64},
                   Ref{Float64}, Ptr{Float64}, Ptr{Cvoid}, Ptr{Cvoid}),
                  ctx.handle, precision_code(precision), opts, Co, o, Cv, v, n, eps, E, T2, C_NULL, C_NULL)
        else
            ccall((:jues_rccsd, libjues), Cint,
                  (Ptr{Cvoid}, Cint, Ref{CCOpts}, Ptr{Float64}, Cint, Ptr{Float64}, Cint, Cint, Ptr{Float64},
                   Ref{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}, Ptr{Cvoid}),
                  ctx.handle, precision_code(precision), opts, Co, o, Cv, v, n, eps, E, T1, T2, C_NULL, C_NULL)
        end
    end
    check(ctx, st; allow_not_converged=!check_convergence)
    if return_T
        return method == 0 ? (E[], T2) : (E[], T1, T2)
    end
    return E[]
end

module RCCD
import ..CoupledCluster: _cc
"RCCD.jl:33-83 (default guess (ij|ab)/D, RCCD.jl:145-160)"
do_rccd(refWfn; return_T2=false, kwargs...) = _cc(0, refWfn; return_T=return_T2, default_guess=:ovov, kwargs...)
end

module mRCCD
import ..CoupledCluster: _cc
"mRCCD.jl:38-118 (DIIS, zero guess mRCCD.jl:245)"
do_rccd(refWfn; maxit=100, doprint=false, return_T2=false, kwargs...) =
    _cc(0, refWfn; return_T=return_T2, default_guess=:zero, maxit=maxit, doprint=doprint, kwargs...)
end

module RCCSD
import ..CoupledCluster: _cc
"RCCSD.jl:33-101 (Fae T1 term of RCCSD.jl:189)"
do_rccsd(refWfn; return_T=false, kwargs...) = _cc(1, refWfn; return_T=return_T, ccsd_eqs=:RCCSD, kwargs...)
end

module mRCCSD
import ..CoupledCluster: _cc
"mRCCSD.jl:35-101 (Fae T1 term of mRCCSD.jl:140)"
do_rccsd(refWfn; return_T=false, kwargs...) = _cc(1, refWfn; return_T=return_T, ccsd_eqs=:mRCCSD, kwargs...)
end

module PerturbativeTriples
import ...JuESHip: context, ensure_ao!, check, libjues, precision_code
export compute_pT, rccsd_pT
"""PerturbativeTriples.jl:35-138 on the device: same keyword arguments (T1, T2, Vvvvo, Vvooo,
Vvovo, fo, fv); `precision=:fp32_mixed` forms W in FP32 MFMA."""
function compute_pT(; T1::Array{Float64,2}, T2::Array{Float64,4}, Vvvvo::Array{Float64,4},
                    Vvooo::Array{Float64,4}, Vvovo::Array{Float64,4}, fo::Array{Float64,1},
                    fv::Array{Float64,1}, precision=:fp64, doprint=false)
    ctx = context()
    o, v = size(T1)
    Et = Ref{Float64}(0.0)
    GC.@preserve T1 T2 Vvvvo Vvooo Vvovo fo fv check(ctx, ccall((:jues_compute_pT, libjues), Cint,
        (Ptr{Cvoid}, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Ptr{Float64}, Cint, Cint, Cint, Ref{Float64}),
        ctx.handle, precision_code(precision), T1, T2, Vvvvo, Vvooo, Vvovo, fo, fv, o, v, doprint, Et))
    return Et[]
end
"(T) after CCSD on canonical orbitals (AutoRCCSD.jl:294-300), integral blocks built on the device."
function rccsd_pT(refWfn, T1::Array{Float64,2}, T2::Array{Float64,4}; precision=:fp64, doprint=false)
    ctx = context()
    n = ensure_ao!(ctx, refWfn.ao_eri)
    Co, Cv, eps = refWfn.Cao, refWfn.Cav, refWfn.epsa
    Et = Ref{Float64}(0.0)
    GC.@preserve Co Cv eps T1 T2 check(ctx, ccall((:jues_rccsd_pT, libjues), Cint,
        (Ptr{Cvoid}, Cint, Ptr{Float64}, Cint, Ptr{Float64}, Cint, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64},
         Cint, Ref{Float64}),
        ctx.handle, precision_code(precision), Co, refWfn.nalpha, Cv, refWfn.nvira, n, eps, T1, T2, doprint, Et))
    return Et[]
end
end

module AutoRCCSD
import ...JuESHip: context, ensure_ao!, check, libjues, precision_code, CCOpts
import ...JuESHip.IntegralTransformation: get_fock
"""AutoRCCSD.jl:193-300: CCSD driven by the Wfn's MO Fock matrix (off-diagonal blocks in the
equations, its diagonal in the resolvents); plain updates as the reference unless diis=true."""
function do_rccsd(wfn; precision=:fp64, return_T=false, diis=false, kwargs...)
    f = get_fock(wfn; spin="alpha")
    kw = Dict{Symbol,Any}(kwargs)
    haskey(kw, :ccsd_eqs) || (kw[:ccsd_eqs] = :RCCSD)
    haskey(kw, :guess) || (kw[:guess] = :mp2)
    opts = Ref(CCOpts(; diis=diis, kw...))
    ctx = context()
    n = ensure_ao!(ctx, wfn.ao_eri)
    o, v = wfn.nalpha, wfn.nvira
    oa = o - opts[].fcn
    Co, Cv = wfn.Cao, wfn.Cav
    T2 = zeros(oa, oa, v, v)
    T1 = zeros(oa, v)
    E = Ref{Float64}(0.0)
    st = GC.@preserve Co Cv f T1 T2 ccall((:jues_rccsd_fock, libjues), Cint,
        (Ptr{Cvoid}, Cint, Ref{CCOpts}, Ptr{Float64}, Cint, Ptr{Float64}, Cint, Cint, Ptr{Float64},
         Ref{Float64}, Ptr{Float64}, Ptr{Float64}, Ptr{Cvoid}, Ptr{Cvoid}),
        ctx.handle, precision_code(precision), opts, Co, o, Cv, v, n, f, E, T1, T2, C_NULL, C_NULL)
    check(ctx, st)
    return return_T ? (E[], T1, T2) : E[]
end
end

module DFRCCD
import ..CoupledCluster: _cc, FactoredERI, DFIntegrals
"""DF-RCCD.jl:11-53: MP2 guess, exactly `maxit` Jacobi sweeps (no DIIS, no convergence
test), pp-ladder recomputed from bvv*bvv; the residual is RCCD.jl's (its WmBeJ exchange
term, RCCD.jl:400-401, where DF-RCCD.jl:250 reads (me|nf))."""
function do_df_rccd(refWfn; maxit=40, doprint=false, return_T2=false, dfbname="default", precision=:fp64,
                    kwargs...)
    dfbname == "default" || error("dfbname: the auxiliary basis is the one the DFIntegrals were built in")
    precision === :fp64 || error("do_df_rccd: fixed sweeps run in FP64 here (the Python Solver steps the fast forms)")
    refWfn.ao_eri isa Union{DFIntegrals,FactoredERI} || error("do_df_rccd needs DFIntegrals or FactoredERI")
    kw = Dict{Symbol,Any}(kwargs)
    haskey(kw, :ladder) || (kw[:ladder] = :recompute)
    _cc(0, refWfn; return_T=return_T2, default_guess=:mp2, maxit=maxit, doprint=doprint, diis=false,
        cc_e_conv=0.0, cc_max_rms=0.0, check_convergence=false, kw...)
end
end
end

end # module JuESHip
