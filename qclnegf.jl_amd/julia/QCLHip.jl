"""
    QCLHip

Julia binding of libqclhip.so, the MI355X operator backend for the SCBA
inner loop of QCLNEGF.jl. Every function is a thin `ccall` into the C-ABI of
`include/qclhip.h`; device fields are `AMDGPU.ROCArray{ComplexF64,4}` in the
reference layout `(N_E, N_k, N_b, N_b)` (column-major, energy fastest), so
`pointer(A)` is passed unchanged. Return codes are rethrown as the
reference's exception types (`SingularException`, `ArgumentError`,
`DimensionMismatch`, `DomainError`).

The module mirrors the production operators it replaces:
`_retarded_green_production` (production_operators.jl:1337-1385),
`_current_green_production` (production_state.jl:1-74),
`_production_static_contraction!` (production_operators.jl:402-658),
`_production_lo_contraction` (:1045-1163), `production_fft_hilbert_transform`
(fft_parallel.jl:247-291), `_retarded_self_energy_fft_production` (:842-884),
`_embedding_component_production` (:886-922), `apply_energy_shift!`
(production_cache.jl:300-325), `_mix_family_production!` (:1475-1491),
`_production_boundary_flux_bar` / `_production_state_populations`
(production_iteration.jl:174-215) and the seed helpers (seed.jl:28-128),
plus what has no reference counterpart: the split residual suite, totals
folded across calls, the batched solve over independent instances, graphs
with deferred results, the autotuner, the allocation-free `current_green`
and the reverse operators, which give a fixed-point adjoint of the device
SCBA loop and a ChainRulesCore `rrule` for its converged current.
"""
module QCLHip

using LinearAlgebra
using Libdl
using AMDGPU
using ChainRulesCore: ChainRulesCore, NoTangent

const libqclhip = joinpath(@__DIR__, "..", "lib", "libqclhip.so")
# entry points chosen at run time are resolved by name
_entry(name::Symbol) = Libdl.dlsym(Libdl.dlopen(libqclhip), name)

const OK = 0
const ERR_ARGUMENT = -1
const ERR_DIMENSION = -2
const ERR_DOMAIN = -3
const ERR_RUNTIME = -4
const ERR_UNSUPPORTED = -5

mutable struct Context
    handle::Ptr{Cvoid}
end

last_error(ctx::Context) = unsafe_string(ccall((:qclhip_ctx_last_error, libqclhip), Cstring, (Ptr{Cvoid},), ctx.handle))

function check(ctx, rc::Integer, what::AbstractString)
    rc == OK && return nothing
    msg = ctx === nothing ? what : string(what, ": ", last_error(ctx))
    rc > 0 && throw(SingularException(Int(rc)))
    rc == ERR_ARGUMENT && throw(ArgumentError(msg))
    rc == ERR_DIMENSION && throw(DimensionMismatch(msg))
    rc == ERR_DOMAIN && throw(DomainError(rc, msg))
    error("qclhip failure $rc: $msg")
end

function Context(device::Integer = 0; stream = AMDGPU.stream())
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:qclhip_ctx_create, libqclhip), Cint, (Cint, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
               device, stream.stream, h)
    check(nothing, rc, "qclhip_ctx_create")
    ctx = Context(h[])
    finalizer(c -> ccall((:qclhip_ctx_destroy, libqclhip), Cint, (Ptr{Cvoid},), c.handle), ctx)
    return ctx
end

synchronize(ctx::Context) = check(ctx, ccall((:qclhip_ctx_synchronize, libqclhip), Cint, (Ptr{Cvoid},), ctx.handle), "synchronize")

"""Device copy of `ModelGrids` (reference/grids.jl:10-47)."""
mutable struct Grid
    ctx::Context
    handle::Ptr{Cvoid}
    NE::Int
    Nk::Int
    Nb::Int
end

function Grid(ctx::Context, ε::Vector{Float64}, wᴱ::Vector{Float64}, wᵏ::Vector{Float64}, Nb::Integer, g_s::Integer)
    length(ε) == length(wᴱ) || throw(DimensionMismatch("energy data differ"))
    h = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:qclhip_grid_create, libqclhip), Cint,
               (Ptr{Cvoid}, Cint, Cint, Cint, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{Ptr{Cvoid}}),
               ctx.handle, length(ε), length(wᵏ), Nb, g_s, ε, wᴱ, wᵏ, h)
    check(ctx, rc, "qclhip_grid_create")
    g = Grid(ctx, h[], length(ε), length(wᵏ), Int(Nb))
    finalizer(x -> ccall((:qclhip_grid_destroy, libqclhip), Cint, (Ptr{Cvoid},), x.handle), g)
    return g
end

_ptr(A) = A === nothing ? C_NULL : Ptr{Cvoid}(pointer(A))

function _shape(g::Grid, A, name)
    A === nothing && return nothing
    size(A) == (g.NE, g.Nk, g.Nb, g.Nb) || throw(DimensionMismatch("$name has the wrong shape"))
    return nothing
end

"""`_retarded_green_production` on the device; returns `(Gᴿ, scale)`."""
function retarded_green(g::Grid, h::ROCArray{ComplexF64,3}, Σᴿ; η::Real = 0.0)
    _shape(g, Σᴿ, "Σᴿ")
    Gᴿ = ROCArray{ComplexF64}(undef, g.NE, g.Nk, g.Nb, g.Nb)
    scale = ROCArray{Float64}(undef, g.NE, g.Nk)
    rc = ccall((:qclhip_retarded_green, libqclhip), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}, Ptr{Cvoid}),
               g.ctx.handle, g.handle, _ptr(h), _ptr(Σᴿ), η, _ptr(Gᴿ), _ptr(scale))
    check(g.ctx, rc, "retarded_green")
    synchronize(g.ctx)
    return Gᴿ, scale
end

"""
    dyson_keldysh_strided!(g, h, Σᴿ, Σˡ, Σᵍ, Gᴿ, Gˡ, Gᵍ)

The Dyson + Keldysh launch pair (raw G< and G> products) on fields stored as
`plane × Nb²` matrices with `plane ≥ NE*Nk`: element `(e, m, a, b)` of a field
is entry `(e + NE*(m-1), a + Nb*(b-1))` (one-based), the layout of a
`reshape(view(F, 1:NE*Nk, :), NE, Nk, Nb, Nb)` into a larger array.
"""
function dyson_keldysh_strided!(g::Grid, h::ROCArray{ComplexF64,3}, Σᴿ::ROCArray{ComplexF64,2},
                                Σˡ::ROCArray{ComplexF64,2}, Σᵍ::ROCArray{ComplexF64,2},
                                Gᴿ::ROCArray{ComplexF64,2}, Gˡ::ROCArray{ComplexF64,2},
                                Gᵍ::ROCArray{ComplexF64,2})
    plane = size(Σᴿ, 1)
    for F in (Σᴿ, Σˡ, Σᵍ, Gᴿ, Gˡ, Gᵍ)
        size(F) == (plane, g.Nb * g.Nb) || throw(DimensionMismatch("fields must be plane × Nb² with one plane"))
    end
    rc = ccall((:qclhip_dyson_keldysh_strided, libqclhip), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble,
                Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
               g.ctx.handle, g.handle, plane, _ptr(h), _ptr(Σᴿ), _ptr(Σˡ), _ptr(Σᵍ), 0.0,
               _ptr(Gᴿ), _ptr(Gˡ), _ptr(Gᵍ), C_NULL, C_NULL)
    check(g.ctx, rc, "dyson_keldysh_strided")
    return Gᴿ, Gˡ, Gᵍ
end

"""
    current_green!(g, h, families, target, Gᴿ, Gˡ, Gᵍ, A, scale)

`_current_green_production`: `families` is a vector of `(Σᴿ, Σˡ, Σᵍ)` device
triples, embedding first, then the mechanisms in `problem.kernels.enabled`
order. Returns `(λ, number, holes)`; `current_green_coefficients!` takes the
same arguments and returns all five results `(λ, number, holes, a, c)`.
"""
current_green!(args...; kw...) = Tuple(current_green_coefficients!(args...; kw...)[1:3])

function current_green_coefficients!(g::Grid, h, families, target::Real, Gᴿ, Gˡ, Gᵍ, A, scale;
                        occupation_normalization::Symbol = :paired_convex, allocation_free::Bool = false)
    occupation_normalization in (:paired_convex, :scalar_lesser) ||
        throw(ArgumentError("occupation_normalization must be :paired_convex or :scalar_lesser"))
    n = length(families)
    fr = Ptr{Cvoid}[_ptr(f[1]) for f in families]
    fl = Ptr{Cvoid}[_ptr(f[2]) for f in families]
    fg = Ptr{Cvoid}[_ptr(f[3]) for f in families]
    out = zeros(Float64, 5)
    # :scalar_lesser is production_state.jl:54-61; the allocation-free entry
    # takes the block traces from the library's fixed per-device scratch
    entry = occupation_normalization === :scalar_lesser ? :qclhip_current_green_scalar_lesser :
            allocation_free ? :qclhip_current_green_noalloc : :qclhip_current_green
    (occupation_normalization === :scalar_lesser && allocation_free) &&
        throw(ArgumentError("the allocation-free call is the :paired_convex current_green"))
    rc = ccall(_entry(entry), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}},
                Cdouble, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}),
               g.ctx.handle, g.handle, _ptr(h), n, fr, fl, fg, target, _ptr(Gᴿ), _ptr(Gˡ), _ptr(Gᵍ), _ptr(A),
               _ptr(scale), out)
    check(g.ctx, rc, "current_green")
    return out
end

"""`current_green_coefficients!` that also writes the totals (`totals`: the
(Σᴿ, Σˡ, Σᵍ) fields, or `nothing`) and the raw-product screen word of every
(e,m) block to `raw_screen` (a device `UInt8` array of NE*Nk entries), for
`residual_suite(...; raw_screen)` / `residual_state(...; raw_screen)` on the
same state: the measured SCBA iteration's fused path (FP64 vector engine,
N_b <= 32). No counterpart in the reference."""
function current_green_screened!(g::Grid, h, families, target::Real, Gᴿ, Gˡ, Gᵍ, A, scale, raw_screen;
                                 totals = nothing)
    n = length(families)
    fr = Ptr{Cvoid}[_ptr(f[1]) for f in families]
    fl = Ptr{Cvoid}[_ptr(f[2]) for f in families]
    fg = Ptr{Cvoid}[_ptr(f[3]) for f in families]
    tp = totals === nothing ? (C_NULL, C_NULL, C_NULL) : map(_ptr, totals)
    out = zeros(Float64, 5)
    check(g.ctx, ccall((:qclhip_current_green_screened, libqclhip), Cint,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}},
                        Cdouble, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64},
                        Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                       g.ctx.handle, g.handle, _ptr(h), n, fr, fl, fg, target, _ptr(Gᴿ), _ptr(Gˡ), _ptr(Gᵍ), _ptr(A),
                       _ptr(scale), out, tp[1], tp[2], tp[3], _ptr(raw_screen)), "current_green_screened")
    return out
end

"""Device `EnergyShiftPlan` (production_cache.jl:258-288)."""
mutable struct ShiftPlan
    ctx::Context
    handle::Ptr{Cvoid}
end

function ShiftPlan(g::Grid, δ::Real)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(g.ctx, ccall((:qclhip_shift_create, libqclhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ref{Ptr{Cvoid}}),
                       g.ctx.handle, g.handle, δ, h), "shift_create")
    p = ShiftPlan(g.ctx, h[])
    finalizer(x -> ccall((:qclhip_shift_destroy, libqclhip), Cint, (Ptr{Cvoid},), x.handle), p)
    return p
end

function apply_energy_shift!(dst, plan::ShiftPlan, X)
    dst === X && throw(ArgumentError("in-place aliasing is not supported"))
    ncols = length(X) ÷ size(X, 1)
    check(plan.ctx, ccall((:qclhip_apply_shift, libqclhip), Cint,
                          (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{Cvoid}),
                          plan.ctx.handle, plan.handle, ncols, _ptr(X), _ptr(dst)), "apply_shift")
    return dst
end

"""
Device representation of an `AbstractProductionKernel`
(production_contracts.jl:262-323): `kind` 0 dense, 1 momentum independent,
2 separable IFR, 3 low rank. Built from the host matrices of the reference
representation.
"""
mutable struct HipProductionKernel
    grid::Grid
    handle::Ptr{Cvoid}
end

function HipProductionKernel(g::Grid, kind::Integer, a::Matrix{ComplexF64}, b = nothing, c = nothing; rank::Integer = 0)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    pb = b === nothing ? C_NULL : pointer(b)
    pc = c === nothing ? C_NULL : pointer(c)
    GC.@preserve a b c begin
        rc = ccall((:qclhip_kernel_create, libqclhip), Cint,
                   (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Ptr{ComplexF64}, Ptr{ComplexF64}, Ptr{ComplexF64},
                    Ref{Ptr{Cvoid}}), g.ctx.handle, g.handle, kind, rank, a, pb, pc, h)
    end
    check(g.ctx, rc, "kernel_create")
    k = HipProductionKernel(g, h[])
    finalizer(x -> ccall((:qclhip_kernel_destroy, libqclhip), Cint, (Ptr{Cvoid},), x.handle), k)
    return k
end

"""`_production_static_contraction!(Σ, op, G, wᵏ, qᴷ, chunk; α, β)` on the device."""
function production_static_contraction!(Σ, op::HipProductionKernel, G, qᴷ::Real; α::Real = 1, β::Real = 0)
    Σ === G && throw(ArgumentError("in-place aliasing is not supported"))
    check(op.grid.ctx, ccall((:qclhip_contract, libqclhip), Cint,
                             (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Cdouble, Ptr{Cvoid}),
                             op.grid.ctx.handle, op.handle, _ptr(G), qᴷ, α, β, _ptr(Σ)), "contract")
    return Σ
end

function lo_contraction!(Σˡ, Σᵍ, op::HipProductionKernel, Gˡ, Gᵍ, qᴷ::Real, Nᴸᴼ::Real, plus::ShiftPlan, minus::ShiftPlan)
    check(op.grid.ctx, ccall((:qclhip_lo_contraction, libqclhip), Cint,
                             (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Ptr{Cvoid}, Ptr{Cvoid},
                              Ptr{Cvoid}, Ptr{Cvoid}),
                             op.grid.ctx.handle, op.handle, _ptr(Gˡ), _ptr(Gᵍ), qᴷ, Nᴸᴼ, plus.handle, minus.handle,
                             _ptr(Σˡ), _ptr(Σᵍ)), "lo_contraction")
    return Σˡ, Σᵍ
end

"""Device `ProductionFFTHilbertPlan` (fft_parallel.jl:27-143)."""
mutable struct HilbertPlan
    grid::Grid
    handle::Ptr{Cvoid}
end

function HilbertPlan(g::Grid; verify_roundoff::Bool = true)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(g.ctx, ccall((:qclhip_hilbert_create, libqclhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int64, Cint, Ref{Ptr{Cvoid}}),
                       g.ctx.handle, g.handle, g.Nk * g.Nb^2, verify_roundoff, h), "hilbert_create")
    p = HilbertPlan(g, h[])
    finalizer(x -> ccall((:qclhip_hilbert_destroy, libqclhip), Cint, (Ptr{Cvoid},), x.handle), p)
    return p
end

"""Implementation of the plan's transform: `:auto` (by L), `:lds` (one
workgroup per column, L ≤ 16384) or `:hbm` (four-step through HBM, L ≥ 4096)."""
function set_path!(plan::HilbertPlan, path::Symbol)
    code = Dict(:auto => 0, :lds => 1, :hbm => 2)
    haskey(code, path) || throw(ArgumentError("unknown Hilbert path $path"))
    check(plan.grid.ctx, ccall((:qclhip_hilbert_set_path, libqclhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint),
                               plan.grid.ctx.handle, plan.handle, code[path]), "hilbert_set_path")
    return plan
end

"""`_retarded_self_energy_fft_production`: returns `(Σᴿ, roundoff)`."""
function retarded_self_energy!(Σᴿ, plan::HilbertPlan, Σˡ, Σᵍ)
    r = Ref{Float64}(0.0)
    check(plan.grid.ctx, ccall((:qclhip_retarded_self_energy, libqclhip), Cint,
                               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{Float64}),
                               plan.grid.ctx.handle, plan.handle, _ptr(Σˡ), _ptr(Σᵍ), _ptr(Σᴿ), C_NULL, r),
          "retarded_self_energy")
    return Σᴿ, r[]
end

"""`_embedding_component_production(T, G, plan)`; `T` on the device."""
function embedding_component!(Σ, g::Grid, T::ROCArray{ComplexF64,2}, G, plan::ShiftPlan)
    check(g.ctx, ccall((:qclhip_embedding_component, libqclhip), Cint,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                       g.ctx.handle, g.handle, _ptr(T), _ptr(G), plan.handle, _ptr(Σ)), "embedding")
    return Σ
end

function mix!(old, cand, α::Real, ctx::Context)
    check(ctx, ccall((:qclhip_mix, libqclhip), Cint, (Ptr{Cvoid}, Int64, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble),
                     ctx.handle, length(old), _ptr(old), _ptr(cand), α), "mix")
    return old
end

function boundary_flux_bar(g::Grid, plusˡ, plusᵍ, Gˡ, Gᵍ)
    J = Ref{Float64}(0.0)
    check(g.ctx, ccall((:qclhip_boundary_flux, libqclhip), Cint,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{Float64}),
                       g.ctx.handle, g.handle, _ptr(plusˡ), _ptr(plusᵍ), _ptr(Gˡ), _ptr(Gᵍ), J), "boundary_flux")
    return J[]
end

function state_populations(g::Grid, Gˡ, target::Real)
    out = zeros(Float64, g.Nb)
    check(g.ctx, ccall((:qclhip_state_populations, libqclhip), Cint,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ptr{Float64}),
                       g.ctx.handle, g.handle, _ptr(Gˡ), target, out), "populations")
    return out
end

"""`production_residual_suite` (residuals.jl:537-772): (r_D, r_A, r_K, r_PSD, r_caus).
`A` may be `nothing` (formed from Gᴿ); with `raw_screen` (from
`current_green_screened!` on the same state and totals) the raw-product
screens are taken from it, with the same results."""
function residual_suite(g::Grid, h, Gᴿ, Gˡ, Gᵍ, A, families, candidate_lesser; raw_screen = nothing)
    n = length(families)
    fr = Ptr{Cvoid}[_ptr(f[1]) for f in families]
    fl = Ptr{Cvoid}[_ptr(f[2]) for f in families]
    fg = Ptr{Cvoid}[_ptr(f[3]) for f in families]
    cl = Ptr{Cvoid}[_ptr(c) for c in candidate_lesser]
    out = zeros(Float64, 5)
    if raw_screen === nothing
        check(g.ctx, ccall((:qclhip_residual_suite, libqclhip), Cint,
                           (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint,
                            Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Float64}),
                           g.ctx.handle, g.handle, _ptr(h), _ptr(Gᴿ), _ptr(Gˡ), _ptr(Gᵍ), _ptr(A), n, fr, fl, fg,
                           length(cl), cl, out), "residual_suite")
    else
        check(g.ctx, ccall((:qclhip_residual_suite_prescreened, libqclhip), Cint,
                           (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint,
                            Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Cvoid},
                            Ptr{Float64}),
                           g.ctx.handle, g.handle, _ptr(h), _ptr(Gᴿ), _ptr(Gˡ), _ptr(Gᵍ), _ptr(A), n, fr, fl, fg,
                           length(cl), cl, _ptr(raw_screen), out), "residual_suite_prescreened")
    end
    return (r_D = out[1], r_A = out[2], r_K = out[3], r_PSD = out[4], r_caus = out[5])
end

"""Linear mixing fused with `production_selfenergy_residual` of the unmixed
pairs; `partner[p]` (0-based, -1 for none) couples plus and minus. α = 0
measures only. With `totals = [(T, members), ...]` (members: 0-based pair
indices in summation order, every pair in one total) each `T` receives the
`_total_family_production` sum of its members' mixed values."""
function mix_fields_residual!(ctx::Context, olds, cands, α::Real, partner::Vector{Cint}; totals = nothing)
    po = Ptr{Cvoid}[_ptr(o) for o in olds]
    pc = Ptr{Cvoid}[_ptr(c) for c in cands]
    nterms = length(olds) + count(p -> partner[p] > p - 1, eachindex(partner))
    out = zeros(Float64, nterms)
    if totals === nothing
        check(ctx, ccall((:qclhip_mix_fields_residual, libqclhip), Cint,
                         (Ptr{Cvoid}, Int64, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Cdouble, Ptr{Cint}, Ptr{Float64}),
                         ctx.handle, length(first(olds)), length(olds), po, pc, α, partner, out), "mix_fields_residual")
    else
        counts = Cint[length(m) for (_, m) in totals]
        members = Cint[p for (_, m) in totals for p in m]
        pt = Ptr{Cvoid}[_ptr(T) for (T, _) in totals]
        check(ctx, ccall((:qclhip_mix_fields_residual_totals, libqclhip), Cint,
                         (Ptr{Cvoid}, Int64, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Cdouble, Ptr{Cint}, Cint, Ptr{Cint},
                          Ptr{Cint}, Ptr{Ptr{Cvoid}}, Ptr{Float64}),
                         ctx.handle, length(first(olds)), length(olds), po, pc, α, partner, length(totals), counts,
                         members, pt, out), "mix_fields_residual_totals")
    end
    return out
end

"""The Poisson step (production_solver.jl:1044-1047): density, bordered
Poisson solve, U mixing and projected Hamiltonians; all arrays on the device."""
function electron_density!(n̄, g::Grid, Gˡ, χ, gs::Real)
    check(g.ctx, ccall((:qclhip_electron_density, libqclhip), Cint,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}),
                       g.ctx.handle, g.handle, _ptr(Gˡ), size(χ, 1), _ptr(χ), gs, _ptr(n̄)), "electron_density")
    return n̄
end

function periodic_poisson!(U, ctx::Context, εᵣ, wˣ, Nᴰ, λ_P::Real, n̄)
    out = zeros(Float64, 3)
    check(ctx, ccall((:qclhip_periodic_poisson, libqclhip), Cint,
                     (Ptr{Cvoid}, Cint, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}),
                     ctx.handle, length(n̄), _ptr(εᵣ), _ptr(wˣ), _ptr(Nᴰ), λ_P, _ptr(n̄), _ptr(U), out), "poisson")
    return U, out[1], out[2], out[3]   # U, ζ, r_P, neutrality
end

"""`_anderson_guard` (production_mixing.jl:82-111): (first_failure, code)."""
function anderson_guard(g::Grid, families)
    n = length(families)
    fr = Ptr{Cvoid}[_ptr(f[1]) for f in families]
    fl = Ptr{Cvoid}[_ptr(f[2]) for f in families]
    fg = Ptr{Cvoid}[_ptr(f[3]) for f in families]
    first = Ref{Int64}(0)
    code = Ref{Cint}(0)
    check(g.ctx, ccall((:qclhip_anderson_guard, libqclhip), Cint,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}},
                        Ref{Int64}, Ref{Cint}),
                       g.ctx.handle, g.handle, n, fr, fl, fg, first, code), "anderson_guard")
    return first[], code[]
end

# ---------------------------------------------------------------------------
# Remaining operators of the C-ABI

"""`_number_functional_production` (production_operators.jl:1417-1445)."""
function number_functional(g::Grid, G)
    v = Ref{Float64}(0.0)
    check(g.ctx, ccall((:qclhip_number_functional, libqclhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ref{Float64}),
                       g.ctx.handle, g.handle, _ptr(G), v), "number_functional")
    return v[]
end

"""`_green_state_production` without a given G> (production_operators.jl:1387-1415):
G> = Gᴿ - Gᴿ' + G<, A = i(Gᴿ - Gᴿ'); either output may be `nothing`."""
function green_state!(Gᵍ, A, g::Grid, Gᴿ, Gˡ)
    check(g.ctx, ccall((:qclhip_green_state, libqclhip), Cint,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                       g.ctx.handle, g.handle, _ptr(Gᴿ), _ptr(Gˡ), _ptr(Gᵍ), _ptr(A)), "green_state")
    return Gᵍ, A
end

_fermi(ε, μ, kBT) = (x = (ε - μ) / kBT; x > 40 ? 0.0 : (x < -40 ? 1.0 : 1 / (exp(x) + 1)))

"""Bisection of Σₑ qₑ f(εₑ; μ) = target on [ε₁ - 50kBT, ε_N + 50kBT]
(seed.jl:175-182), down to adjacent floating-point numbers."""
function seed_chemical_potential(ε::Vector{Float64}, q::Vector{Float64}, kBT::Real, target::Real)
    number(μ) = sum(q[e] * _fermi(ε[e], μ, kBT) for e in eachindex(ε)) - target
    lo, hi = ε[1] - 50kBT, ε[end] + 50kBT
    flo, fhi = number(lo), number(hi)
    flo * fhi <= 0 || throw(ArgumentError("seed chemical-potential root is not bracketed"))
    flo == 0 && return lo
    fhi == 0 && return hi
    for _ = 1:4000
        mid = lo + (hi - lo) / 2
        (mid <= lo || mid >= hi) && break
        fm = number(mid)
        fm == 0 && return mid
        if (fm < 0) == (flo < 0)
            lo, flo = mid, fm
        else
            hi, fhi = mid, fm
        end
    end
    return abs(flo) <= abs(fhi) ? lo : hi
end

"""`_seed_green_production` (seed.jl:151-197): Dyson with η and Σ = 0, the
spectral number weights on the device, the chemical potential by bisection
on [ε₁ - 50kBT, ε_N + 50kBT] on the host, G< = -f (Gᴿ - Gᴿ') on the device.
Returns (Gᴿ, G<, G>, μ)."""
function seed_green(g::Grid, h, ε::Vector{Float64}, η::Real, kBT::Real, target::Real)
    Gᴿ, _ = retarded_green(g, h, nothing; η = η)
    q = zeros(Float64, g.NE)
    check(g.ctx, ccall((:qclhip_seed_weights, libqclhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}),
                       g.ctx.handle, g.handle, _ptr(Gᴿ), q), "seed_weights")
    μ = seed_chemical_potential(ε, q, kBT, target)
    Gˡ = similar(Gᴿ)
    check(g.ctx, ccall((:qclhip_seed_lesser, libqclhip), Cint,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cdouble, Ptr{Cvoid}),
                       g.ctx.handle, g.handle, _ptr(Gᴿ), μ, kBT, _ptr(Gˡ)), "seed_lesser")
    Gᵍ = similar(Gᴿ)
    green_state!(Gᵍ, nothing, g, Gᴿ, Gˡ)
    synchronize(g.ctx)
    return Gᴿ, Gˡ, Gᵍ, μ
end

"""GF-solve precision: `:fp64`, `:fp32_refine` or `:fp32_products` (see qclhip.h)."""
function set_precision!(g::Grid, precision::Symbol; refine_steps::Integer = 2, refine_tol::Real = 0.0)
    code = Dict(:fp64 => 0, :fp32_refine => 1, :fp32_products => 2)
    haskey(code, precision) || throw(ArgumentError("unknown precision $precision"))
    check(g.ctx, ccall((:qclhip_grid_set_precision, libqclhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint, Cdouble),
                       g.ctx.handle, g.handle, code[precision], refine_steps, refine_tol), "set_precision")
    return g
end

"""Keldysh engine: `:vector`, `:matrix` or `:fused` (Dyson and the vector
products in one kernel)."""
function set_engine!(g::Grid, engine::Symbol)
    code = Dict(:vector => 0, :matrix => 1, :fused => 2)
    haskey(code, engine) || throw(ArgumentError("unknown engine $engine"))
    check(g.ctx, ccall((:qclhip_grid_set_engine, libqclhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint), g.ctx.handle,
                       g.handle, code[engine]), "set_engine")
    return g
end

set_keldysh_split!(g::Grid, waves_per_block::Integer) = (check(g.ctx, ccall((:qclhip_grid_set_keldysh_split,
    libqclhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint), g.ctx.handle, g.handle, waves_per_block), "set_keldysh_split"); g)

set_tiles!(g::Grid, dyson_waves::Integer, keldysh_waves::Integer) = (check(g.ctx, ccall((:qclhip_grid_set_tiles,
    libqclhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Cint), g.ctx.handle, g.handle, dyson_waves, keldysh_waves),
    "set_tiles"); g)

"""The autotuner (engine and tile shapes, cached per shape in the context);
returns its record."""
function autotune!(g::Grid)
    ms = zeros(Float64, 2)
    check(g.ctx, ccall((:qclhip_grid_autotune, libqclhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}),
                       g.ctx.handle, g.handle, ms), "autotune")
    t = zeros(Float64, 8)
    ch = zeros(Cint, 4)
    check(g.ctx, ccall((:qclhip_grid_tuning, libqclhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}, Ptr{Cint}),
                       g.ctx.handle, g.handle, t, ch), "grid_tuning")
    return (engine = (:vector, :matrix, :fused)[ch[1] + 1], dyson_waves = Int(ch[2]), keldysh_waves = Int(ch[3]),
            keldysh_split = Int(ch[4]), ms = t)
end

# -- deferred results and graphs ---------------------------------------------
defer_results!(ctx::Context, buf::Union{Nothing,ROCArray{Float64,1}}) =
    check(ctx, ccall((:qclhip_ctx_defer_results, libqclhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Cint), ctx.handle,
                     _ptr(buf), buf === nothing ? 0 : length(buf)), "defer_results")

mutable struct Graph
    ctx::Context
    handle::Ptr{Cvoid}
end

graph_begin(ctx::Context) = check(ctx, ccall((:qclhip_graph_begin, libqclhip), Cint, (Ptr{Cvoid},), ctx.handle),
                                  "graph_begin")

function graph_end(ctx::Context)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ctx, ccall((:qclhip_graph_end, libqclhip), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}), ctx.handle, h), "graph_end")
    gr = Graph(ctx, h[])
    finalizer(x -> ccall((:qclhip_graph_destroy, libqclhip), Cint, (Ptr{Cvoid},), x.handle), gr)
    return gr
end

launch(gr::Graph) = check(gr.ctx, ccall((:qclhip_graph_launch, libqclhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}),
                                        gr.ctx.handle, gr.handle), "graph_launch")

# -- batched GF solve over independent instances ------------------------------
mutable struct Batch
    ctx::Context
    handle::Ptr{Cvoid}
    grids::Vector{Grid}
end

function Batch(grids::Vector{Grid})
    ctx = first(grids).ctx
    h = Ref{Ptr{Cvoid}}(C_NULL)
    pg = Ptr{Cvoid}[g.handle for g in grids]
    check(ctx, ccall((:qclhip_batch_create, libqclhip), Cint, (Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Ref{Ptr{Cvoid}}),
                     ctx.handle, length(grids), pg, h), "batch_create")
    b = Batch(ctx, h[], grids)
    finalizer(x -> ccall((:qclhip_batch_destroy, libqclhip), Cint, (Ptr{Cvoid},), x.handle), b)
    return b
end

"""current_green for every instance: `hs[i]`, `families[i]` (same count for
all), `targets[i]`, outputs `Gᴿs[i]`, `Gˡs[i]`, `Gᵍs[i]`; returns a 5 × n
matrix of (λ, number, holes, a, c)."""
function batch_current_green!(b::Batch, hs, families, targets::Vector{Float64}, Gᴿs, Gˡs, Gᵍs)
    n, nf = length(b.grids), length(first(families))
    fam(k) = Ptr{Cvoid}[_ptr(families[i][f][k]) for i = 1:n for f = 1:nf]
    out = zeros(Float64, 5, n)
    check(b.ctx, ccall((:qclhip_batch_current_green, libqclhip), Cint,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Ptr{Cvoid}}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}},
                        Ptr{Ptr{Cvoid}}, Ptr{Float64}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Float64}),
                       b.ctx.handle, b.handle, Ptr{Cvoid}[_ptr(h) for h in hs], nf, fam(1), fam(2), fam(3), targets,
                       Ptr{Cvoid}[_ptr(x) for x in Gᴿs], Ptr{Cvoid}[_ptr(x) for x in Gˡs],
                       Ptr{Cvoid}[_ptr(x) for x in Gᵍs], out), "batch_current_green")
    return out
end

"""Launch granularity of the batched solve: `:persistent` or `:per_instance`."""
set_mode!(b::Batch, mode::Symbol) = check(b.ctx, ccall((:qclhip_batch_set_mode, libqclhip), Cint,
    (Ptr{Cvoid}, Ptr{Cvoid}, Cint), b.ctx.handle, b.handle, mode === :per_instance ? 1 : 0), "batch_set_mode")

# -- split residual suite -----------------------------------------------------
"""Every residual but r_K, before the candidates exist; `A` may be `nothing`;
`raw_screen` as for `residual_suite`."""
function residual_state(g::Grid, h, Gᴿ, Gˡ, Gᵍ, A, families; raw_screen = nothing)
    n = length(families)
    out = zeros(Float64, 5)
    fr = Ptr{Cvoid}[_ptr(f[1]) for f in families]
    fl = Ptr{Cvoid}[_ptr(f[2]) for f in families]
    fg = Ptr{Cvoid}[_ptr(f[3]) for f in families]
    if raw_screen === nothing
        check(g.ctx, ccall((:qclhip_residual_state, libqclhip), Cint,
                           (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint,
                            Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Float64}),
                           g.ctx.handle, g.handle, _ptr(h), _ptr(Gᴿ), _ptr(Gˡ), _ptr(Gᵍ), _ptr(A), n, fr, fl, fg, out),
              "residual_state")
    else
        check(g.ctx, ccall((:qclhip_residual_state_prescreened, libqclhip), Cint,
                           (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint,
                            Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}, Ptr{Float64}),
                           g.ctx.handle, g.handle, _ptr(h), _ptr(Gᴿ), _ptr(Gˡ), _ptr(Gᵍ), _ptr(A), n, fr, fl, fg,
                           _ptr(raw_screen), out), "residual_state_prescreened")
    end
    return (r_D = out[1], r_A = out[2], r_PSD = out[4], r_caus = out[5])
end

function residual_keldysh(g::Grid, Gᴿ, Gˡ, candidate_lesser)
    out = zeros(Float64, 5)
    check(g.ctx, ccall((:qclhip_residual_keldysh, libqclhip), Cint,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Float64}),
                       g.ctx.handle, g.handle, _ptr(Gᴿ), _ptr(Gˡ), length(candidate_lesser),
                       Ptr{Cvoid}[_ptr(c) for c in candidate_lesser], out), "residual_keldysh")
    return out[3]
end

"""`mix_fields_residual!` with totals folded across calls: `totals =
[(T, members, flags), ...]`, flags a combination of 1 (accumulate: continue
the fold from T) and 2 (sum the members' candidate values)."""
function mix_fields_residual_accumulate!(ctx::Context, olds, cands, α::Real, partner::Vector{Cint}, totals)
    nterms = length(olds) + count(p -> partner[p] > p - 1, eachindex(partner))
    out = zeros(Float64, nterms)
    counts = Cint[length(m) for (_, m, _) in totals]
    members = Cint[p for (_, m, _) in totals for p in m]
    flags = Cint[f for (_, _, f) in totals]
    check(ctx, ccall((:qclhip_mix_fields_residual_accumulate, libqclhip), Cint,
                     (Ptr{Cvoid}, Int64, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Cdouble, Ptr{Cint}, Cint, Ptr{Cint},
                      Ptr{Cint}, Ptr{Cint}, Ptr{Ptr{Cvoid}}, Ptr{Float64}),
                     ctx.handle, length(first(olds)), length(olds), Ptr{Cvoid}[_ptr(o) for o in olds],
                     Ptr{Cvoid}[_ptr(c) for c in cands], α, partner, length(totals), counts, members, flags,
                     Ptr{Cvoid}[_ptr(T) for (T, _, _) in totals], out), "mix_fields_residual_accumulate")
    return out
end

sum_fields!(out, ctx::Context, xs) = check(ctx, ccall((:qclhip_sum_fields, libqclhip), Cint,
    (Ptr{Cvoid}, Int64, Cint, Ptr{Ptr{Cvoid}}, Ptr{Cvoid}), ctx.handle, length(out), length(xs),
    Ptr{Cvoid}[_ptr(x) for x in xs], _ptr(out)), "sum_fields")

fill_zero!(X, ctx::Context) = check(ctx, ccall((:qclhip_fill_zero, libqclhip), Cint, (Ptr{Cvoid}, Int64, Ptr{Cvoid}),
                                               ctx.handle, length(X), _ptr(X)), "fill_zero")

# ---------------------------------------------------------------------------
# Adjoint (no reference counterpart; algorithm_registry.jl:226-249 lists it
# as missing). Gradient convention, as ChainRulesCore's for a real function
# of complex arguments: Z̄ = ∂J/∂Re Z + i ∂J/∂Im Z.

"""Reverse of `current_green!`: the gradients of the total self-energy from
those of (Gᴿ, G<, G>) and, when `h̄ !== nothing`, of h."""
function current_green_adjoint!(Σ̄ᴿ, Σ̄ˡ, Σ̄ᵍ, h̄, g::Grid, families, target::Real, coef::Vector{Float64}, Gᴿ, Gˡ, Gᵍ,
                                Ḡᴿ, Ḡˡ, Ḡᵍ)
    n = length(families)
    check(g.ctx, ccall((:qclhip_current_green_adjoint, libqclhip), Cint,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Cdouble, Ptr{Float64},
                        Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                        Ptr{Cvoid}, Ptr{Cvoid}),
                       g.ctx.handle, g.handle, n, Ptr{Cvoid}[_ptr(f[2]) for f in families],
                       Ptr{Cvoid}[_ptr(f[3]) for f in families], target, coef, _ptr(Gᴿ), _ptr(Gˡ), _ptr(Gᵍ),
                       _ptr(Ḡᴿ), _ptr(Ḡˡ), _ptr(Ḡᵍ), _ptr(Σ̄ᴿ), _ptr(Σ̄ˡ), _ptr(Σ̄ᵍ), _ptr(h̄)), "current_green_adjoint")
    return h̄
end

boundary_flux_adjoint!(g::Grid, J̄::Real, plusˡ, plusᵍ, Gˡ, Gᵍ, p̄ˡ, p̄ᵍ, Ḡˡ, Ḡᵍ) =
    check(g.ctx, ccall((:qclhip_boundary_flux_adjoint, libqclhip), Cint,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid},
                        Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                       g.ctx.handle, g.handle, J̄, _ptr(plusˡ), _ptr(plusᵍ), _ptr(Gˡ), _ptr(Gᵍ), _ptr(p̄ˡ), _ptr(p̄ᵍ),
                       _ptr(Ḡˡ), _ptr(Ḡᵍ)), "boundary_flux_adjoint")

shift_transpose!(X, g::Grid, plan::ShiftPlan, Y, α::Real; accumulate::Bool = false) =
    check(g.ctx, ccall((:qclhip_shift_transpose, libqclhip), Cint,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cdouble, Cint, Ptr{Cvoid}),
                       g.ctx.handle, g.handle, plan.handle, _ptr(Y), α, accumulate, _ptr(X)), "shift_transpose")

"""(K diag(wᵏ))† in the same representation."""
function kernel_adjoint(op::HipProductionKernel)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(op.grid.ctx, ccall((:qclhip_kernel_adjoint, libqclhip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ref{Ptr{Cvoid}}),
                             op.grid.ctx.handle, op.handle, h), "kernel_adjoint")
    k = HipProductionKernel(op.grid, h[])
    finalizer(x -> ccall((:qclhip_kernel_destroy, libqclhip), Cint, (Ptr{Cvoid},), x.handle), k)
    return k
end

retarded_self_energy_adjoint!(Σ̄ˡ, Σ̄ᵍ, plan::HilbertPlan, Σ̄ᴿ) =
    check(plan.grid.ctx, ccall((:qclhip_retarded_self_energy_adjoint, libqclhip), Cint,
                               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                               plan.grid.ctx.handle, plan.handle, _ptr(Σ̄ᴿ), _ptr(Σ̄ˡ), _ptr(Σ̄ᵍ)),
          "retarded_self_energy_adjoint")

antihermitian_part!(X, g::Grid) = check(g.ctx, ccall((:qclhip_antihermitian_part, libqclhip), Cint,
    (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}), g.ctx.handle, g.handle, _ptr(X)), "antihermitian_part")

embedding_adjoint!(X̄, g::Grid, T, plan::ShiftPlan, Σ̄) =
    check(g.ctx, ccall((:qclhip_embedding_adjoint, libqclhip), Cint,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                       g.ctx.handle, g.handle, _ptr(T), plan.handle, _ptr(Σ̄), _ptr(X̄)), "embedding_adjoint")

"""Mix every (old, new) pair and return (Σ‖new - old‖², Σ‖new‖²); α = 0 measures only."""
function mix_fields_change!(ctx::Context, olds, news, α::Real)
    out = zeros(Float64, 2)
    check(ctx, ccall((:qclhip_mix_fields_change, libqclhip), Cint,
                     (Ptr{Cvoid}, Int64, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Cdouble, Ptr{Float64}),
                     ctx.handle, length(first(olds)), length(olds), Ptr{Cvoid}[_ptr(o) for o in olds],
                     Ptr{Cvoid}[_ptr(c) for c in news], α, out), "mix_fields_change")
    return out[1], out[2]
end

# ---------------------------------------------------------------------------
# The production SCBA loop on the device and its fixed-point adjoint
# (production_solver.jl:435-844 with linear mixing), for the rrule below.

struct HipMechanism
    kernel::HipProductionKernel
    adjoint::HipProductionKernel
    qᴷ::Float64
    is_lo::Bool
end

HipMechanism(k::HipProductionKernel, qᴷ::Real, is_lo::Bool) = HipMechanism(k, kernel_adjoint(k), qᴷ, is_lo)

"""Everything the loop needs on the device: grid, T± (device Nb × Nb),
mechanisms in `problem.kernels.enabled` order, shift plans for ±Eᵖ and
±ħωᴸᴼ, the Hilbert plan, target N_D²ᴰ, thermal Nᴸᴼ and the mixing α."""
struct HipScbaProblem
    grid::Grid
    T₊::ROCArray{ComplexF64,2}
    T₋::ROCArray{ComplexF64,2}
    mechanisms::Vector{HipMechanism}
    plusE::ShiftPlan
    minusE::ShiftPlan
    plusLO::ShiftPlan
    minusLO::ShiftPlan
    hilbert::HilbertPlan
    target::Float64
    Nᴸᴼ::Float64
    α::Float64
end

_field(g::Grid) = ROCArray{ComplexF64}(undef, g.NE, g.Nk, g.Nb, g.Nb)
_zfield(g::Grid) = (X = _field(g); fill_zero!(X, g.ctx); X)

"""The SCBA state: one (Σᴿ, Σ<, Σ>) triple per family (plus, minus, then
the mechanisms) and the Green functions of the last solve."""
mutable struct HipScbaState
    families::Vector{NTuple{3,ROCArray{ComplexF64,4}}}
    Gᴿ::ROCArray{ComplexF64,4}
    Gˡ::ROCArray{ComplexF64,4}
    Gᵍ::ROCArray{ComplexF64,4}
    coef::Vector{Float64}
end

"""Fresh candidates of every family from (Gᴿ, G<, G>)
(_scattering_candidate_production, _embedding_self_energy_production)."""
function candidates!(out, p::HipScbaProblem, Gᴿ, Gˡ, Gᵍ)
    g = p.grid
    for (X, c) in ((Gᴿ, 1), (Gˡ, 2), (Gᵍ, 3))
        embedding_component!(out[1][c], g, p.T₊, X, p.plusE)
        embedding_component!(out[2][c], g, p.T₋, X, p.minusE)
    end
    for (i, m) in enumerate(p.mechanisms)
        R, L, G = out[2+i]
        if m.is_lo
            lo_contraction!(L, G, m.kernel, Gˡ, Gᵍ, m.qᴷ, p.Nᴸᴼ, p.plusLO, p.minusLO)
        else
            production_static_contraction!(L, m.kernel, Gˡ, m.qᴷ)
            production_static_contraction!(G, m.kernel, Gᵍ, m.qᴷ)
        end
        retarded_self_energy!(R, p.hilbert, L, G)
    end
    return out
end

"""One production iteration with linear mixing; returns (λ, J̄)."""
function scba_iteration!(s::HipScbaState, p::HipScbaProblem, h, scratch)
    g = p.grid
    λ, number, holes = current_green!(g, h, s.families, p.target, s.Gᴿ, s.Gˡ, s.Gᵍ, nothing, nothing)
    J̄ = boundary_flux_bar(g, s.families[1][2], s.families[1][3], s.Gˡ, s.Gᵍ)
    candidates!(scratch, p, s.Gᴿ, s.Gˡ, s.Gᵍ)
    for (f, c) in zip(s.families, scratch), k = 1:3
        mix!(f[k], c[k], p.α, g.ctx)
    end
    return λ, J̄
end

"""Iterate to |ΔJ̄| ≤ tol |J̄|; returns (J̄, iterations)."""
function converge!(s::HipScbaState, p::HipScbaProblem, h; tol = 1e-12, max_iterations = 5000)
    g = p.grid
    scratch = [(_field(g), _field(g), _field(g)) for _ in s.families]
    prev, J̄ = NaN, NaN
    it = 0
    for outer it = 1:max_iterations
        _, J̄ = scba_iteration!(s, p, h, scratch)
        it >= 5 && abs(J̄ - prev) <= tol * abs(J̄) && break
        prev = J̄
    end
    return J̄, it
end

"""dJ̄/dh at a converged state by the damped fixed-point adjoint
μ ← (1-α)μ + α[(∂F/∂s)† μ + ∂J̄/∂s] on the anti-Hermitian tangent space of
Σ< and Σ> (the mirror of qclhip/adjoint.py); h̄ has the layout of h."""
function adjoint_gradient(s::HipScbaState, p::HipScbaProblem, h; tol = 1e-13, max_iterations = 5000)
    g, ctx = p.grid, p.grid.ctx
    # (λ, number, holes, a, c) of the solve at the converged state
    coef = current_green_coefficients!(g, h, s.families, p.target, s.Gᴿ, s.Gˡ, s.Gᵍ, nothing, nothing)
    plus = s.families[1]
    Jˡ, Jᵍ, p̄ˡ, p̄ᵍ = (_zfield(g) for _ = 1:4)
    boundary_flux_adjoint!(g, 1.0, plus[2], plus[3], s.Gˡ, s.Gᵍ, p̄ˡ, p̄ᵍ, Jˡ, Jᵍ)
    antihermitian_part!(p̄ˡ, g)
    antihermitian_part!(p̄ᵍ, g)
    μ = [(_zfield(g), _zfield(g), _zfield(g)) for _ in s.families]
    Ḡᴿ, Ḡˡ, Ḡᵍ = _field(g), _field(g), _field(g)
    σ = (_field(g), _field(g), _field(g))
    pnew = (_field(g), _field(g))
    tmpL, tmpG, Z = _field(g), _field(g), _field(g)
    function reverse_candidates!()
        fill_zero!(Ḡᴿ, ctx)
        copyto!(Ḡˡ, Jˡ)
        copyto!(Ḡᵍ, Jᵍ)
        for (i, m) in enumerate(p.mechanisms)
            R, L, G = μ[2+i]
            copyto!(tmpL, L)
            copyto!(tmpG, G)
            retarded_self_energy_adjoint!(tmpL, tmpG, p.hilbert, R)
            if m.is_lo
                shift_transpose!(Z, g, p.plusLO, tmpL, p.Nᴸᴼ + 1)
                shift_transpose!(Z, g, p.minusLO, tmpL, p.Nᴸᴼ; accumulate = true)
                production_static_contraction!(Ḡˡ, m.adjoint, Z, m.qᴷ; β = 1)
                shift_transpose!(Z, g, p.minusLO, tmpG, p.Nᴸᴼ + 1)
                shift_transpose!(Z, g, p.plusLO, tmpG, p.Nᴸᴼ; accumulate = true)
                production_static_contraction!(Ḡᵍ, m.adjoint, Z, m.qᴷ; β = 1)
            else
                production_static_contraction!(Ḡˡ, m.adjoint, tmpL, m.qᴷ; β = 1)
                production_static_contraction!(Ḡᵍ, m.adjoint, tmpG, m.qᴷ; β = 1)
            end
        end
        for (T, plan, f) in ((p.T₊, p.plusE, μ[1]), (p.T₋, p.minusE, μ[2]))
            embedding_adjoint!(Ḡᴿ, g, T, plan, f[1])
            embedding_adjoint!(Ḡˡ, g, T, plan, f[2])
            embedding_adjoint!(Ḡᵍ, g, T, plan, f[3])
        end
    end
    for it = 1:max_iterations
        reverse_candidates!()
        current_green_adjoint!(σ[1], σ[2], σ[3], nothing, g, s.families, p.target, coef, s.Gᴿ, s.Gˡ, s.Gᵍ, Ḡᴿ, Ḡˡ, Ḡᵍ)
        antihermitian_part!(σ[2], g)
        antihermitian_part!(σ[3], g)
        sum_fields!(pnew[1], ctx, (σ[2], p̄ˡ))
        sum_fields!(pnew[2], ctx, (σ[3], p̄ᵍ))
        olds, news = ROCArray{ComplexF64,4}[], ROCArray{ComplexF64,4}[]
        for (k, f) in enumerate(μ)
            nl, ng = k == 1 ? pnew : (σ[2], σ[3])
            append!(olds, f)
            append!(news, (σ[1], nl, ng))
        end
        num, den = mix_fields_change!(ctx, olds, news, p.α)
        sqrt(num / max(den, 1e-300)) <= tol && break
    end
    reverse_candidates!()
    h̄ = ROCArray{ComplexF64}(undef, g.Nk, g.Nb, g.Nb)
    current_green_adjoint!(σ[1], σ[2], σ[3], h̄, g, s.families, p.target, coef, s.Gᴿ, s.Gˡ, s.Gᵍ, Ḡᴿ, Ḡˡ, Ḡᵍ)
    return h̄
end

"""The converged boundary flux J̄(h) of the device SCBA loop started from
`state` (which it updates in place)."""
converged_current(p::HipScbaProblem, s::HipScbaState, h) = first(converge!(s, p, h))

"""ChainRulesCore rrule of `converged_current`: the forward converges the
loop; the pullback runs the fixed-point adjoint at that state and returns
J̄̄ · dJ̄/dh (the implicit adjoint needs only the converged state, so no
iterate is checkpointed)."""
function ChainRulesCore.rrule(::typeof(converged_current), p::HipScbaProblem, s::HipScbaState, h)
    J̄ = converged_current(p, s, h)
    function converged_current_pullback(Δ)
        h̄ = adjoint_gradient(s, p, h)
        return NoTangent(), NoTangent(), NoTangent(), h̄ .* ChainRulesCore.unthunk(Δ)
    end
    return J̄, converged_current_pullback
end

"""dJ̄/dU on a real-space grid from h̄ (basis.jl:300: ∂h_m/∂U_g = Φ[g,:]* Φ[g,:]ᵀ):
dJ̄/dU_g = Σ_m Re Σ_ab conj(h̄_m[a,b]) conj(Φ[g,a]) Φ[g,b]."""
function potential_gradient(h̄, Φ::Matrix{ComplexF64})
    H = dropdims(sum(Array(h̄); dims = 1); dims = 1)  # Σ_m h̄_m, (Nb, Nb)
    return [real(sum(conj(H[a, b]) * conj(Φ[g, a]) * Φ[g, b] for a in axes(H, 1), b in axes(H, 2)))
            for g in axes(Φ, 1)]
end

"""dJ̄/dL_i of layer thicknesses by the chain rule through central
differences of a host assembly L ↦ h(L) (Nk, Nb, Nb)."""
function thickness_gradient(h̄, assemble, L::Vector{Float64}; step = 1e-3)
    hb = Array(h̄)
    return [begin
        d = zeros(length(L)); d[i] = step
        real(sum(conj.(hb) .* ((assemble(L .+ d) .- assemble(L .- d)) ./ (2step))))
    end for i in eachindex(L)]
end

end # module
