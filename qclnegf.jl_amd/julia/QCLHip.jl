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
(production_iteration.jl:174-215) and the seed helpers (seed.jl:28-128).
"""
module QCLHip

using LinearAlgebra
using AMDGPU

const libqclhip = joinpath(@__DIR__, "..", "lib", "libqclhip.so")

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
order. Returns `(λ, number, holes)`.
"""
function current_green!(g::Grid, h, families, target::Real, Gᴿ, Gˡ, Gᵍ, A, scale)
    n = length(families)
    fr = Ptr{Cvoid}[_ptr(f[1]) for f in families]
    fl = Ptr{Cvoid}[_ptr(f[2]) for f in families]
    fg = Ptr{Cvoid}[_ptr(f[3]) for f in families]
    out = zeros(Float64, 5)
    rc = ccall((:qclhip_current_green, libqclhip), Cint,
               (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}},
                Cdouble, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Float64}),
               g.ctx.handle, g.handle, _ptr(h), n, fr, fl, fg, target, _ptr(Gᴿ), _ptr(Gˡ), _ptr(Gᵍ), _ptr(A),
               _ptr(scale), out)
    check(g.ctx, rc, "current_green")
    return out[1], out[2], out[3]
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

"""`production_residual_suite` (residuals.jl:537-772): (r_D, r_A, r_K, r_PSD, r_caus)."""
function residual_suite(g::Grid, h, Gᴿ, Gˡ, Gᵍ, A, families, candidate_lesser)
    n = length(families)
    fr = Ptr{Cvoid}[_ptr(f[1]) for f in families]
    fl = Ptr{Cvoid}[_ptr(f[2]) for f in families]
    fg = Ptr{Cvoid}[_ptr(f[3]) for f in families]
    cl = Ptr{Cvoid}[_ptr(c) for c in candidate_lesser]
    out = zeros(Float64, 5)
    check(g.ctx, ccall((:qclhip_residual_suite, libqclhip), Cint,
                       (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Cint,
                        Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Ptr{Ptr{Cvoid}}, Cint, Ptr{Ptr{Cvoid}}, Ptr{Float64}),
                       g.ctx.handle, g.handle, _ptr(h), _ptr(Gᴿ), _ptr(Gˡ), _ptr(Gᵍ), _ptr(A), n, fr, fl, fg,
                       length(cl), cl, out), "residual_suite")
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

end # module
