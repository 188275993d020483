"""
    QCLHipExport

Writes an assembled QCLNEGF.jl problem in the directory format read by
`qclhip/interchange.py` (`manifest.json` plus one `.npy` file per array), so
that reference-assembled problems run on the device backend. Arrays are
written as the reference holds them: column-major bytes behind a `.npy`
header with `fortran_order` true. Needs only the standard library.

    using QCLNEGF, .QCLHipExport
    problem = build_problem(...)                     # solver.jl:13-64
    cache = build_production_cache(problem, ...)     # production_cache.jl:759-840
    QCLHipExport.export_problem("problem_dir", problem, cache; Uᴴ = zeros(N_z))

Fields used (reference types): `problem.grids` (ModelGrids, types.jl:515-526),
`problem.profiles` (MaterialProfiles, :539-547), `problem.basis` (BasisData,
:558-578), `cache.kernels` (production_contracts.jl:262-323),
`project_hamiltonians(problem, Uᴴ)` (basis.jl:288-308), and the scaled
physics constants of `_scaled_physics` / `problem.scales`.
"""
module QCLHipExport

using LinearAlgebra

const FORMAT = "qclhip-problem"
const VERSION = 1

_descr(::Type{ComplexF64}) = "<c16"
_descr(::Type{Float64}) = "<f8"

"""Write `A` as a `.npy` file (format 1.0), column-major (`fortran_order`)."""
function write_npy(path::AbstractString, A::AbstractArray{T}) where {T<:Union{Float64,ComplexF64}}
    dims = size(A)
    shape = length(dims) == 1 ? "($(dims[1]),)" : "(" * join(dims, ", ") * ")"
    header = "{'descr': '$(_descr(T))', 'fortran_order': True, 'shape': $shape, }"
    # magic (6) + version (2) + header length (2) + header, padded to 64 bytes
    total = 10 + length(header) + 1
    header *= " "^((64 - total % 64) % 64) * "\n"
    open(path, "w") do io
        write(io, UInt8(0x93))
        write(io, "NUMPY")
        write(io, UInt8(1), UInt8(0))
        write(io, htol(UInt16(length(header))))
        write(io, header)
        for x in A                      # column-major order, little-endian host
            write(io, x)
        end
    end
    return basename(path)
end

_kind(k) = k isa QCLNEGF.DenseProductionKernel ? "dense" :
           k isa QCLNEGF.MomentumIndependentProductionKernel ? "momentum_independent" :
           k isa QCLNEGF._SeparableIFRProductionKernel ? "separable_ifr" :
           k isa QCLNEGF.LowRankProductionKernel ? "low_rank" :
           throw(ArgumentError("kernel representation $(typeof(k)) has no device counterpart"))

function _kernel_arrays(k)
    k isa QCLNEGF.DenseProductionKernel && return (a = k.matrix,)
    k isa QCLNEGF.MomentumIndependentProductionKernel && return (a = k.block,)
    k isa QCLNEGF._SeparableIFRProductionKernel && return (a = k.left, b = k.right, c = k.radial)
    return (a = k.left, b = k.right_adjoint)
end

_json(x::AbstractString) = "\"" * escape_string(x) * "\""
_json(x::Bool) = x ? "true" : "false"
_json(x::Real) = isfinite(x) ? repr(Float64(x)) : throw(ArgumentError("non-finite scalar"))
_json(d::AbstractDict) = "{" * join([_json(string(k)) * ": " * _json(v) for (k, v) in d], ", ") * "}"
_json(v::AbstractVector) = "[" * join(_json.(v), ", ") * "]"

"""
    export_problem(dir, problem, cache; Uᴴ, mechanisms, scalars)

`mechanisms`: pairs `name => (qK, is_lo)` in the enabled order (defaults:
every kernel of the cache with qK = 1, LO when the name contains "LO");
`scalars`: target, E_p, ħω_LO, N_LO, kBT, α_Σ, g_s and, for the Poisson
step, λ_P, F, z₀, C (the kinetic prefactor of κ²) and α_P.
"""
function export_problem(dir::AbstractString, problem, cache; Uᴴ = zeros(length(problem.grids.x)),
                        mechanisms = [name => (1.0, occursin("LO", String(name))) for name in keys(cache.kernels)],
                        scalars::AbstractDict)
    mkpath(dir)
    g, basis, prof = problem.grids, problem.basis, problem.profiles
    arrays = Dict{String,String}()
    put(name, A) = (arrays[name] = write_npy(joinpath(dir, name * ".npy"), A))
    put("eps", collect(Float64, g.ε))
    put("wE", collect(Float64, g.wᴱ))
    put("wk", collect(Float64, g.wᵏ))
    put("kappa", collect(Float64, g.κ))
    put("x", collect(Float64, g.x))
    put("wx", collect(Float64, g.wˣ))
    put("eps_r", collect(Float64, prof.εᵣ))
    put("ND", collect(Float64, prof.Nᴰ))
    put("h", ComplexF64.(QCLNEGF.project_hamiltonians(problem, Uᴴ)))
    put("Tplus", ComplexF64.(basis.T₊))
    put("Tminus", ComplexF64.(basis.T₋))
    put("Phi", ComplexF64.(basis.Φ))
    put("chi", ComplexF64.(basis.χ))
    put("H0", ComplexF64.(basis.H₀))
    put("Z", ComplexF64.(basis.Z))
    put("Minv", ComplexF64.(basis.M⁻¹))
    mechs = Dict{String,Any}[]
    for (i, (name, (qK, is_lo))) in enumerate(mechanisms)
        k = cache.kernels[name]
        files = Dict{String,String}()
        for (field, A) in pairs(_kernel_arrays(k))
            files[String(field)] = write_npy(joinpath(dir, "mech$(i - 1)_$(field).npy"), ComplexF64.(A))
        end
        push!(mechs, Dict{String,Any}("name" => String(name), "kind" => _kind(k), "qK" => Float64(qK),
                                      "is_lo" => Bool(is_lo), "rank" => k isa QCLNEGF.LowRankProductionKernel ? k.rank : 0,
                                      "arrays" => files))
    end
    manifest = "{\"format\": " * _json(FORMAT) * ", \"version\": $VERSION, \"scalars\": " *
               _json(Dict(string(k) => Float64(v) for (k, v) in scalars)) * ", \"arrays\": " * _json(arrays) *
               ", \"mechanisms\": " * _json(mechs) * "}"
    write(joinpath(dir, "manifest.json"), manifest)
    return dir
end

end # module
