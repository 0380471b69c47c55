"""
    FC3AMD

Julia binding of libfc3.so, the MI355X (gfx950) library behind `include/fc3.h`, for
ForceConstants.jl users.  The functions below keep the reference's signatures verbatim
(`src/ifc.jl:3-35`, `src/analytical/third_order_analytical.jl:185`, `src/mcc.jl:29,60,89`,
`src/io.jl:9,39`, `src/asr.jl:43,62`) and its assertion messages; the device array type
`DeviceArray` takes the place of `CuArray` in the `mcc3` methods (the library needs no
CUDA.jl or AMDGPU.jl: a `DeviceArray` is a device pointer, its dimensions and a
finalizer).

    using ForceConstants, FC3AMD
    Ψ = FC3AMD.third_order(sys, LJ(3.4u"Å", 0.24037u"kcal/mol", 8.5u"Å"),
                           AnalyticalCalculator(1e-12, 8.5u"Å"); check_asr = true)
    FC3AMD.mass_weight_third_order!(Ψ, masses(sys))
    FC3AMD.set_precision!(FC3AMD.BF16X3)                  # default FP32, as the reference
    K3 = FC3AMD.mcc3(DeviceArray(Float32.(Ψ.values)), DeviceArray(Float32.(phi)))

Supported: `LJ` with `AnalyticalCalculator` and `StillingerWeberSilicon` with
`AutoDiffCalculator` in 3 dimensions; other combinations raise an `ArgumentError` (the
library has no kernels for them).  Every call runs on the task-local handle (one HIP
stream and workspace per task), as the reference's `mcc3` runs on the task-local CUDA
stream.  `n_threads` is accepted and unused: the work runs on the GPU.
"""
module FC3AMD

using LinearAlgebra: Hermitian, eigen
using Unitful: ustrip, unit
using ForceConstants: SuperCellSystem, Potential, ForceConstantCalculator, LJ,
                      StillingerWeberSilicon, AnalyticalCalculator, AutoDiffCalculator,
                      DenseForceConstants, SecondOrderForceConstants, ThirdOrderForceConstants,
                      energy_unit, length_unit, n_atoms, positions
import ForceConstants

export DeviceArray, second_order, third_order, mass_weight_third_order!, mcc3, mcc3!,
       mcc3_unique, asr_satisfied, save_second_order, save_third_order, set_precision!,
       get_modes_device, weighted_square_loss, third_order_mcc3, FP32, BF16X3

const libfc3 = get(ENV, "FC3_LIBRARY", joinpath(@__DIR__, "..", "fc3", "libfc3.so"))

# ---- status: the reference raises AssertionError for its @asserts ---------------------
const FC3_OK = Cint(0)
const FC3_ERR_ARGUMENT, FC3_ERR_CUTOFF, FC3_ERR_BLOCK_SIZE = Cint(1), Cint(2), Cint(3)

function check(status::Cint)
    status == FC3_OK && return nothing
    msg = unsafe_string(ccall((:fc3_last_error, libfc3), Cstring, ()))
    status in (FC3_ERR_ARGUMENT, FC3_ERR_CUTOFF, FC3_ERR_BLOCK_SIZE) && throw(AssertionError(msg))
    error("fc3: " * msg)
end

# ---- handle: one HIP stream + workspace, one per task -----------------------------------
mutable struct Handle
    ptr::Ptr{Cvoid}
    function Handle(device::Integer = 0)
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:fc3_create, libfc3), Cint, (Ref{Ptr{Cvoid}}, Cint), r, device))
        h = new(r[])
        finalizer(x -> ccall((:fc3_destroy, libfc3), Cint, (Ptr{Cvoid},), x.ptr), h)
    end
end

handle() = get!(() -> Handle(0), task_local_storage(), :fc3_handle)::Handle
synchronize(h::Handle = handle()) = check(ccall((:fc3_synchronize, libfc3), Cint, (Ptr{Cvoid},), h.ptr))

# ---- device arrays -----------------------------------------------------------------------
mutable struct DeviceArray{T,N}
    ptr::Ptr{T}
    dims::NTuple{N,Int}
    h::Handle
    function DeviceArray{T}(h::Handle, dims::Vararg{Int,N}) where {T,N}
        r = Ref{Ptr{Cvoid}}(C_NULL)
        check(ccall((:fc3_device_alloc, libfc3), Cint, (Ptr{Cvoid}, Ref{Ptr{Cvoid}}, Csize_t),
                    h.ptr, r, sizeof(T) * prod(dims; init = 1)))
        a = new{T,N}(Ptr{T}(r[]), dims, h)
        finalizer(x -> ccall((:fc3_device_free, libfc3), Cint, (Ptr{Cvoid}, Ptr{Cvoid}),
                             x.h.ptr, x.ptr), a)
    end
end
DeviceArray{T}(dims::Vararg{Int,N}) where {T,N} = DeviceArray{T}(handle(), dims...)

function DeviceArray(h::Handle, x::Array{T,N}) where {T,N}
    a = DeviceArray{T}(h, size(x)...)
    check(ccall((:fc3_memcpy_h2d, libfc3), Cint, (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Csize_t),
                h.ptr, a.ptr, x, sizeof(x)))
    return a
end
DeviceArray(x::Array) = DeviceArray(handle(), x)

Base.size(a::DeviceArray) = a.dims
Base.size(a::DeviceArray, i::Integer) = a.dims[i]

function Base.Array(a::DeviceArray{T,N}) where {T,N}
    out = Array{T,N}(undef, a.dims)
    check(ccall((:fc3_memcpy_d2h, libfc3), Cint, (Ptr{Cvoid}, Ptr{T}, Ptr{T}, Csize_t),
                a.h.ptr, out, a.ptr, sizeof(out)))
    return out
end

# ---- systems and potentials as the C ABI takes them -----------------------------------------
# positions: 3 x N (column-major = atom-major x, y, z) in the potential's length unit
positions_matrix(sys, lu) = Float64.(reduce(hcat, [ustrip.(lu, p) for p in positions(sys)]))
box3(sys, lu) = Float64.(ustrip.(lu, sys.box_sizes_SC))
mass_vector(m) = Float64.(ustrip.(m))

# SW_Params in field order (SW.jl:3-14), then the cos(theta0) that the reference's
# generated derivative blocks bake in (SW/H3_exec_ijk.jl:39)
function sw_abi_params(pot::StillingerWeberSilicon)
    p = pot.params
    eu, lu = energy_unit(pot), length_unit(pot)
    return Float64[ustrip(eu, p.ϵ), ustrip(lu, p.σ), p.a, p.λ, p.γ, p.cosθ₀, p.A, p.B, p.p,
                   p.q, -0.333333333333]
end

unsupported(pot, calc) = throw(ArgumentError(
    "the MI355X library computes LJ with AnalyticalCalculator and StillingerWeberSilicon " *
    "with AutoDiffCalculator, not $(nameof(typeof(pot))) with $(nameof(typeof(calc)))"))

function second_order_device!(Φ::DeviceArray{Float64,2}, sys, pot::LJ, calc::AnalyticalCalculator)
    lu, eu = length_unit(pot), energy_unit(pot)
    check(ccall((:fc3_second_order_lj, libfc3), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cint, Ptr{Float64}, Float64, Float64, Float64,
                 Float64, Float64, Ptr{Float64}),
                Φ.h.ptr, positions_matrix(sys, lu), n_atoms(sys), box3(sys, lu),
                ustrip(lu, pot.σ), ustrip(eu, pot.ϵ), ustrip(lu, pot.r_cut),
                ustrip(lu, calc.r_cut), calc.tol, Φ.ptr))
    return Φ
end
function second_order_device!(Φ::DeviceArray{Float64,2}, sys, pot::StillingerWeberSilicon,
                              calc::AutoDiffCalculator)
    lu = length_unit(pot)
    check(ccall((:fc3_second_order_sw, libfc3), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cint, Ptr{Float64}, Ptr{Float64}, Float64, Float64,
                 Ptr{Float64}),
                Φ.h.ptr, positions_matrix(sys, lu), n_atoms(sys), box3(sys, lu),
                sw_abi_params(pot), ustrip(lu, calc.r_cut), calc.tol, Φ.ptr))
    return Φ
end
second_order_device!(Φ, sys, pot, calc) = unsupported(pot, calc)

# FP64 Psi (psi64) and/or mass-weighted FP32 Psi (psi32, masses !== nothing)
function third_order_device!(psi64, psi32, masses, sys, pot::LJ, calc::AnalyticalCalculator)
    lu, eu = length_unit(pot), energy_unit(pot)
    h = psi64 === nothing ? psi32.h : psi64.h
    m = masses === nothing ? C_NULL : mass_vector(masses)
    check(ccall((:fc3_third_order_lj, libfc3), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cint, Ptr{Float64}, Float64, Float64, Float64,
                 Float64, Float64, Ptr{Float64}, Ptr{Float64}, Ptr{Float32}),
                h.ptr, positions_matrix(sys, lu), n_atoms(sys), box3(sys, lu),
                ustrip(lu, pot.σ), ustrip(eu, pot.ϵ), ustrip(lu, pot.r_cut),
                ustrip(lu, calc.r_cut), calc.tol, m,
                psi64 === nothing ? C_NULL : psi64.ptr, psi32 === nothing ? C_NULL : psi32.ptr))
end
function third_order_device!(psi64, psi32, masses, sys, pot::StillingerWeberSilicon,
                             calc::AutoDiffCalculator)
    lu = length_unit(pot)
    h = psi64 === nothing ? psi32.h : psi64.h
    m = masses === nothing ? C_NULL : mass_vector(masses)
    check(ccall((:fc3_third_order_sw, libfc3), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cint, Ptr{Float64}, Ptr{Float64}, Float64, Float64,
                 Ptr{Float64}, Ptr{Float64}, Ptr{Float32}),
                h.ptr, positions_matrix(sys, lu), n_atoms(sys), box3(sys, lu),
                sw_abi_params(pot), ustrip(lu, calc.r_cut), calc.tol, m,
                psi64 === nothing ? C_NULL : psi64.ptr, psi32 === nothing ? C_NULL : psi32.ptr))
end
third_order_device!(psi64, psi32, masses, sys, pot, calc) = unsupported(pot, calc)

# ---- checks (src/asr.jl:43-74; ifc.jl:15) -------------------------------------------------
function asr_violations(ifc::DeviceArray{Float64,O}, N_atoms, D, tol) where {O}
    cnt, first, s = Ref{Clonglong}(0), Ref{Clonglong}(0), Ref{Float64}(0.0)
    check(ccall((:fc3_asr_satisfied, libfc3), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cint, Cint, Cint, Float64, Ref{Clonglong},
                 Ref{Clonglong}, Ref{Float64}),
                ifc.h.ptr, ifc.ptr, O, N_atoms, D, tol, cnt, first, s))
    return cnt[], first[], s[]
end

function report_first(first, s, n, D, order)
    b, rs = first % D, first ÷ D
    row, sl = rs % n, rs ÷ n
    order == 3 && println("=======\n On k = $(sl ÷ D + 1), γ = $(sl % D + 1) \n=======")
    println("Set i = $(row ÷ D + 1), α = $(row % D + 1), β = $(b + 1) does not satisfy ASR. Sum was: $s")
end

function asr_satisfied(ifc2::DeviceArray{Float64,2}, N_atoms, D, tol; verbose = false)
    cnt, first, s = asr_violations(ifc2, N_atoms, D, tol)
    verbose && cnt > 0 && report_first(first, s, N_atoms * D, D, 2)
    return cnt == 0
end
function asr_satisfied(ifc3::DeviceArray{Float64,3}, N_atoms, D, tol; verbose = false)
    cnt, first, s = asr_violations(ifc3, N_atoms, D, tol)
    verbose && cnt > 0 && report_first(first, s, N_atoms * D, D, 3)
    return cnt == 0
end
asr_satisfied(ifc2::AbstractArray{T,2}, N_atoms, D, tol; verbose = false) where {T} =
    asr_satisfied(DeviceArray(Array{Float64}(ifc2)), N_atoms, D, tol; verbose)
asr_satisfied(ifc3::Array{T,3}, N_atoms, D, tol; verbose = false) where {T} =
    asr_satisfied(DeviceArray(Array{Float64}(ifc3)), N_atoms, D, tol; verbose)

function device_issymmetric(Φ::DeviceArray{Float64,2})
    c = Ref{Clonglong}(0)
    check(ccall((:fc3_issymmetric2, libfc3), Cint, (Ptr{Cvoid}, Ptr{Float64}, Cint, Ref{Clonglong}),
                Φ.h.ptr, Φ.ptr, size(Φ, 1), c))
    return c[] == 0
end

# ---- second_order / third_order (src/ifc.jl:3-35) -----------------------------------------
function second_order(sys::SuperCellSystem{D}, pot::Potential,
    calc::ForceConstantCalculator; n_threads::Int = Threads.nthreads(),
    check_asr = false, check_symmetry = false) where D

    D == 3 || throw(ArgumentError("the MI355X library computes 3-dimensional systems"))
    N_atoms = n_atoms(sys)
    Φ = DeviceArray{Float64}(D * N_atoms, D * N_atoms)
    second_order_device!(Φ, sys, pot, calc)          # includes the tolerance pass

    check_asr && @assert asr_satisfied(Φ, N_atoms, D, calc.tol) "ASR not satisfied for second order force constants, please report this"
    check_symmetry && @assert device_issymmetric(Φ) "IFC were not symmetric, please report this"

    return DenseForceConstants(Array(Φ), energy_unit(pot) / length_unit(pot)^2, calc.tol)
end

function third_order(sys::SuperCellSystem{D}, pot::Potential,
    calc::ForceConstantCalculator; n_threads::Int = Threads.nthreads(),
    check_asr = false) where D

    D == 3 || throw(ArgumentError("the MI355X library computes 3-dimensional systems"))
    N_atoms = n_atoms(sys)
    Ψ = DeviceArray{Float64}(D * N_atoms, D * N_atoms, D * N_atoms)
    third_order_device!(Ψ, nothing, nothing, sys, pot, calc)   # ASR and tolerance fused

    Ψ_unit = energy_unit(pot) / length_unit(pot)^3

    check_asr && @assert asr_satisfied(Ψ, N_atoms, D, calc.tol) "ASR not satisfied for third order force constants, please report this"

    return DenseForceConstants(Array(Ψ), Ψ_unit, calc.tol)
end

# Psi left on the device: FP64 and/or mass-weighted FP32 (the mcc3 operand) in one pass
function third_order_device(sys::SuperCellSystem{3}, pot::Potential, calc::ForceConstantCalculator;
                            fp64 = true, fp32_masses = nothing)
    n = 3 * n_atoms(sys)
    psi64 = fp64 ? DeviceArray{Float64}(n, n, n) : nothing
    psi32 = fp32_masses === nothing ? nothing : DeviceArray{Float32}(n, n, n)
    third_order_device!(psi64, psi32, fp32_masses, sys, pot, calc)
    return psi64, psi32
end

# ---- mass_weight_third_order! (third_order_analytical.jl:185-214) --------------------------
function mass_weight_third_order!(Ψ::ThirdOrderForceConstants, masses::AbstractVector)
    N_modes = size(Ψ)[1]
    D = Int(N_modes / length(masses))
    N_atoms = length(masses)

    @assert D ∈ [1, 2, 3]

    mass_unit = unit(masses[1])
    d = DeviceArray(Array{Float64}(Ψ.values))
    check(ccall((:fc3_mass_weight_third_order, libfc3), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cint, Cint, Ptr{Float64}),
                d.h.ptr, d.ptr, N_atoms, D, mass_vector(masses)))
    Ψ.values .= Array(d)

    Ψ.units /= mass_unit
    return Ψ
end

# ---- mcc3 (src/mcc.jl:29,60,89) ------------------------------------------------------------
const FP32, BF16X3 = Cint(0), Cint(1)
const PRECISION = Ref(FP32)
"""FP32 (the reference's arithmetic, default) or BF16X3 (BF16 matrix cores, 3 products,
relative error <= 1e-5) for the mcc3 methods."""
set_precision!(p::Cint) = (PRECISION[] = p; p)

function mcc3(Ψ::DeviceArray{Float32, 3}, phi::DeviceArray{Float32, 2})
    n = size(phi)[1]
    K3 = DeviceArray{Float32}(Ψ.h, n, n, n)
    check(ccall((:fc3_mcc3, libfc3), Cint,
                (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Cint, Cint, Ptr{Float32}),
                Ψ.h.ptr, Ψ.ptr, phi.ptr, n, PRECISION[], K3.ptr))
    return K3
end

function mcc3!(Ψ::DeviceArray{Float32, 3}, phi::DeviceArray{Float32, 2})
    check(ccall((:fc3_mcc3_inplace, libfc3), Cint,
                (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Cint, Cint),
                Ψ.h.ptr, Ψ.ptr, phi.ptr, size(phi)[1], PRECISION[]))
    return Ψ
end

function mcc3(Ψ::DeviceArray{Float32, 3}, phi::DeviceArray{Float32, 2}, block_size::Int)
    n = size(phi)[1]
    K3_CPU = Array{Float32, 3}(undef, n, n, n)        # host result, as the reference returns
    check(ccall((:fc3_mcc3_blocked, libfc3), Cint,
                (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Cint, Cint, Cint, Ptr{Float32}),
                Ψ.h.ptr, Ψ.ptr, phi.ptr, n, block_size, PRECISION[], K3_CPU))
    return K3_CPU
end

"""
    mcc3_unique(Ψ, phi) -> DeviceArray{Float32,1}

The unique part of K3 (the TODO at src/mcc.jl:116-117): K3[a,b,l] for a <= b <= l (1-based
here) at position (l-1)l(l+1)/6 + (b-1)b/2 + a, each entry bitwise the one `mcc3` computes.
"""
function mcc3_unique(Ψ::DeviceArray{Float32, 3}, phi::DeviceArray{Float32, 2})
    n = size(phi)[1]
    K3u = DeviceArray{Float32}(Ψ.h, n * (n + 1) * (n + 2) ÷ 6)
    check(ccall((:fc3_mcc3_unique, libfc3), Cint,
                (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Cint, Cint, Ptr{Float32}),
                Ψ.h.ptr, Ψ.ptr, phi.ptr, n, PRECISION[], K3u.ptr))
    return K3u
end

"""
    weighted_square_loss(K3, seed; grad = true) -> (L, dK3 or nothing)

L = sum w .* K3.^2 with w ~ U[0, 1) from Philox4x32-10 keyed by `seed`, generated on the
device and reproducible on any host (the differentiation workload's scalar; no
counterpart in the reference). `dK3 = 2 w .* K3` is the cotangent for the mcc3 pullback.
"""
function weighted_square_loss(K3::DeviceArray{Float32}, seed::Integer; grad::Bool = true)
    dK3 = grad ? DeviceArray{Float32}(K3.h, size(K3)...) : nothing
    L = Ref{Float64}(0.0)
    check(ccall((:fc3_weighted_square_loss, libfc3), Cint,
                (Ptr{Cvoid}, Ptr{Float32}, Clonglong, Culonglong, Ptr{Float32}, Ref{Float64}),
                K3.h.ptr, K3.ptr, prod(size(K3)), UInt64(seed),
                grad ? dK3.ptr : Ptr{Float32}(C_NULL), L))
    return L[], dK3
end

"""
    third_order_mcc3(sys, pot, calc, phi; masses = nothing) -> DeviceArray{Float32,3}

K3 = mcc3(Float32.(mass_weight_third_order!(third_order(sys, pot, calc), masses)), phi)
without forming the n^3 Psi (the sparse mode-1 intent of src/todo/old_mcc.jl:17-65): mode 1
runs over Psi's nonzero 3x3x3 blocks as the assembly produces them, modes 2 and 3 in
BF16X3 whatever `set_precision!` says.  `masses = nothing` leaves Psi unweighted.  Needs
3N % 4 == 0; LJ with AnalyticalCalculator or StillingerWeberSilicon with AutoDiffCalculator.
"""
function third_order_mcc3(sys::SuperCellSystem{3}, pot::LJ, calc::AnalyticalCalculator,
                          phi::DeviceArray{Float32, 2}; masses = nothing)
    lu, eu = length_unit(pot), energy_unit(pot)
    n = 3 * n_atoms(sys)
    K3 = DeviceArray{Float32}(phi.h, n, n, n)
    m = masses === nothing ? C_NULL : mass_vector(masses)
    check(ccall((:fc3_third_order_mcc3_lj, libfc3), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cint, Ptr{Float64}, Float64, Float64, Float64,
                 Float64, Float64, Ptr{Float64}, Ptr{Float32}, Ptr{Float32}),
                phi.h.ptr, positions_matrix(sys, lu), n_atoms(sys), box3(sys, lu),
                ustrip(lu, pot.σ), ustrip(eu, pot.ϵ), ustrip(lu, pot.r_cut),
                ustrip(lu, calc.r_cut), calc.tol, m, phi.ptr, K3.ptr))
    return K3
end
function third_order_mcc3(sys::SuperCellSystem{3}, pot::StillingerWeberSilicon,
                          calc::AutoDiffCalculator, phi::DeviceArray{Float32, 2};
                          masses = nothing)
    lu = length_unit(pot)
    n = 3 * n_atoms(sys)
    K3 = DeviceArray{Float32}(phi.h, n, n, n)
    m = masses === nothing ? C_NULL : mass_vector(masses)
    check(ccall((:fc3_third_order_mcc3_sw, libfc3), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cint, Ptr{Float64}, Ptr{Float64}, Float64, Float64,
                 Ptr{Float64}, Ptr{Float32}, Ptr{Float32}),
                phi.h.ptr, positions_matrix(sys, lu), n_atoms(sys), box3(sys, lu),
                sw_abi_params(pot), ustrip(lu, calc.r_cut), calc.tol, m, phi.ptr, K3.ptr))
    return K3
end
third_order_mcc3(sys, pot, calc, phi; masses = nothing) = unsupported(pot, calc)

# get_modes(dynamical_matrix(sys, pot, calc)) (src/dynmat.jl:5-33,123-131) on the device,
# phi left there in FP32 for mcc3
function get_modes_device(sys::SuperCellSystem{3}, pot::LJ, calc::AnalyticalCalculator,
                          num_rigid_translation = 3)
    lu, eu = length_unit(pot), energy_unit(pot)
    n = 3 * n_atoms(sys)
    freqs_sq = Vector{Float64}(undef, n)
    phi = DeviceArray{Float32}(n, n)
    check(ccall((:fc3_modes_lj, libfc3), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cint, Ptr{Float64}, Float64, Float64, Float64,
                 Float64, Float64, Ptr{Float64}, Cint, Ptr{Float64}, Ptr{Float64}, Ptr{Float32}),
                phi.h.ptr, positions_matrix(sys, lu), n_atoms(sys), box3(sys, lu),
                ustrip(lu, pot.σ), ustrip(eu, pot.ϵ), ustrip(lu, pot.r_cut),
                ustrip(lu, calc.r_cut), calc.tol, mass_vector(ForceConstants.masses(sys)),
                num_rigid_translation, freqs_sq, C_NULL, phi.ptr))
    return freqs_sq, phi
end

# ---- output (src/io.jl:9-80): the text format on the device, JLD2 by the reference ---------
function save_second_order(Φ::SecondOrderForceConstants, N_atoms, D, outpath; filename = "second_order", fmt = :JLD2)
    fmt == :txt || return ForceConstants.save_second_order(Φ, N_atoms, D, outpath; filename, fmt)
    @assert ((N_atoms*D) == size(Φ.values)[1] && (N_atoms*D) == size(Φ.values)[2]) "Incorrect dimensions"
    d = DeviceArray(Array{Float64}(Φ.values))
    path = joinpath(outpath, filename * ".txt")
    check(ccall((:fc3_save_second_order_txt, libfc3), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cint, Cint, Cstring, Cstring),
                d.h.ptr, d.ptr, N_atoms, D, string(Φ.units[1, 1]), path))
    return nothing
end

function save_third_order(Ψ::ThirdOrderForceConstants, N_atoms, D, outpath; filename = "third_order", fmt = :JLD2)
    fmt == :txt || return ForceConstants.save_third_order(Ψ, N_atoms, D, outpath; filename, fmt)
    @assert ((N_atoms*D) == size(Ψ.values)[1] && (N_atoms*D) == size(Ψ)[2] && (N_atoms*D) == size(Ψ.values)[3]) "Incorrect dimensions"
    d = DeviceArray(Array{Float64}(Ψ.values))
    path = joinpath(outpath, filename * ".txt")
    check(ccall((:fc3_save_third_order_txt, libfc3), Cint,
                (Ptr{Cvoid}, Ptr{Float64}, Cint, Cint, Cstring, Cstring),
                d.h.ptr, d.ptr, N_atoms, D, string(Ψ.units[1, 1, 1]), path))
    return nothing
end

end # module
