"""
    FC3AMDChainRules

ChainRulesCore rules for FC3AMD (the reference has no adjoints; `Project.toml` lists no
ChainRulesCore): the pullback of `mcc3` (dL/dPsi and dL/dphi, fc3_mcc3_backward) and of
the LJ third-order tensor with respect to the potential's sigma and epsilon
(fc3_third_order_lj_param_backward).  Load it beside FC3AMD (as a package extension on
ChainRulesCore, or `include` it after FC3AMD).
"""
module FC3AMDChainRules

using ChainRulesCore
using Unitful: ustrip
using ForceConstants: SuperCellSystem, LJ, AnalyticalCalculator, energy_unit, length_unit, n_atoms
using ..FC3AMD: FC3AMD, DeviceArray, check, libfc3, positions_matrix, box3, mass_vector,
                PRECISION

function ChainRulesCore.rrule(::typeof(FC3AMD.mcc3), Ψ::DeviceArray{Float32, 3},
                              phi::DeviceArray{Float32, 2})
    K3 = FC3AMD.mcc3(Ψ, phi)
    prec = PRECISION[]
    function mcc3_pullback(ΔK3)
        n = size(phi)[1]
        dK3 = ΔK3 isa DeviceArray ? ΔK3 : DeviceArray(Array{Float32}(unthunk(ΔK3)))
        dΨ = DeviceArray{Float32}(Ψ.h, n, n, n)
        dphi = DeviceArray{Float32}(Ψ.h, n, n)
        check(ccall((:fc3_mcc3_backward, libfc3), Cint,
                    (Ptr{Cvoid}, Ptr{Float32}, Ptr{Float32}, Ptr{Float32}, Cint, Cint,
                     Ptr{Float32}, Ptr{Float32}),
                    Ψ.h.ptr, Ψ.ptr, phi.ptr, dK3.ptr, n, prec, dΨ.ptr, dphi.ptr))
        return NoTangent(), dΨ, dphi
    end
    return K3, mcc3_pullback
end

"""
    third_order_mw32(sys, pot::LJ, calc::AnalyticalCalculator, masses)

The mass-weighted FP32 third-order tensor on the device (the mcc3 operand), as a function
of the LJ parameters, with a ChainRulesCore rule returning `Tangent{LJ}(σ = ..., ϵ = ...)`.
"""
function third_order_mw32(sys::SuperCellSystem{3}, pot::LJ, calc::AnalyticalCalculator, masses)
    _, psi32 = FC3AMD.third_order_device(sys, pot, calc; fp64 = false, fp32_masses = masses)
    return psi32
end

function ChainRulesCore.rrule(::typeof(third_order_mw32), sys::SuperCellSystem{3}, pot::LJ,
                              calc::AnalyticalCalculator, masses)
    Ψ = third_order_mw32(sys, pot, calc, masses)
    function third_order_pullback(ΔΨ)
        lu, eu = length_unit(pot), energy_unit(pot)
        dΨ = ΔΨ isa DeviceArray ? ΔΨ : DeviceArray(Array{Float32}(unthunk(ΔΨ)))
        g = zeros(2)
        check(ccall((:fc3_third_order_lj_param_backward, libfc3), Cint,
                    (Ptr{Cvoid}, Ptr{Float64}, Cint, Ptr{Float64}, Float64, Float64, Float64,
                     Float64, Float64, Ptr{Float64}, Ptr{Float32}, Ptr{Float64}, Ptr{Float64}),
                    dΨ.h.ptr, positions_matrix(sys, lu), n_atoms(sys), box3(sys, lu),
                    ustrip(lu, pot.σ), ustrip(eu, pot.ϵ), ustrip(lu, pot.r_cut),
                    ustrip(lu, calc.r_cut), calc.tol, mass_vector(masses), dΨ.ptr, C_NULL, g))
        return NoTangent(), NoTangent(), Tangent{LJ}(; σ = g[1], ϵ = g[2]), NoTangent(),
               NoTangent()
    end
    return Ψ, third_order_pullback
end

end # module
