"""ChainRulesCore rules for the GPU energy (loaded with ChainRulesCore): the
reverse-mode rules of total_energy (hamiltonian.jl:70-132) with respect to
the spins, and of CSMCGPU.coupling_energy with respect to the couplings,
from the library's adjoint kernels (csmc_energy_grad_spins,
csmc_energy_grad_couplings), in FP64 on the device."""
module CSMCGPUChainRulesCoreExt

using ChainRulesCore
using ClassicalSpinMC
using ClassicalSpinMC: Lattice
using CSMCGPU

# dE/dS: the energy and its gradient (3 x N, the layout of lat.spins) from one
# device evaluation; the pullback scales the gradient by the cotangent of E
function ChainRulesCore.rrule(::typeof(ClassicalSpinMC.total_energy), lat::Lattice)
    b = CSMCGPU.Batch(CSMCGPU.default_context(), [lat]; precision=64)
    CSMCGPU.set_spins!(b, 1, lat.spins)
    E = CSMCGPU.total_energy(b)[1]
    g = CSMCGPU.energy_grad_spins(b)
    function total_energy_pullback(dE)
        return (NoTangent(), Tangent{Lattice}(; spins=unthunk(dE) .* g))
    end
    return E, total_energy_pullback
end

# dE/dtheta for the couplings theta (layout of CSMCGPU.coupling_vector)
function ChainRulesCore.rrule(::typeof(CSMCGPU.coupling_energy), lat::Lattice,
                              theta::AbstractVector{Float64}; ctx=CSMCGPU.default_context())
    CSMCGPU.update_couplings!(ctx, lat, theta)
    b = CSMCGPU.Batch(ctx, [lat]; precision=64)
    CSMCGPU.set_spins!(b, 1, lat.spins)
    E = CSMCGPU.total_energy(b)[1]
    gJ = vec(CSMCGPU.energy_grad_couplings(b))
    gS = CSMCGPU.energy_grad_spins(b)
    function coupling_energy_pullback(dE)
        d = unthunk(dE)
        return (NoTangent(), Tangent{Lattice}(; spins=d .* gS), d .* gJ)
    end
    return E, coupling_energy_pullback
end

end # module
