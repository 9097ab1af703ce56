# VespaHIPChainRulesExt — ChainRulesCore rules for the HIP reverse mode.
#
# A weak-dependency extension of VespaHIPFluxExt: registered in Project.toml as
#   [weakdeps]   ChainRulesCore = "d360d2e6-b24c-11e9-a2a3-2a2ae2dbcce4"
#   [extensions] VespaHIPChainRulesExt = "ChainRulesCore"
# so it loads only when ChainRulesCore does. The reference has no derivative of
# its step (no AD anywhere in the tree); these rules make step!(sim, dt)
# (src/driver.jl:333-359) differentiable with respect to the conserved state,
# dt held fixed, through the hand-written adjoint kernels:
#   * hip_step_state(g, U, dt): one step; pullback = vh_grid_step_vjp;
#   * hip_rollout_state(r, U₀, cfl): a checkpointed rollout with the device CFL
#     dt; pullback = vh_adjoint_backward (segments recomputed from checkpoints).
# Non-smooth points take the active branch (minmod's selected argument or zero,
# the active Davis min/max argument, the upwind HLLC branch, the pressure floor
# when it is strictly the larger argument).
module VespaHIPChainRulesExt

using ChainRulesCore
using ..VespaHIPFluxExt: HIPGrid, HIPRollout, hip_step_state, hip_step_vjp, hip_rollout_state, hip_backward

function ChainRulesCore.rrule(::typeof(hip_step_state), g::HIPGrid, U::Vector{Float64}, dt::Real)
    Unew = hip_step_state(g, U, dt)
    function hip_step_state_pullback(Ū)
        Ūv = Vector{Float64}(unthunk(Ū))
        return NoTangent(), NoTangent(), hip_step_vjp(g, U, dt, Ūv), NoTangent()
    end
    return Unew, hip_step_state_pullback
end

function ChainRulesCore.rrule(::typeof(hip_rollout_state), r::HIPRollout, U₀::Vector{Float64}, cfl::Real)
    UN = hip_rollout_state(r, U₀, cfl)
    # the pullback replays the checkpoints and dt schedule of this forward pass;
    # a later forward on the same HIPRollout replaces them
    function hip_rollout_state_pullback(Ū)
        Ūv = Vector{Float64}(unthunk(Ū))
        return NoTangent(), NoTangent(), hip_backward(r, Ūv), NoTangent()
    end
    return UN, hip_rollout_state_pullback
end

end # module
