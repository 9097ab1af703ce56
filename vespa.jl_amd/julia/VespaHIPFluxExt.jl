# VespaHIPFluxExt — Julia binding of libvespahip (gfx950) behind Vespa.jl's flux seam.
#
# Drop-in for the KernelAbstractions device backend (ext/VespaKAFluxExt.jl):
#   * tier 1: `HIPFlux <: Vespa.AbstractFluxBackend` with a
#     `Vespa.compute_face_fluxes!` method (contract src/ka_flux.jl:40-53), so
#     `sim.flux = HIPFlux()` routes `accumulate_flux!` (src/driver.jl:210-211)
#     through the batched gather → GPU HLLC → scatter path unchanged;
#   * tier 2: `HIPGrid`, a device-resident uniform grid whose `hip_step!` /
#     `hip_evolve!` replace `step!` / `evolve!` (src/driver.jl:333-359, 483-508)
#     for RefMesh `UniformMesh` simulations, and `hip_set_mode!` selecting the
#     exact (bit-identical) or fast arithmetic profile;
#   * `HIPBlocks`: block-structured AMR meshes (hanging-node faces, flux
#     registers, subcycling) advanced by one persistent launch per RK stage;
#   * `hip_step_vjp`: the pullback of one step, and `HIPRollout` (`hip_forward!` /
#     `hip_backward`): a checkpointed rollout; `hip_step_state` / `hip_rollout_state`
#     are the state-to-state functions whose ChainRulesCore `rrule`s live in
#     VespaHIPChainRulesExt.jl (a weak-dependency extension);
#   * `hip_step_bflux!`: step! with the boundary-flux register (src/reflux.jl);
#   * `HIPMuscl`: PPMKernels' unsplit RK2 MUSCL step with flux recording
#     (`hip_fluxrec`, `hip_flux_register`);
#   * `hip_enable_gravity!` / `hip_solve_poisson!`: self-gravity of a HIPGrid
#     (enable_gravity! / solve_poisson!, src/gravity.jl:63-80, 248-322): the
#     device CG Poisson solve, the gravity source in every RK stage and the
#     free-fall limiter in the device dt.
#
# Load with `include` after `using Vespa`; set ENV["VESPAHIP_LIB"] or keep
# libvespahip.so next to the package's vespa.jl_amd/ directory.
module VespaHIPFluxExt

using Vespa
using Vespa: AbstractFluxBackend, adiabatic_index
using MeshInterface: Outflow, ParentGhost, Periodic, Reflecting, bc_on

const libvespahip = get(ENV, "VESPAHIP_LIB", joinpath(@__DIR__, "..", "libvespahip.so"))

const VH_F64 = Int32(0)
const VH_F32 = Int32(1)
const VH_EXACT = Int32(0)
const VH_FAST = Int32(1)

_check(rc) = rc == 0 || error("vespahip: ", unsafe_string(ccall((:vh_last_error, libvespahip), Cstring, ())))

# ------------------------------------------------------------------ tier 1

"GPU flux backend: the per-face HLLC of `compute_face_fluxes!` on gfx950."
struct HIPFlux <: AbstractFluxBackend end

function Vespa.compute_face_fluxes!(::HIPFlux, model, F::Vector{NTuple{5,T}}, WL::Vector{NTuple{5,T}},
                                    WR::Vector{NTuple{5,T}}, axisv::Vector{Int}) where {T<:Union{Float64,Float32}}
    nf = length(F)
    nf == 0 && return F
    ax = convert(Vector{Int32}, axisv)
    γ = Float64(adiabatic_index(model))
    fn = T === Float64 ? :vh_face_flux_f64 : :vh_face_flux_f32
    GC.@preserve F WL WR ax begin
        pF = Ptr{T}(pointer(F)); pL = Ptr{T}(pointer(WL)); pR = Ptr{T}(pointer(WR))
        rc = T === Float64 ?
            ccall((:vh_face_flux_f64, libvespahip), Cint,
                  (Ptr{Float64}, Ptr{Float64}, Ptr{Int32}, Int64, Float64, Ptr{Float64}),
                  pL, pR, ax, nf, γ, pF) :
            ccall((:vh_face_flux_f32, libvespahip), Cint,
                  (Ptr{Float32}, Ptr{Float32}, Ptr{Int32}, Int64, Float64, Ptr{Float32}),
                  pL, pR, ax, nf, γ, pF)
        _check(rc)
    end
    return F
end

# ------------------------------------------------------------------ tier 2

_bc_code(::Periodic) = Int32(0)
_bc_code(::Outflow) = Int32(1)
_bc_code(::Reflecting) = Int32(2)
_bc_code(::ParentGhost) = Int32(3)
_bc_code(bc) = error("vespahip: boundary condition $(typeof(bc)) is not supported on the device path")

"Device-resident state of a `Simulation` on a RefMesh `UniformMesh`."
mutable struct HIPGrid
    ptr::Ptr{Cvoid}
    ncell::Int
    eltype::DataType
    rank::Int
end

function HIPGrid(sim::Vespa.Simulation)
    b = sim.backend
    N = Vespa.rank(b)
    dims = Int64[i <= N ? b.dims[i] : 1 for i in 1:3]
    lo = Float64[i <= N ? b.lo[i] : 0.0 for i in 1:3]
    hi = Float64[i <= N ? b.hi[i] : 1.0 for i in 1:3]
    bc = Int32[]
    for d in 1:3
        if d <= N
            push!(bc, _bc_code(bc_on(sim.bcs, d, :lo)), _bc_code(bc_on(sim.bcs, d, :hi)))
        else
            push!(bc, Int32(1), Int32(1))
        end
    end
    T = eltype(sim.sv[1])
    h = Ref{Ptr{Cvoid}}(C_NULL)
    _check(ccall((:vh_grid_create, libvespahip), Cint,
                 (Ref{Ptr{Cvoid}}, Int32, Ptr{Int64}, Ptr{Float64}, Ptr{Float64}, Ptr{Int32}, Float64, Int32),
                 h, Int32(N), dims, lo, hi, bc, Float64(sim.γ), T === Float64 ? VH_F64 : VH_F32))
    g = HIPGrid(h[], prod(dims), T, N)
    finalizer(g) do x
        x.ptr == C_NULL || ccall((:vh_grid_destroy, libvespahip), Cint, (Ptr{Cvoid},), x.ptr)
        x.ptr = C_NULL
    end
    hip_upload!(g, sim)
    return g
end

"Copy the simulation's SoA conserved fields (one dense array per field) to the device."
function hip_upload!(g::HIPGrid, sim)
    buf = Vector{g.eltype}(undef, 5 * g.ncell)
    for f in 1:5
        copyto!(buf, (f - 1) * g.ncell + 1, vec(sim.sv[f]), 1, g.ncell)
    end
    GC.@preserve buf _check(ccall((:vh_grid_upload, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), g.ptr, buf))
    return g
end

"Copy the device state back into the simulation's field arrays."
function hip_download!(sim, g::HIPGrid)
    buf = Vector{g.eltype}(undef, 5 * g.ncell)
    GC.@preserve buf _check(ccall((:vh_grid_download, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), g.ptr, buf))
    for f in 1:5
        copyto!(vec(sim.sv[f]), 1, buf, (f - 1) * g.ncell + 1, g.ncell)
    end
    return sim
end

hip_set_mode!(g::HIPGrid, fast::Bool) =
    (_check(ccall((:vh_grid_set_mode, libvespahip), Cint, (Ptr{Cvoid}, Int32), g.ptr, fast ? VH_FAST : VH_EXACT)); g)

"""
    hip_step!(sim, g::HIPGrid, dt)

step!(sim, dt) on the device (src/driver.jl:333-359). Advances `sim.t` and
`sim.step` only: the new state stays in device memory, so `sim`'s field arrays
are stale until `hip_download!(sim, g)` (keeping the state resident is what
removes the per-step PCIe round trip). `hip_evolve!` downloads at its end.
"""
# The ParentGhost closures of sim (src/driver.jl:127-134) evaluated for every
# boundary face of their sides, in the boundary-flux slot order, handed to the device.
function _refresh_parent!(sim, g::HIPGrid)
    N = g.rank
    b = sim.backend
    dims = ntuple(d -> b.dims[d], N)
    any(bc_on(sim.bcs, d, s) isa ParentGhost for d in 1:N for s in (:lo, :hi)) || return g
    states = zeros(g.eltype, 5 * hip_bflux_count(g))
    slot = 0
    for d in 1:N, s in (:lo, :hi)
        bc = bc_on(sim.bcs, d, s)
        other = [e for e in 1:N if e != d]
        plane = CartesianIndices(ntuple(i -> 1:dims[other[i]], N - 1))
        for J in plane
            if bc isa ParentGhost
                I = ntuple(e -> e == d ? (s === :lo ? 1 : dims[d]) : J[findfirst(==(e), other)], N)
                W = bc.ghost(d, s, CartesianIndex(I))
                for f in 1:5
                    states[5 * slot + f] = W[f]
                end
            end
            slot += 1
        end
    end
    return hip_set_parent_ghost!(g, states)
end

function hip_step!(sim, g::HIPGrid, dt::Real)
    _refresh_parent!(sim, g)
    _check(ccall((:vh_grid_step, libvespahip), Cint, (Ptr{Cvoid}, Float64), g.ptr, Float64(dt)))
    sim.t += dt
    sim.step += 1
    return sim
end

"evolve!(sim) on the device (src/driver.jl:483-508): CFL dt reduced on the GPU each step."
function hip_evolve!(sim, g::HIPGrid; max_steps::Integer = 0)
    t = Ref(Float64(sim.t))
    k = Ref{Int64}(0)
    _check(ccall((:vh_grid_run, libvespahip), Cint, (Ptr{Cvoid}, Float64, Float64, Int64, Ref{Float64}, Ref{Int64}),
                 g.ptr, Float64(sim.problem.cfl), Float64(sim.problem.tfinal), Int64(max_steps), t, k))
    sim.t = t[]
    sim.step += k[]
    hip_download!(sim, g)
    return sim
end

"conserved_totals(sim) (src/diagnostics.jl:16-31) from the device state."
function hip_conserved_totals(g::HIPGrid)
    out = zeros(Float64, 5)
    _check(ccall((:vh_grid_totals, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Float64}), g.ptr, out))
    return (mass = out[1], momentum_x = out[2], momentum_y = out[3], momentum_z = out[4], energy = out[5])
end

# ------------------------------------------------------------------ batched blocks

"""
    HIPBlocks(rank, maxlevel, h0, γ, level, dims, origin, bc; eltype = Float64)

Device-resident block mesh (block-structured AMR with hanging nodes): `level[b]`,
`dims[:, b]` (cells per axis), `origin[:, b]` (integer lattice of the finest cells)
and `bc[:, b]` (6 codes, used where no block touches a face). Leaves are ordered
block by block, x fastest; faces follow HGBackend (lib/HGBackend/src/HGBackend.jl:202-276).
"""
mutable struct HIPBlocks
    ptr::Ptr{Cvoid}
    ncell::Int
    eltype::DataType
end

function HIPBlocks(rank::Integer, maxlevel::Integer, h0::AbstractVector, γ::Real, level::AbstractVector,
                   dims::AbstractMatrix, origin::AbstractMatrix, bc::AbstractMatrix; eltype = Float64)
    nb = length(level)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    lv = Int32.(level); dm = Int32.(dims); og = Int64.(origin); bcs = Int32.(bc)
    h0v = Float64[i <= length(h0) ? h0[i] : 1.0 for i in 1:3]
    _check(ccall((:vh_blocks_create, libvespahip), Cint,
                 (Ref{Ptr{Cvoid}}, Int32, Int32, Int32, Ptr{Float64}, Float64, Ptr{Int32}, Ptr{Int32}, Ptr{Int64},
                  Ptr{Int32}, Int32),
                 h, Int32(rank), Int32(nb), Int32(maxlevel), h0v, Float64(γ), lv, dm, og, bcs,
                 eltype === Float64 ? VH_F64 : VH_F32))
    n = Int(ccall((:vh_blocks_ncell, libvespahip), Int64, (Ptr{Cvoid},), h[]))
    m = HIPBlocks(h[], n, eltype)
    finalizer(m) do x
        x.ptr == C_NULL || ccall((:vh_blocks_destroy, libvespahip), Cint, (Ptr{Cvoid},), x.ptr)
        x.ptr = C_NULL
    end
    return m
end

"Upload the SoA conserved state (5 × ncell, field-major) of a block mesh."
hip_upload!(m::HIPBlocks, U::Vector) = (GC.@preserve U _check(ccall((:vh_blocks_upload, libvespahip), Cint,
    (Ptr{Cvoid}, Ptr{Cvoid}), m.ptr, U)); m)

function hip_download(m::HIPBlocks)
    U = Vector{m.eltype}(undef, 5 * m.ncell)
    GC.@preserve U _check(ccall((:vh_blocks_download, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), m.ptr, U))
    return U
end

"step!(sim, dt; level) (src/driver.jl:333-378); level < 0 updates every leaf."
hip_step!(m::HIPBlocks, dt::Real; level::Integer = -1) =
    (_check(ccall((:vh_blocks_step, libvespahip), Cint, (Ptr{Cvoid}, Float64, Int32), m.ptr, Float64(dt),
                  Int32(level))); m)

"evolve! single-rate or subcycled with flux registers (src/driver.jl:483-508, 553-605; src/reflux.jl:51-103)."
function hip_evolve!(m::HIPBlocks, cfl::Real, tfinal::Real; t0::Real = 0.0, subcycle::Bool = false,
                     max_steps::Integer = 0)
    t = Ref(Float64(t0)); k = Ref{Int64}(0)
    _check(ccall((:vh_blocks_run, libvespahip), Cint,
                 (Ptr{Cvoid}, Float64, Float64, Int64, Int32, Ref{Float64}, Ref{Int64}),
                 m.ptr, Float64(cfl), Float64(tfinal), Int64(max_steps), Int32(subcycle), t, k))
    return t[], k[]
end

"""
    hip_indicator(m::HIPBlocks, nblocks) -> Vector{Float64}

density_gradient_indicator (src/driver.jl:400-414) of every leaf, evaluated on the
device; returns its maximum per block, all `regrid!`'s tagging needs.
"""
function hip_indicator(m::HIPBlocks, nblocks::Integer)
    bmax = Vector{Float64}(undef, nblocks)
    _check(ccall((:vh_blocks_indicator, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}),
                 m.ptr, bmax, C_NULL))
    return bmax
end

"""
    hip_remap!(dst::HIPBlocks, src::HIPBlocks, src_block)

Conservative transfer of `src`'s state onto `dst`'s refined mesh on the device:
block b of dst takes the cells of block src_block[b] (1-based: the same block or
its ancestor) by injection (lib/RefMesh/src/RefMesh.jl:284-302).
"""
function hip_remap!(dst::HIPBlocks, src::HIPBlocks, src_block::AbstractVector{<:Integer})
    sb = Int32.(src_block) .- Int32(1)
    _check(ccall((:vh_blocks_remap, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Int32}), dst.ptr, src.ptr, sb))
    return dst
end

"""
    hip_regrid!(m::HIPBlocks, level, dims, origin, bc, src_block)

regrid! in place (src/driver.jl:440-462): `m` moves onto the new block table
(same rank, maxlevel, h0, γ), block b taking the cells of current block
src_block[b] (1-based: the same block or its ancestor) by injection
(lib/RefMesh/src/RefMesh.jl:284-302); the device buffers are reused.
"""
function hip_regrid!(m::HIPBlocks, level::AbstractVector, dims::AbstractMatrix, origin::AbstractMatrix,
                     bc::AbstractMatrix, src_block::AbstractVector{<:Integer})
    lv = Int32.(level); dm = Int32.(dims); og = Int64.(origin); bcs = Int32.(bc)
    sb = Int32.(src_block) .- Int32(1)
    _check(ccall((:vh_blocks_regrid, libvespahip), Cint,
                 (Ptr{Cvoid}, Int32, Ptr{Int32}, Ptr{Int32}, Ptr{Int64}, Ptr{Int32}, Ptr{Int32}),
                 m.ptr, Int32(length(lv)), lv, dm, og, bcs, sb))
    m.ncell = Int(ccall((:vh_blocks_ncell, libvespahip), Int64, (Ptr{Cvoid},), m.ptr))
    return m
end

# ------------------------------------------------------------------ PPMKernels MUSCL

"""
    HIPMuscl(dims, ng; gamma, theta=1.5, dx=1.0, small_rho=1e-10, eltype=Float64)

Device context for PPMKernels' `muscl_step_3d!` on a ghosted 3D grid
(lib/PPMKernels/src/muscl_grid.jl:314-370): state SoA (D, S1, S2, S3, τ), x fastest.
"""
mutable struct HIPMuscl
    ptr::Ptr{Cvoid}
    ncell::Int
    eltype::DataType
    nfields::Int        # 5, or 6 with the gas energy ρe (dual energy)
end

function HIPMuscl(dims::NTuple{3,Int}, ng::Integer; gamma::Real, theta::Real = 1.5, dx::Real = 1.0,
                  small_rho::Real = 1e-10, eltype = Float64, dual::Bool = false, eta1::Real = 1e-3)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    _check(ccall((:vh_muscl_create, libvespahip), Cint,
                 (Ref{Ptr{Cvoid}}, Ptr{Int32}, Int32, Float64, Float64, Float64, Float64, Int32),
                 h, Int32[dims...], Int32(ng), Float64(gamma), Float64(theta), Float64(small_rho), Float64(dx),
                 eltype === Float64 ? VH_F64 : VH_F32))
    m = HIPMuscl(h[], prod(dims), eltype, dual ? 6 : 5)
    finalizer(m) do x
        x.ptr == C_NULL || ccall((:vh_muscl_destroy, libvespahip), Cint, (Ptr{Cvoid},), x.ptr)
        x.ptr = C_NULL
    end
    dual && _check(ccall((:vh_muscl_enable_dual_energy, libvespahip), Cint, (Ptr{Cvoid}, Float64), m.ptr,
                         Float64(eta1)))
    return m
end

hip_upload!(m::HIPMuscl, U::Vector) = (GC.@preserve U _check(ccall((:vh_muscl_upload, libvespahip), Cint,
    (Ptr{Cvoid}, Ptr{Cvoid}), m.ptr, U)); m)

function hip_download(m::HIPMuscl)
    U = Vector{m.eltype}(undef, m.nfields * m.ncell)
    GC.@preserve U _check(ccall((:vh_muscl_download, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), m.ptr, U))
    return U
end

"muscl_step_3d!(…; dt, fluxrec) on the device state; record = fluxrec !== nothing."
hip_muscl_step!(m::HIPMuscl, dt::Real; record::Bool = false) =
    (_check(ccall((:vh_muscl_step, libvespahip), Cint, (Ptr{Cvoid}, Float64, Int32), m.ptr, Float64(dt),
                  Int32(record))); m)

"Exact (bit-identical, default) or fast (contracted, refined reciprocals) stage kernels."
hip_set_mode!(m::HIPMuscl, fast::Bool) =
    (_check(ccall((:vh_muscl_set_mode, libvespahip), Cint, (Ptr{Cvoid}, Int32), m.ptr, fast ? VH_FAST : VH_EXACT)); m)

hip_fill_periodic!(m::HIPMuscl) = (_check(ccall((:vh_muscl_fill_periodic, libvespahip), Cint, (Ptr{Cvoid},),
                                                m.ptr)); m)

"""
    hip_set_bc!(m::HIPMuscl, bc)

bc! of muscl_step_3d! (lib/PPMKernels/src/muscl_grid.jl:316-322, 339, 350), run
before each RK stage: `nothing`, `:periodic` (fill_periodic! on the device), or a
C function pointer `fn` (from `@cfunction(f, Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}, Int32))`)
called as `f(user, state, stream, stage)` with the device state of the stage and the
library's HIP stream; `user` is passed through (`hip_set_bc!(m, fn, user)`).
"""
function hip_set_bc!(m::HIPMuscl, bc, user::Ptr{Cvoid} = C_NULL)
    kind, fn = bc === nothing ? (Int32(0), C_NULL) :
               bc === :periodic ? (Int32(1), C_NULL) :
               bc isa Ptr{Cvoid} ? (Int32(2), bc) : throw(ArgumentError("bc must be nothing, :periodic or a Ptr{Cvoid}"))
    _check(ccall((:vh_muscl_set_bc, libvespahip), Cint, (Ptr{Cvoid}, Int32, Ptr{Cvoid}, Ptr{Cvoid}),
                 m.ptr, kind, fn, kind == 2 ? user : C_NULL))
    return m
end

function hip_max_wavespeed(m::HIPMuscl)
    s = Ref{Float64}(0.0)
    _check(ccall((:vh_muscl_max_wavespeed, libvespahip), Cint, (Ptr{Cvoid}, Ref{Float64}), m.ptr, s))
    return s[]
end

# ------------------------------------------------------------------ ParentGhost

"""
    hip_set_parent_ghost!(g::HIPGrid, states::Vector) -> g

The states of the grid's ParentGhost sides (`VH_BC_PARENT`; MeshInterface's
`ParentGhost(ghost)`, src/driver.jl:127-134): `ghost(axis, side, cell)` evaluated for
every boundary face, 5 primitives (ρ, vx, vy, vz, p) of the state type per face in the
slot order of the boundary-flux register (`hip_bflux_count(g)` faces). Held for every
step until the next call; call it before each step when the parent data change.
"""
function hip_set_parent_ghost!(g::HIPGrid, states::Vector)
    length(states) == 5 * hip_bflux_count(g) || error("states must hold 5 × hip_bflux_count(g) values")
    GC.@preserve states _check(ccall((:vh_grid_set_parent_ghost, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}),
                                     g.ptr, states))
    return g
end

# ------------------------------------------------------------------ boundary-flux register

"""
    hip_step_bflux!(g::HIPGrid, dt, reg::Vector) -> reg

step!(sim, dt; bflux) (src/driver.jl:333-359 with src/reflux.jl:105-132): one step that
also adds ½dt·F·area of every domain-boundary face, per RK stage, to `reg` (5 values of
the state type per slot, `hip_bflux_count(g)` slots; planes axis 1 lo, axis 1 hi, axis 2
lo, …, each column-major over the other axes). The register accumulates, as the
reference's does; zero it to start a new interval.
"""
function hip_step_bflux!(g::HIPGrid, dt::Real, reg::Vector)
    nb = sizeof(eltype(reg)) * length(reg)
    length(reg) == 5 * hip_bflux_count(g) || error("register must hold 5 × hip_bflux_count(g) values")
    d = Ref{Ptr{Cvoid}}(C_NULL)
    _check(ccall((:vh_buffer_alloc, libvespahip), Cint, (Ref{Ptr{Cvoid}}, Int64), d, nb))
    try
        GC.@preserve reg begin
            _check(ccall((:vh_buffer_upload, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int64), d[], reg, nb))
            _check(ccall((:vh_grid_step_bflux, libvespahip), Cint, (Ptr{Cvoid}, Float64, Ptr{Cvoid}), g.ptr,
                         Float64(dt), d[]))
            _check(ccall((:vh_buffer_download, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int64), reg, d[], nb))
        end
    finally
        ccall((:vh_buffer_free, libvespahip), Cint, (Ptr{Cvoid},), d[])
    end
    return reg
end

hip_bflux_count(g::HIPGrid) = Int(ccall((:vh_grid_bflux_count, libvespahip), Int64, (Ptr{Cvoid},), g.ptr))

# ------------------------------------------------------------------ MUSCL flux recording

"The fluxes recorded by the last `hip_muscl_step!(…; record = true)`: 3 × nfields × ncell, [axis][field][cell]."
function hip_fluxrec(m::HIPMuscl)
    out = Vector{m.eltype}(undef, 3 * m.nfields * m.ncell)
    GC.@preserve out _check(ccall((:vh_muscl_fluxrec_download, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), m.ptr, out))
    return out
end

"""
    hip_flux_register(m::HIPMuscl, dt, act) -> (planes, interior)

boundary_flux_register (lib/PPMKernels/src/reflux_capture.jl:86-131) of the last recording
step as Float64 ∫F·area·dt arrays, axis by axis (layout in include/vespahip.h); act = the
active cells per axis (dims .- 2ng).
"""
function hip_flux_register(m::HIPMuscl, dt::Real, act::NTuple{3,Int})
    np = sum(10 * prod(act[[b for b in 1:3 if b != a]]) for a in 1:3)
    ni = sum(5 * (act[a] - 1) * prod(act[[b for b in 1:3 if b != a]]) for a in 1:3)
    planes = Vector{Float64}(undef, np)
    interior = Vector{Float64}(undef, max(ni, 1))
    GC.@preserve planes interior _check(ccall((:vh_muscl_flux_register, libvespahip), Cint,
        (Ptr{Cvoid}, Float64, Ptr{Float64}, Ptr{Float64}), m.ptr, Float64(dt), planes, interior))
    return planes, interior
end

# ------------------------------------------------------------------ reverse mode

"""
    hip_step_vjp(g::HIPGrid, U, dt, Ū) -> ∂L/∂U

Pullback of `step!(sim, dt)` with dt held fixed (the body of a ChainRulesCore
`rrule`): U and Ū are host vectors of 5 × ncell Float64 (field-major).
"""
function hip_step_vjp(g::HIPGrid, U::Vector{Float64}, dt::Real, Ū::Vector{Float64})
    nb = 8 * length(U)
    dU = Ref{Ptr{Cvoid}}(C_NULL); dW = Ref{Ptr{Cvoid}}(C_NULL); dO = Ref{Ptr{Cvoid}}(C_NULL)
    for r in (dU, dW, dO)
        _check(ccall((:vh_buffer_alloc, libvespahip), Cint, (Ref{Ptr{Cvoid}}, Int64), r, nb))
    end
    out = similar(U)
    try
        GC.@preserve U Ū out begin
            _check(ccall((:vh_buffer_upload, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int64), dU[], U, nb))
            _check(ccall((:vh_buffer_upload, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int64), dW[], Ū, nb))
            _check(ccall((:vh_grid_step_vjp, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Float64, Ptr{Cvoid}, Ptr{Cvoid}),
                         g.ptr, dU[], Float64(dt), dW[], dO[]))
            _check(ccall((:vh_buffer_download, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int64), out, dO[], nb))
        end
    finally
        for r in (dU, dW, dO)
            ccall((:vh_buffer_free, libvespahip), Cint, (Ptr{Cvoid},), r[])
        end
    end
    return out
end

"""
    hip_enable_gravity!(g::HIPGrid; G=1.0, bcs=Periodic(), maxiter=500, tol, project_mean)

enable_gravity! (src/gravity.jl:63-80) for the device grid: φ starts at zero;
non-periodic φ sides are homogeneous Neumann. `hip_evolve!` then solves the
Poisson equation before each step (src/driver.jl:498).
"""
function hip_enable_gravity!(g::HIPGrid; G::Real = 1.0, bcs = Periodic(), maxiter::Integer = 500,
                             tol::Real = (g.eltype === Float32 ? 1.0e-5 : 1.0e-8),
                             project_mean::Bool = (bcs isa Periodic))
    rank = g.rank
    mask = 0
    for d in 1:rank
        b = bcs isa Tuple ? bcs[d] : bcs
        b isa Periodic && (mask |= 1 << (d - 1))
    end
    _check(ccall((:vh_grid_enable_gravity, libvespahip), Cint, (Ptr{Cvoid}, Float64, Int32, Int64, Float64, Int32),
                 g.ptr, Float64(G), Int32(mask), Int64(maxiter), Float64(tol), Int32(project_mean)))
    return g
end

"solve_poisson! (src/gravity.jl:248-322) on the device: returns (iterations, relative residual)."
function hip_solve_poisson!(g::HIPGrid)
    it = Ref{Int64}(0)
    rr = Ref{Float64}(0.0)
    _check(ccall((:vh_grid_solve_poisson, libvespahip), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Float64}), g.ptr, it, rr))
    return (it[], rr[])
end

"φ of the device grid as a flat host vector (x fastest)."
function hip_phi(g::HIPGrid)
    out = Vector{g.eltype}(undef, g.ncell)
    GC.@preserve out _check(ccall((:vh_grid_phi_download, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), g.ptr, out))
    return out
end

"""
    hip_enable_gravity!(m::HIPBlocks; G=1.0, maxiter=500, tol, project_mean=true)

enable_gravity! (src/gravity.jl:63-80) on a block mesh: the composite Poisson
operator of the mesh (coarse↔fine rows, src/gravity.jl:455-600) built once, φ = 0.
`hip_evolve!` (single rate) then solves before each step; every stage adds the
gravity source and the device dt folds in the free-fall limiter.
"""
function hip_enable_gravity!(m::HIPBlocks; G::Real = 1.0, maxiter::Integer = 500,
                             tol::Real = (m.eltype === Float32 ? 1.0e-5 : 1.0e-8), project_mean::Bool = true)
    _check(ccall((:vh_blocks_enable_gravity, libvespahip), Cint, (Ptr{Cvoid}, Float64, Int64, Float64, Int32),
                 m.ptr, Float64(G), Int64(maxiter), Float64(tol), Int32(project_mean)))
    return m
end

function hip_solve_poisson!(m::HIPBlocks)
    it = Ref{Int64}(0)
    rr = Ref{Float64}(0.0)
    _check(ccall((:vh_blocks_solve_poisson, libvespahip), Cint, (Ptr{Cvoid}, Ref{Int64}, Ref{Float64}), m.ptr, it, rr))
    return (it[], rr[])
end

"φ of the block mesh as a flat host vector in leaf order."
function hip_phi(m::HIPBlocks)
    out = Vector{m.eltype}(undef, m.ncell)
    GC.@preserve out _check(ccall((:vh_blocks_phi_download, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), m.ptr, out))
    return out
end

"""
    HIPRollout(g::HIPGrid, nsteps; every = 20)

Checkpointed rollout of `nsteps` steps with the device CFL dt: `hip_forward!` keeps the
state every `every` steps and each step's dt, `hip_backward` recomputes each segment and
applies the step VJPs in reverse (the pullback of the whole rollout).
"""
mutable struct HIPRollout
    ptr::Ptr{Cvoid}
    grid::HIPGrid
    nsteps::Int
end

function HIPRollout(g::HIPGrid, nsteps::Integer; every::Integer = 20)
    h = Ref{Ptr{Cvoid}}(C_NULL)
    _check(ccall((:vh_adjoint_create, libvespahip), Cint, (Ref{Ptr{Cvoid}}, Ptr{Cvoid}, Int64, Int32),
                 h, g.ptr, Int64(nsteps), Int32(every)))
    r = HIPRollout(h[], g, Int(nsteps))
    finalizer(r) do x
        x.ptr == C_NULL || ccall((:vh_adjoint_destroy, libvespahip), Cint, (Ptr{Cvoid},), x.ptr)
        x.ptr = C_NULL
    end
    return r
end

"Run the rollout forward from the grid's current state (left at the final state)."
hip_forward!(r::HIPRollout, cfl::Real) =
    (_check(ccall((:vh_adjoint_forward, libvespahip), Cint, (Ptr{Cvoid}, Float64), r.ptr, Float64(cfl))); r)

"∂L/∂U₀ for a loss whose gradient with respect to the final state is W̄ (host vectors, 5 × ncell Float64)."
function hip_backward(r::HIPRollout, W̄::Vector{Float64})
    nb = 8 * length(W̄)
    dW = Ref{Ptr{Cvoid}}(C_NULL); dU = Ref{Ptr{Cvoid}}(C_NULL)
    for d in (dW, dU)
        _check(ccall((:vh_buffer_alloc, libvespahip), Cint, (Ref{Ptr{Cvoid}}, Int64), d, nb))
    end
    out = similar(W̄)
    try
        GC.@preserve W̄ out begin
            _check(ccall((:vh_buffer_upload, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int64), dW[], W̄, nb))
            _check(ccall((:vh_adjoint_backward, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Ptr{Cvoid}),
                         r.ptr, dW[], dU[]))
            _check(ccall((:vh_buffer_download, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}, Int64), out, dU[], nb))
        end
    finally
        for d in (dW, dU)
            ccall((:vh_buffer_free, libvespahip), Cint, (Ptr{Cvoid},), d[])
        end
    end
    return out
end

"""
    hip_step_state(g::HIPGrid, U, dt) -> U′

The step as a function of the state (the form `rrule` differentiates): uploads U
(5 × ncell Float64, field-major), runs one step!(·, dt) on the device and returns
the new state as a fresh host vector. The grid's device state is overwritten.
"""
function hip_step_state(g::HIPGrid, U::Vector{Float64}, dt::Real)
    g.eltype === Float64 || error("hip_step_state: the reverse mode is Float64 only")
    length(U) == 5 * g.ncell || error("hip_step_state: U must hold 5 × ncell values")
    GC.@preserve U _check(ccall((:vh_grid_upload, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), g.ptr, U))
    _check(ccall((:vh_grid_step, libvespahip), Cint, (Ptr{Cvoid}, Float64), g.ptr, Float64(dt)))
    out = similar(U)
    GC.@preserve out _check(ccall((:vh_grid_download, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), g.ptr, out))
    return out
end

"""
    hip_rollout_state(r::HIPRollout, U₀, cfl) -> U_N

The whole checkpointed rollout as a function of its initial state: uploads U₀,
runs `r.nsteps` steps with the device CFL dt (checkpoints kept on the device) and
returns the final state. Its pullback is `hip_backward(r, Ū)` on the dt schedule
of this forward pass.
"""
function hip_rollout_state(r::HIPRollout, U₀::Vector{Float64}, cfl::Real)
    g = r.grid
    length(U₀) == 5 * g.ncell || error("hip_rollout_state: U₀ must hold 5 × ncell values")
    GC.@preserve U₀ _check(ccall((:vh_grid_upload, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), g.ptr, U₀))
    hip_forward!(r, cfl)
    out = similar(U₀)
    GC.@preserve out _check(ccall((:vh_grid_download, libvespahip), Cint, (Ptr{Cvoid}, Ptr{Cvoid}), g.ptr, out))
    return out
end

# The ChainRulesCore rules for hip_step_state and hip_rollout_state live in
# VespaHIPChainRulesExt.jl (a weak-dependency extension: loaded only together
# with ChainRulesCore).

export HIPFlux, HIPGrid, hip_step!, hip_evolve!, hip_set_mode!, hip_upload!, hip_download!, hip_conserved_totals,
       HIPBlocks, hip_download, hip_indicator, hip_remap!, hip_regrid!, hip_step_vjp, hip_enable_gravity!, hip_solve_poisson!,
       hip_phi, HIPMuscl, hip_muscl_step!, hip_fill_periodic!, hip_set_bc!, hip_max_wavespeed, hip_step_bflux!, hip_bflux_count, hip_set_parent_ghost!,
       hip_fluxrec, hip_flux_register, HIPRollout, hip_forward!, hip_backward, hip_step_state, hip_rollout_state

end # module
