/*
 * vespahip.h — C ABI of the MI355X (gfx950) flux-divergence library.
 *
 * The library replaces the device flux backend of Vespa.jl behind its
 * MeshInterface seam. Plain pointers and sizes only; every entry point returns
 * 0 on success and a negative code on failure, with a message available from
 * vh_last_error() (the reference raises Julia exceptions instead:
 * src/ka_flux.jl:63-66, src/problem.jl:39-46).
 *
 * Conventions
 *   - Primitive W = (rho, vx, vy, vz, p); conserved U = (rho, rho*vx, rho*vy,
 *     rho*vz, E) (src/eos.jl:1-5).
 *   - Face batches are AoS: 5 contiguous values per face (the reinterpret of
 *     Vector{NTuple{5,T}} in ext/VespaKAFluxExt.jl:46-48,60).
 *   - Grid state is SoA: 5 fields, each a dense column-major array (x fastest),
 *     the RefMesh SoA layout (lib/RefMesh/src/RefMesh.jl:220-226), stored one
 *     after another: field f of cell c at U[f*ncell + c].
 *   - dtype: VH_F64 (Float64 state) or VH_F32 (Float32 state with Float64
 *     geometry, clock, CFL reduction, totals and pressure floor).
 *   - Boundary codes per (axis, side): VH_BC_PERIODIC, VH_BC_OUTFLOW,
 *     VH_BC_REFLECTING (lib/MeshInterface/src/MeshInterface.jl:73-89).
 *   - Calls on one context are stream-ordered on the context's own HIP stream;
 *     functions that return host data synchronise that stream first.
 */
#ifndef VESPAHIP_H
#define VESPAHIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { VH_F64 = 0, VH_F32 = 1 };
/* Boundary codes. VH_BC_PARENT is ParentGhost (lib/MeshInterface/src/MeshInterface.jl:86-99,
 * src/driver.jl:127-134), uniform grids only: the side's boundary ghost comes from the
 * caller (vh_grid_set_parent_ghost). */
enum { VH_BC_PERIODIC = 0, VH_BC_OUTFLOW = 1, VH_BC_REFLECTING = 2, VH_BC_PARENT = 3 };
/* Arithmetic profile of a grid context's stage kernels:
 *   VH_EXACT  the reference's evaluation order, no FMA contraction, correctly
 *             rounded division and square root: bit-identical to the reference
 *             arithmetic (the default);
 *   VH_FAST   FMA contraction, refined hardware reciprocals/rsqrt and fewer
 *             divisions in the HLLC star state: agrees with VH_EXACT to a few
 *             ulps per face. */
enum { VH_EXACT = 0, VH_FAST = 1 };

/* Last error message of the calling thread ("" if none). */
const char* vh_last_error(void);

/* ABI version of this header (bumped on any signature change). */
int vh_abi_version(void);

/* Number of visible HIP devices (0 when none; never an error). */
int vh_device_count(void);

/* Compute units of the current device (no counterpart in the reference: the
 * launch autotuner sizes its workgroup candidates and keys its cache by it). */
int vh_device_cu_count(int32_t* cus);

/* ---------------------------------------------------------------------------
 * Tier 1 — exact drop-in for compute_face_fluxes!(::KAFlux, model, F, WL, WR, axisv)
 * (ext/VespaKAFluxExt.jl:37-63; contract src/ka_flux.jl:40-53): HLLC flux
 * (src/riemann.jl:33-88) for nf faces. WL, WR, F: [nf][5]; axis: [nf] in {1,2,3}.
 * nf == 0 is a no-op (ext/VespaKAFluxExt.jl:40).
 *
 * Host-pointer form: copies in, runs on the device, copies back (the reference
 * kernel's own host-to-device crossing, ext/VespaKAFluxExt.jl:51-61), reusing
 * device staging buffers across calls. An axis value outside 1..3 is rejected
 * (-1) before anything is copied, as the reference's tuple index raises
 * (src/riemann.jl:36-37).
 */
int vh_face_flux_f64(const double* WL, const double* WR, const int32_t* axis, int64_t nf,
                     double gamma, double* F);
int vh_face_flux_f32(const float* WL, const float* WR, const int32_t* axis, int64_t nf,
                     double gamma, float* F);

/* Device-pointer form of the same contract, enqueued on `stream` (a
 * hipStream_t, NULL = the default stream). No synchronisation. The device
 * axis array is not validated (that would need a host copy): a value <= 1
 * computes an x face and a value >= 3 a z face. */
int vh_face_flux_dev_f64(const double* dWL, const double* dWR, const int32_t* daxis, int64_t nf,
                         double gamma, double* dF, void* stream);
int vh_face_flux_dev_f32(const float* dWL, const float* dWR, const int32_t* daxis, int64_t nf,
                         double gamma, float* dF, void* stream);

/* No counterpart in the reference. Allocation-free form of vh_face_flux_f64/_f32
 * for callers that cannot allocate device memory (tight serving loops, harnesses
 * that forbid hipMalloc): the same exact HLLC fluxes, with WL, WR, F and axis
 * staged in a fixed device array that the library holds on each GPU it runs on
 * until the process ends (VH_NOALLOC_MAX_FACES faces: 31 MiB). dtype VH_F64 or
 * VH_F32 selects the element type of WL, WR and F (host memory, [nf][5]).
 * nf > VH_NOALLOC_MAX_FACES is rejected (-1) before anything is copied, as is an
 * axis outside 1..3. Enqueued on `stream` (a hipStream_t, NULL = the default
 * stream); returns after F holds the fluxes. Calls no hipMalloc/hipFree. Calls
 * that overlap on one GPU share the staging array and race. */
#define VH_NOALLOC_MAX_FACES 262144
int vh_face_flux_noalloc(const void* WL, const void* WR, const int32_t* axis, int64_t nf,
                         double gamma, int32_t dtype, void* F, void* stream);

/* Device buffers for callers without a GPU array package of their own (the
 * reference's KA.allocate / copyto!, ext/VespaKAFluxExt.jl:51-54,59). */
int vh_buffer_alloc(void** dptr, int64_t bytes);
int vh_buffer_free(void* dptr);
int vh_buffer_upload(void* dptr, const void* host, int64_t bytes);
int vh_buffer_download(void* host, const void* dptr, int64_t bytes);
int vh_buffer_copy(void* ddst, const void* dsrc, int64_t bytes);

/* ---------------------------------------------------------------------------
 * Tier 2 — device-resident uniform grid (the RefMesh single-level path).
 * Replaces the whole-stage unit step!(sim, dt) (src/driver.jl:333-359) with its
 * accumulate_flux! (src/driver.jl:209-223), _euler_apply!/_rk2_combine!
 * (src/driver.jl:297-321), compute_dt (src/driver.jl:169-199), evolve!
 * (src/driver.jl:483-508) and conserved_totals (src/diagnostics.jl:16-31).
 */
typedef struct vh_grid vh_grid;

/* Create a context over a UniformMesh(dims, domain) (lib/RefMesh/src/RefMesh.jl:95-100)
 * of rank 1..3. dims[d], lo[d], hi[d] for d < rank; bc[2*d + side] (side 0 lo, 1 hi,
 * VH_BC_* codes). Periodic must be set on both sides of an axis, and a periodic axis
 * needs >= 2 cells. Allocates the state and the stage buffer on the current device. */
int vh_grid_create(vh_grid** out, int32_t rank, const int64_t* dims, const double* lo,
                   const double* hi, const int32_t* bc, double gamma, int32_t dtype);
int vh_grid_destroy(vh_grid* g);

/* Number of cells of the grid. */
int64_t vh_grid_ncell(const vh_grid* g);

/* Select / query the arithmetic profile (VH_EXACT or VH_FAST) of the stage
 * kernels of this context. The tier-1 face-flux entry points are always exact. */
int vh_grid_set_mode(vh_grid* g, int32_t mode);
int vh_grid_get_mode(const vh_grid* g);

/* Copy the conserved state (5*ncell values, layout above) from / to the host.
 * Replaces set_initial_conditions!/field views (src/driver.jl:101-112). */
int vh_grid_upload(vh_grid* g, const void* U);
int vh_grid_download(vh_grid* g, void* U);

/* Device address of the state (5*ncell values) for zero-copy callers. */
int vh_grid_state_device_ptr(vh_grid* g, void** dptr);

/* compute_dt(sim) (src/driver.jl:169-199): cfl / max(|v_d|+c)/w_d, +Inf if the
 * max is 0. Synchronises. */
int vh_grid_compute_dt(vh_grid* g, double cfl, double* dt);

/* step!(sim, dt) (src/driver.jl:333-359): one SSP-RK2 step with the given dt.
 * Stream-ordered; does not synchronise. */
int vh_grid_step(vh_grid* g, double dt);

/* step!(sim, dt; bflux) (src/driver.jl:333-359 with src/reflux.jl:105-132):
 * one step that also adds ½dt·F·area of every domain-boundary face, per RK
 * stage, to the device register dbflux (5 values of the state dtype per slot,
 * vh_grid_bflux_count slots: planes ordered axis 0 lo, axis 0 hi, axis 1 lo, ...,
 * each column-major over the other axes; periodic planes stay untouched). The
 * register accumulates; the caller zeroes it. Returns after the step has
 * completed, so the register can be read at once. */
int vh_grid_step_bflux(vh_grid* g, double dt, void* dbflux);

/* ParentGhost states (the closure ghost(axis, side, cell) of a ParentGhost side,
 * src/driver.jl:127-134): the primitive state (ρ, vx, vy, vz, p) at the state's
 * precision of every boundary face, 5 values per face in the boundary-flux
 * register's slot order (axis 0 lo plane, axis 0 hi, axis 1 lo, ...; each plane
 * column-major over the other axes; vh_grid_bflux_count faces in all, entries of
 * non-ParentGhost planes are not read). A ParentGhost side takes that state as
 * the outer neighbour of its boundary cells' slopes and as the outer state of its
 * boundary faces, in both RK stages of every step until the next call (the
 * reference's closure is evaluated per use; a caller whose parent data change
 * between steps uploads them before each step). Steps of a grid with a
 * ParentGhost side fail until this has been called; the reverse mode refuses
 * such grids. Synchronises. */
int vh_grid_set_parent_ghost(vh_grid* g, const void* host);
int64_t vh_grid_bflux_count(const vh_grid* g);

/* evolve!-style loop with the CFL dt computed on the device each step
 * (src/driver.jl:493-506). tfinal > 0: run while t < tfinal*(1-1e-12) with
 * dt = min(cfl/invdt, tfinal - t), synchronising once per step to test t;
 * max_steps > 0 bounds the count. tfinal <= 0: exactly max_steps steps with no
 * clip and no host synchronisation inside the loop. *t_io is the clock in/out,
 * *steps receives the number of steps taken. */
int vh_grid_run(vh_grid* g, double cfl, double tfinal, int64_t max_steps, double* t_io,
                int64_t* steps);

/* Self-gravity (src/gravity.jl). With gravity on, every RK stage adds the source
 * (0, -rho*g*V, -(rho*v.g)*V) of g = -grad(phi) to the flux sum
 * (apply_gravity_source!, src/gravity.jl:412-447), compute_dt folds in the
 * free-fall limiter sqrt(4*pi*G*rho)/0.5 (src/gravity.jl:449-452), and
 * vh_grid_run solves the Poisson equation before each step (src/driver.jl:498). */

/* enable_gravity!(sim; G, bcs, maxiter, tol, project_mean) (src/gravity.jl:63-80):
 * allocates phi (zero) and the CG work vectors. phi_periodic: bit d set when phi
 * is periodic along axis d (else homogeneous Neumann). Synchronises. */
int vh_grid_enable_gravity(vh_grid* g, double G, int32_t phi_periodic, int64_t maxiter, double tol,
                           int32_t project_mean);
int vh_grid_disable_gravity(vh_grid* g);

/* solve_poisson!(sim, grav) (src/gravity.jl:248-322): conjugate gradients on
 * A = -lap*V for the current density, warm-started from phi, Float64 dot
 * products, mean projection when enabled. Returns iterations and the final
 * relative residual; synchronises. */
int vh_grid_solve_poisson(vh_grid* g, int64_t* iters, double* relres);

/* phi (ncell values of the state dtype, x fastest) to / from host memory. */
int vh_grid_phi_upload(vh_grid* g, const void* host);
int vh_grid_phi_download(vh_grid* g, void* host);

/* apply_laplacian!(sim, grav, x, out) (src/gravity.jl:99-128): out = A*x on
 * device arrays of ncell values. Synchronises. */
int vh_grid_apply_laplacian(vh_grid* g, const void* dx, void* dout);

/* gravity_accel for every cell (src/gravity.jl:367-410): 3*ncell values (g_x,
 * g_y, g_z fields) into device memory. Synchronises. */
int vh_grid_gravity_accel(vh_grid* g, void* dg);

/* conserved_totals(sim) (src/diagnostics.jl:16-31): sum of U*V per field in
 * Float64 (tree order). Synchronises. */
int vh_grid_totals(vh_grid* g, double* out5);

/* Net flux out, acc = accumulate_flux!(sim) (src/driver.jl:209-223), for the
 * current state, copied to the host (5*ncell values of the state dtype).
 * Diagnostic/parity entry point; synchronises. */
int vh_grid_accumulate_flux(vh_grid* g, void* acc_host);

/* Wait for all work queued on the context's stream. */
int vh_grid_sync(vh_grid* g);

/* Launch shape of the fused stage kernels: tile variant (0 = library default)
 * and z planes per workgroup (0 = default). See vh_grid_tile_count for the
 * available tile variants of the grid's rank. Used by the autotuner. */
int vh_grid_set_launch(vh_grid* g, int32_t tile_variant, int32_t zchunk);
int vh_grid_tile_count(const vh_grid* g);
int vh_grid_get_launch(const vh_grid* g, int32_t* tile_x, int32_t* tile_y, int32_t* zchunk,
                       int64_t* workgroups);

/* Time the fused stage kernels with HIP events on the context's stream: runs
 * `nsteps` steps (device dt) and returns the mean duration in milliseconds of
 * each stage kernel and of the whole step. */
int vh_grid_time_steps(vh_grid* g, double cfl, int64_t nsteps, double* ms_stage1, double* ms_stage2,
                       double* ms_step);

/* ---------------------------------------------------------------------------
 * Reverse mode (Float64 uniform grids). The reference has no derivative of its
 * step (no AD anywhere in the tree); these entry points are what a
 * ChainRulesCore rrule over step!(sim, dt) (src/driver.jl:333-359) would call.
 * dt is held fixed; non-smooth points take the active branch (minmod's
 * selected argument or zero, the active Davis min/max argument, the upwind
 * HLLC branch, the pressure floor when it is the larger argument).
 */

/* dUbar = (∂ step(U, dt)/∂U)ᵀ dWbar for device arrays of 5*ncell doubles (the
 * grid's layout). dUbar may alias dWbar. Uses the grid's arithmetic profile for
 * the recomputed stage 1. Synchronises. */
int vh_grid_step_vjp(vh_grid* g, const void* dU, double dt, const void* dWbar, void* dUbar);

/* Kernel form of A'(U)ᵀλ inside the reverse mode: fused = 1 (default) uses one
 * tiled kernel per evaluation on 2D grids periodic on both axes with sizes that
 * are multiples of 32×8, fused = 0 always uses the three-pass form (any rank and
 * boundary). Both give the same values to round-off. */
int vh_grid_set_adjoint_form(vh_grid* g, int32_t fused);

/* Checkpointed rollout of nsteps steps with the device CFL dt: the forward
 * keeps the state every `every` steps and each step's dt; the backward
 * recomputes each segment from its checkpoint and applies the step VJPs in
 * reverse. */
typedef struct vh_adjoint vh_adjoint;
int vh_adjoint_create(vh_adjoint** out, vh_grid* g, int64_t nsteps, int32_t every);
int vh_adjoint_destroy(vh_adjoint* a);
/* nsteps steps from the grid's current state (left at the final state). */
int vh_adjoint_forward(vh_adjoint* a, double cfl);
/* dUbar (device) = gradient w.r.t. the initial state of a loss whose gradient
 * w.r.t. the final state is dWbar (device). */
int vh_adjoint_backward(vh_adjoint* a, const void* dWbar, void* dUbar);
/* Event times of the last forward/backward (ms), device bytes held by the
 * rollout and the grid's reverse-mode work buffers, and (optional) the dt of
 * every step (nsteps doubles). */
int vh_adjoint_info(const vh_adjoint* a, double* ms_forward, double* ms_backward, int64_t* device_bytes,
                    double* dts_host);

/* ---------------------------------------------------------------------------
 * Tier 2 — batched blocks (block-structured AMR with hanging nodes).
 * A mesh is a list of blocks: block b has a refinement level level[b] (0 =
 * base, at most maxlevel), cells per axis dims[3*b + d] and an integer origin
 * origin[3*b + d] on the lattice of the finest cells (one unit = one cell of
 * level maxlevel; a level-L cell spans 2^(maxlevel-L) units and h0[d]/2^L
 * physical units). Leaves are ordered block by block, column-major (x fastest)
 * inside a block; the state is SoA over that order, U[f*ncell + leaf]. bc[6*b +
 * 2*d + side] applies where no other block touches the face; periodic wraps
 * the block onto itself. Blocks that touch must differ by at most one level,
 * and each coarse face cell must see the 2^(rank-1) finer cells of one block.
 * Faces follow HGBackend (lib/HGBackend/src/HGBackend.jl:202-276):
 * conforming faces once from the hi side; coarse-fine faces once from the fine
 * side with the fine area; reconstruction across a block face uses one
 * representative neighbour (lib/HGBackend/src/HGBackend.jl:178-200): the
 * neighbour cell, the coarse cell containing the position, or the finer cell at
 * the lo corner of the coarse face. One persistent launch per RK stage covers
 * every block.
 */
typedef struct vh_blocks vh_blocks;

/* Validate a block mesh and count its derived tables without a device:
 * counts6 = {cells, tiles, link faces, link-cell records, flux-register cells,
 * face slots}. zchunk: z planes per tile (0 = whole block). */
int vh_blocks_describe(int32_t rank, int32_t nblocks, int32_t maxlevel, const double* h0,
                       const int32_t* level, const int32_t* dims, const int64_t* origin,
                       const int32_t* bc, int32_t zchunk, int64_t* counts6);

/* A 64-bit digest (FNV-1a over the bytes) of every table the builder derives
 * for a mesh (block records, representative and link-cell maps, link faces and
 * records, register and tile lists), built with nthreads host threads (0: the
 * builder's default, 1: serial). For checking that builds agree: the tables do
 * not depend on the thread count. No device needed. (No counterpart in the
 * reference: a check of this library's parallel builder.) */
int vh_blocks_table_digest(int32_t rank, int32_t nblocks, int32_t maxlevel, const double* h0, const int32_t* level,
                           const int32_t* dims, const int64_t* origin, const int32_t* bc, int32_t zchunk,
                           int32_t nthreads, uint64_t* digest);

/* Create a block-mesh context (the HGBackend Simulation, src/driver.jl:27-79
 * with lib/HGBackend/src/HGBackend.jl:300-340) on the current device. */
int vh_blocks_create(vh_blocks** out, int32_t rank, int32_t nblocks, int32_t maxlevel,
                     const double* h0, double gamma, const int32_t* level, const int32_t* dims,
                     const int64_t* origin, const int32_t* bc, int32_t dtype);
int vh_blocks_destroy(vh_blocks* m);
int64_t vh_blocks_ncell(const vh_blocks* m);
int vh_blocks_set_mode(vh_blocks* m, int32_t mode);
int vh_blocks_upload(vh_blocks* m, const void* U);
int vh_blocks_download(vh_blocks* m, void* U);

/* compute_dt(sim; level) (src/driver.jl:169-199); level < 0: every leaf. */
int vh_blocks_compute_dt(vh_blocks* m, double cfl, int32_t level, double* dt);

/* step!(sim, dt; level) (src/driver.jl:333-378): level < 0 updates every leaf;
 * level >= 0 updates that level's leaves only (others are read). No flux
 * registers. Stream-ordered. */
int vh_blocks_step(vh_blocks* m, double dt, int32_t level);

/* evolve! (src/driver.jl:483-508): subcycle = 0, a global dt each step from the
 * device CFL reduction (tfinal <= 0: exactly max_steps steps, no host
 * synchronisation inside the loop); subcycle = 1, evolve_level! recursion with
 * per-level dt and flux registers (src/driver.jl:553-605, src/reflux.jl:51-103),
 * which needs tfinal > 0. *t_io in/out, *steps = root steps taken. */
int vh_blocks_run(vh_blocks* m, double cfl, double tfinal, int64_t max_steps, int32_t subcycle,
                  double* t_io, int64_t* steps);

/* conserved_totals(sim) (src/diagnostics.jl:16-31): sum of U*V in Float64. */
int vh_blocks_totals(vh_blocks* m, double* out5);

/* Launch shape: z planes per tile (0 = whole block), cap on the persistent
 * workgroups (0 = 3 per compute unit), and whether the single-rate step is
 * replayed from a captured HIP graph (1, the default) or enqueued eagerly. */
int vh_blocks_set_launch(vh_blocks* m, int32_t zchunk, int32_t max_workgroups, int32_t use_graph);
/* Widest tile along x (4..32, default 32); tiles take min(nx, max_tile_x)
 * columns and as many rows as fit 256 lanes. */
int vh_blocks_set_tiling(vh_blocks* m, int32_t max_tile_x);
int vh_blocks_info(const vh_blocks* m, int64_t* counts6, int32_t* zchunk, int32_t* workgroups);
int vh_blocks_sync(vh_blocks* m);

/* Regrid support (regrid!, src/driver.jl:440-462).
 *
 * vh_blocks_indicator: density_gradient_indicator (src/driver.jl:400-414) of
 * every leaf on the device, the largest |ρn − ρc| / max(ρc, ρn, eps) over the
 * representative neighbours across the 2·rank faces that are interior (domain
 * boundaries are skipped, periodic sides wrap inside the block; the difference
 * at the state's precision, the ratio in Float64). block_max (host, nblocks
 * doubles) receives the per-block maximum; cell_g (host, ncell doubles, may be
 * NULL) the per-leaf values. Returns after the values are on the host. */
int vh_blocks_indicator(vh_blocks* m, double* block_max, double* cell_g);

/* vh_blocks_remap: conservative transfer of src's state onto dst's mesh, device
 * to device (the remap refine! triggers, prolong! by injection,
 * lib/RefMesh/src/RefMesh.jl:284-302). src_block[b] names the block of src that
 * contains block b of dst, at the same or a coarser level; every cell of b takes
 * the conserved values of the src cell containing it, so Σ U·V is unchanged.
 * Both contexts need the same dtype, device, rank, maxlevel and h0. dst's state
 * is replaced; returns after the copy has completed. */
int vh_blocks_remap(vh_blocks* dst, const vh_blocks* src, const int32_t* src_block);

/* vh_blocks_regrid: regrid! in place (src/driver.jl:440-462 with the remap of
 * refine!, lib/RefMesh/src/RefMesh.jl:284-302): the context takes the new block
 * table (nblocks, level, dims, origin, bc as for vh_blocks_create; same rank,
 * maxlevel, h0 and gamma) and its state is carried over by the injection of
 * vh_blocks_remap (src_block[b] = the block of the current mesh containing new
 * block b). The stream, dtype, arithmetic profile, launch shape and clock stay;
 * the tables are built by the host builder's threads and uploaded in one copy
 * into a second device arena, and the state buffers are reused when large enough
 * (allocated at twice the need, at least 4 MiB, otherwise; the arenas and the
 * pinned staging likewise, at least 8 and 4 MiB), so a regrid allocates only
 * when the mesh outgrows twice an earlier one. With self-gravity on, the operator is
 * rebuilt for the new mesh and φ is carried over by the same injection (the next
 * solve's warm start). Invalid tables leave the context unchanged; a HIP error
 * after that leaves it fit only for vh_blocks_destroy. The upload and the remap
 * are queued on the context's stream behind the work already there and the call
 * returns without waiting for them (it waits when the state buffers grow or
 * self-gravity is on): later calls on the context are ordered after them, and
 * vh_blocks_sync waits for them. */
int vh_blocks_regrid(vh_blocks* m, int32_t nblocks, const int32_t* level, const int32_t* dims,
                     const int64_t* origin, const int32_t* bc, const int32_t* src_block);

/* Self-gravity on a block mesh (src/gravity.jl; the composite operator of
 * _build_poisson_ka_struct, src/gravity.jl:455-600, and its device CG,
 * ext/VespaKAGravityExt.jl:21-115). The operator A = −∇²·V is built on the host
 * from the block face walk as CSR rows (one edge per face, coef = area/d with
 * d = ½(w_left + w_right) along the axis; coarse↔fine faces once from the fine
 * side with the fine area). φ's boundaries are the mesh's own: periodic block
 * sides wrap, other uncovered sides are homogeneous Neumann. With gravity on,
 * every RK stage adds the source (0, −ρg·V, −(ρv·g)·V) of g = −∇φ (held over
 * the step) to the flux sum, the CFL dt folds in the free-fall limiter
 * √(4πGρ)/½ (over the leaves of the level asked for), and vh_blocks_run solves
 * the Poisson equation before each root step, g held over every substep when
 * subcycling (src/driver.jl:498,589-592); vh_blocks_step uses the current φ.
 *
 * enable_gravity!(sim; G, maxiter, tol, project_mean) (src/gravity.jl:63-80):
 * builds the operator, allocates φ (zero) and the CG vectors; rebuilds on a
 * second call. Synchronises. */
int vh_blocks_enable_gravity(vh_blocks* m, double G, int64_t maxiter, double tol, int32_t project_mean);
int vh_blocks_disable_gravity(vh_blocks* m);
/* solve_poisson! (src/gravity.jl:248-322): CG warm-started from φ, Float64
 * dots, mean projection when enabled. Synchronises. */
int vh_blocks_solve_poisson(vh_blocks* m, int64_t* iters, double* relres);
/* φ (ncell values of the state dtype, leaf order) to / from host memory. */
int vh_blocks_phi_upload(vh_blocks* m, const void* host);
int vh_blocks_phi_download(vh_blocks* m, void* host);
/* out = A·x on device arrays of ncell values (Float64 row sums in CSR order,
 * _lap_csr_cpu!, src/gravity.jl:326-336). Synchronises. */
int vh_blocks_apply_laplacian(vh_blocks* m, const void* dx, void* dout);
/* g = −∇φ of the current φ for every leaf (_gax, src/gravity.jl:403-410): 3·ncell
 * values (g_x, g_y, g_z fields) into device memory. Synchronises. */
int vh_blocks_gravity_accel(vh_blocks* m, void* dg);

/* Time nsteps single-rate steps (device dt) with HIP events after two untimed
 * steps. method 0: one persistent launch per stage, eager; 1: one launch per
 * block per stage (the comparison loop); 2: persistent launches replayed from a
 * HIP graph. ms_stage2 (optional): mean duration of the persistent stage-2
 * launch over up to 20 further eager steps. */
int vh_blocks_time_steps(vh_blocks* m, double cfl, int64_t nsteps, int32_t method, double* ms_step,
                         double* ms_stage2);

/*
 * PPMKernels' unsplit SSP-RK2 MUSCL step (muscl_step_3d!, lib/PPMKernels/src/muscl_grid.jl:314-370:
 * PLM minmod-θ reconstruction + HLL flux, lib/PPMKernels/src/muscl.jl:13-58,181-206) on a
 * ghosted 3D grid: dims = cells per axis including ng ghost zones on each side (the
 * active region is ng .. dims-ng-1 per axis), column-major (x fastest), state SoA
 * U[f*ncell + cell] with f = D, S1, S2, S3, τ (τ = ρ·etot). Ghost zones are not
 * refilled by a step unless a bc! is set (vh_muscl_set_bc); vh_muscl_fill_periodic
 * is fill_periodic!. Arithmetic at the state's precision, bit-identical to the reference's
 * operation order. The dual-energy field is optional (vh_muscl_enable_dual_energy).
 */
typedef struct vh_muscl vh_muscl;

/* gamma, theta (limiter, reference default 1.5), small_rho (density and eint floor,
 * default 1e-10) and dx are fixed per context; ng >= 2, dims[d] >= 2*ng+1. */
int vh_muscl_create(vh_muscl** out, const int32_t* dims, int32_t ng, double gamma, double theta,
                    double small_rho, double dx, int32_t dtype);
int vh_muscl_destroy(vh_muscl* m);
int64_t vh_muscl_ncell(const vh_muscl* m);
/* The dual-energy formalism (muscl_step_3d!(…; ge, eta1)): the state gains the gas
 * energy ρe as a sixth field (initialised to 0; upload the six fields after this call),
 * eint is selected by _dual_eint, ρe is advected with the HLL flux ρe·u and synced with
 * τ at the end of each step (lib/PPMKernels/src/muscl_grid.jl:34-115); recorded
 * fluxes then hold 6 fields per axis. Idempotent; eta1 may be changed later. The
 * first call moves the state to a new device allocation, so it is rejected (-1,
 * nothing changed) once vh_muscl_state_device_ptr has handed out the address:
 * enable dual energy first, then take the pointer. */
int vh_muscl_enable_dual_energy(vh_muscl* m, double eta1);
/* 5 (or 6) fields × ncell values, field-major. */
int vh_muscl_upload(vh_muscl* m, const void* U);
int vh_muscl_download(vh_muscl* m, void* U);
int vh_muscl_state_device_ptr(vh_muscl* m, void** ptr);
/* Arithmetic profile of the stage kernels: VH_EXACT (default; the reference's operations
 * and order, correctly rounded divisions and square roots: bit-identical) or VH_FAST
 * (contracted a*b+c, one refined reciprocal per cell and per face, refined rsqrt sound
 * speeds: within a few ulps per face). */
int vh_muscl_set_mode(vh_muscl* m, int32_t mode);
/* Tiles: variant 0 = 16x16 cells, 1 = 32x8; zchunk = z planes per workgroup. */
int vh_muscl_set_tiles(vh_muscl* m, int32_t variant, int32_t zchunk);
/* One muscl_step_3d!(…; dt): two launches, stream-ordered. record != 0 is
 * fluxrec: the recorded fluxes (zeroed, then ½F per stage at the lo face of each
 * cell, lib/PPMKernels/src/muscl_grid.jl:278-291) are read with
 * vh_muscl_fluxrec_download, 3·nfields·ncell values [axis][field][cell]. */
int vh_muscl_step(vh_muscl* m, double dt, int32_t record);
int vh_muscl_fluxrec_download(vh_muscl* m, void* frec);
/* boundary_flux_register (lib/PPMKernels/src/reflux_capture.jl:86-131) of the last
 * recording step: face fluxes as ∫F·area·dt = F·dt·dx² in Float64. act = dims − 2·ng;
 * for each axis a (x, y, z) in turn, with b1 < b2 the other two axes and u1, u2 their
 * active cells: planes receives the lo boundary plane (the lo face of the first active
 * cell) then the hi plane (the hi face of the last), each [field][u2][u1] (10·act_b1·act_b2
 * values per axis); interior (may be NULL) the hi faces of active cells 1 .. act_a−1,
 * [field][u2][u1][c] (5·(act_a−1)·act_b1·act_b2 values per axis). */
int vh_muscl_flux_register(vh_muscl* m, double dt, double* planes, double* interior);
/* fill_periodic! (lib/PPMKernels/src/muscl_grid.jl:644-670) on the state's fields. */
int vh_muscl_fill_periodic(vh_muscl* m);

/* bc! of muscl_step_3d! (lib/PPMKernels/src/muscl_grid.jl:316-322, 329, 339, 350):
 * refill the ghost zones before each RK stage. kind VH_MUSCL_BC_NONE is the
 * reference's bc! = nothing (the default); VH_MUSCL_BC_PERIODIC runs
 * fill_periodic! on the device before each stage, with no host round trip;
 * VH_MUSCL_BC_CALLBACK calls fn(user, state, stream, stage) on the host before
 * stage 0 and stage 1. `state` is device memory holding the stage's input
 * (nfields·ncell values, the layout of the state); fn fills its ghost zones in
 * order on `stream` (a hipStream_t: enqueue a kernel there, or synchronise it
 * and copy) and returns 0; any other value abandons the step with an error and
 * leaves the state as it was. As in the reference, the step's combination uses
 * the input as it was before the first refill (D0 is copied before bcfill!): a
 * step with bc! copies the input once to a second buffer, refills it there and
 * refills the stage-1 state in place. */
typedef int (*vh_muscl_bc_fn)(void* user, void* state, void* stream, int32_t stage);
enum { VH_MUSCL_BC_NONE = 0, VH_MUSCL_BC_PERIODIC = 1, VH_MUSCL_BC_CALLBACK = 2 };
int vh_muscl_set_bc(vh_muscl* m, int32_t kind, vh_muscl_bc_fn fn, void* user);
/* max_wavespeed (lib/PPMKernels/src/muscl_grid.jl:672-695) over every cell. */
int vh_muscl_max_wavespeed(vh_muscl* m, double* out);
int vh_muscl_sync(vh_muscl* m);
/* Time nsteps steps with HIP events after one untimed step; ms_stage2: mean of up
 * to 20 stage-2 launches timed one by one. */
int vh_muscl_time_steps(vh_muscl* m, double dt, int64_t nsteps, int32_t record, double* ms_step,
                        double* ms_stage2);

#ifdef __cplusplus
}
#endif

#endif /* VESPAHIP_H */
