/*
 * muscl_oracle.c — CPU oracle for PPMKernels' unsplit SSP-RK2 MUSCL step
 * (TEST INFRASTRUCTURE ONLY; included by vespa_oracle.c).
 *
 * Restates lib/PPMKernels/src/muscl.jl (PLM minmod-θ reconstruction, HLL flux
 * line) and lib/PPMKernels/src/muscl_grid.jl (muscl_step_3d! without the
 * dual-energy formalism, flux recording, fill_periodic!, max_wavespeed) on a
 * ghosted 3D grid, column-major (x fastest), conserved set (D, S1, S2, S3, τ).
 * The reference validates its flux line against the live Enzo solver (absent
 * here); this restatement is pinned in tests/ by known answers (minmod table,
 * exact linear reconstruction, HLL of equal and supersonic states against the
 * Euler flux), the reference's own driver checks (3D step on an axis-only state
 * equals a 1D RK2 built from the flux line, conservation, flux recording
 * reproduces the update) and Sod against the exact Riemann solution.
 */
#define MT double
#define MSUF f64
#define MSQRT sqrt
#define MFABS fabs
#include "muscl_impl.h"
#undef MT
#undef MSUF
#undef MSQRT
#undef MFABS

#define MT float
#define MSUF f32
#define MSQRT sqrtf
#define MFABS fabsf
#include "muscl_impl.h"
#undef MT
#undef MSUF
#undef MSQRT
#undef MFABS
