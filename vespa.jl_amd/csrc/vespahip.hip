// vespahip — C-ABI entry points of the gfx950 flux-divergence library.
//
// Kernels (per arithmetic profile, built from stages.hip):
//   face_flux_kernel   one lane per face, HLLC (tier-1 drop-in for the
//                      KernelAbstractions kernel, ext/VespaKAFluxExt.jl:23-35)
//   rk_stage_kernel    fused RK stage over a uniform grid (stencil.h)
//   cfl_kernel         per-workgroup max of the CFL term (src/driver.jl:178-185)
// Kernels here (profile-independent):
//   dt_finalize_kernel reduce the partial maxima into dt on the device
//   totals_kernel      per-workgroup Σ U·V partial sums (src/diagnostics.jl:16-31)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>

#include "../../include/vespahip.h"
#include "host.h"
#include "profiles.h"

namespace vh {

// ----------------------------------------------------------------------------
// error plumbing
static thread_local std::string g_err;

int fail(int code, const char* fmt, const char* a, long long b) {
  char buf[512];
  snprintf(buf, sizeof buf, fmt, a, b);
  g_err = buf;
  return code;
}

// ----------------------------------------------------------------------------
// reductions

// clock: [0] t, [1] dt, [2] invdt, [3] steps (as double)
__global__ void __launch_bounds__(256) dt_finalize_kernel(const double* __restrict__ partials, int np,
                                                          double cfl, double tfinal, int clip,
                                                          double* clock) {
  __shared__ double red[16];
  double m = 0.0;
  for (int i = threadIdx.x; i < np; i += 256) m = vmax(m, partials[i]);
  m = block_max<256>(m, red);
  if (threadIdx.x == 0) {
    double dt = m == 0.0 ? HUGE_VAL : cfl / m;
    if (clip) dt = vmin(dt, tfinal - clock[0]);
    clock[1] = dt;
    clock[2] = m;
  }
}

__global__ void set_dt_kernel(double* clock, double dt) { clock[1] = dt; }

__global__ void advance_clock_kernel(double* clock) {
  clock[0] = clock[0] + clock[1];
  clock[3] = clock[3] + 1.0;
}

template <typename T>
__global__ void __launch_bounds__(256) totals_kernel(const T* __restrict__ U, long long n, double V,
                                                     double* __restrict__ partials) {
  __shared__ double red[16];
  double s[5] = {0, 0, 0, 0, 0};
  for (long long c = (long long)blockIdx.x * 256 + threadIdx.x; c < n; c += (long long)gridDim.x * 256) {
#pragma unroll
    for (int f = 0; f < 5; ++f) s[f] += double(U[f * n + c]) * V;
  }
#pragma unroll
  for (int f = 0; f < 5; ++f) {
    const double t = block_sum<256>(s[f], red);
    if (threadIdx.x == 0) partials[5 * blockIdx.x + f] = t;
  }
}

// acc = (U − U1)/invV reconstructed exactly is not possible; the diagnostic
// accumulate path instead runs stage 1 with dt = −V (invV = −1 in Tf), so
// U1 = U + acc·1, then subtracts U in a second kernel. With invV = −1 the
// multiply is exact, but the add/subtract round. The parity tests therefore use
// the step path; this entry point is kept for inspection only.
template <typename T>
__global__ void sub_kernel(const T* __restrict__ a, const T* __restrict__ b, T* __restrict__ out,
                           long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a[i] - b[i];
}

void enqueue_dt_finalize(hipStream_t s, const double* partials, int np, double cfl, double tfinal, int clip,
                         double* clock) {
  hipLaunchKernelGGL(dt_finalize_kernel, dim3(1), dim3(256), 0, s, partials, np, cfl, tfinal, clip, clock);
}
void enqueue_set_dt(hipStream_t s, double* clock, double dt) {
  hipLaunchKernelGGL(set_dt_kernel, dim3(1), dim3(1), 0, s, clock, dt);
}
void enqueue_advance_clock(hipStream_t s, double* clock) {
  hipLaunchKernelGGL(advance_clock_kernel, dim3(1), dim3(1), 0, s, clock);
}

}  // namespace vh

using namespace vh;

// The C-ABI profile codes must survive every header of this translation unit.
static_assert(VH_EXACT == 0 && VH_FAST == 1, "profile codes of vespahip.h were redefined");

// ----------------------------------------------------------------------------
// context

struct vh_grid {
  int dtype;
  int device;
  int mode;        // VH_EXACT or VH_FAST
  GridDev G;
  int variant;
  int tx, ty;
  int stage_blocks;
  int cfl_blocks;
  size_t esize;
  void* U;
  void* U1;
  double* partials;
  int npartials;
  double* clock;
  double* totals_partials;
  hipStream_t stream;
  struct AdjWork* adj;   // reverse-mode work buffers (allocated on first use)
  struct GravWork* grav; // self-gravity state (vh_grid_enable_gravity)
  void* pg;              // ParentGhost states (vh_grid_set_parent_ghost), G.pg points here
};

// A step with a ParentGhost side needs its states first.
static int need_parent_states(const vh_grid* g) {
  for (int s = 0; s < 2 * g->G.rank; ++s)
    if (g->G.bc[s] == BC_PARENT && !g->G.pg)
      return fail(-1, "a ParentGhost side needs its states (vh_grid_set_parent_ghost)%s", "");
  return 0;
}
static bool has_parent(const vh_grid* g) {
  for (int s = 0; s < 2 * g->G.rank; ++s)
    if (g->G.bc[s] == BC_PARENT) return true;
  return false;
}

// Self-gravity of a uniform-grid context (the reference's GravityField,
// src/gravity.jl:30-45): φ, CG work vectors and solver settings.
struct GravWork {
  GravParams P;
  GravBuffers B;
};

static void free_grav(vh_grid* g) {
  GravWork* w = g->grav;
  if (!w) return;
  if (w->B.phi) (void)hipFree(w->B.phi);
  if (w->B.work) (void)hipFree(w->B.work);
  if (w->B.state) (void)hipFree(w->B.state);
  if (w->B.scal) (void)hipFree(w->B.scal);
  if (w->B.partials) (void)hipFree(w->B.partials);
  if (w->B.host_state) (void)hipHostFree(w->B.host_state);
  delete w;
  g->grav = nullptr;
  g->G.phi = nullptr;
}

static const int kMaxCflBlocks = 2048;

static int configure(vh_grid* g, int variant, int zchunk) {
  GridDev& G = g->G;
  if (variant < 0 || variant >= tile_count_for_rank(G.rank))
    return fail(-1, "tile variant %s%lld out of range", "", variant);
  const TileShape t = tile_for(G.rank, variant);
  g->variant = variant;
  g->tx = t.tx;
  g->ty = t.ty;
  G.tiles_x = (G.n[0] + t.tx - 1) / t.tx;
  G.tiles_y = (G.n[1] + t.ty - 1) / t.ty;
  if (G.rank == 3) {
    if (zchunk <= 0) zchunk = 32;
    if (zchunk > G.n[2]) zchunk = G.n[2];
  } else {
    zchunk = 1;
  }
  G.zchunk = zchunk;
  G.chunks_z = (G.n[2] + zchunk - 1) / zchunk;
  G.xcd_remap = 1;
  const long long blocks = (long long)G.tiles_x * G.tiles_y * G.chunks_z;
  if (blocks > (1LL << 30)) return fail(-1, "grid too large for one launch%s (%lld workgroups)", "", blocks);
  g->stage_blocks = (int)blocks;
  const int need = g->stage_blocks > kMaxCflBlocks ? g->stage_blocks : kMaxCflBlocks;
  if (need > g->npartials) {
    if (g->partials) VH_HIP(hipFree(g->partials));
    g->partials = nullptr;
    VH_HIP(hipMalloc(&g->partials, sizeof(double) * need));
    g->npartials = need;
  }
  return 0;
}

static void enqueue_stage_dt(vh_grid* g, bool stage2, const void* Us, const void* U0, void* Uout,
                             const double* dtp) {
  if (g->mode == VH_FAST)
    fast::launch_stage(g->dtype, g->G.rank, g->variant, stage2, g->stage_blocks, g->stream, Us, U0, Uout, dtp,
                       g->partials, g->G);
  else
    exact::launch_stage(g->dtype, g->G.rank, g->variant, stage2, g->stage_blocks, g->stream, Us, U0, Uout, dtp,
                        g->partials, g->G);
}

static void enqueue_stage(vh_grid* g, bool stage2, const void* Us, const void* U0, void* Uout) {
  const double* dtp = g->clock + 1;
  if (g->mode == VH_FAST)
    fast::launch_stage(g->dtype, g->G.rank, g->variant, stage2, g->stage_blocks, g->stream, Us, U0, Uout, dtp,
                       g->partials, g->G);
  else
    exact::launch_stage(g->dtype, g->G.rank, g->variant, stage2, g->stage_blocks, g->stream, Us, U0, Uout, dtp,
                        g->partials, g->G);
}

static void enqueue_cfl(vh_grid* g) {
  if (g->mode == VH_FAST)
    fast::launch_cfl(g->dtype, g->G.rank, g->cfl_blocks, g->stream, g->U, g->G, g->partials);
  else
    exact::launch_cfl(g->dtype, g->G.rank, g->cfl_blocks, g->stream, g->U, g->G, g->partials);
}

// One RK2 step reading dt from clock[1]; stage 2 leaves the partial CFL maxima
// of the new state in g->partials (g->stage_blocks entries). Optional events
// bracket the two stage kernels.
static void enqueue_step(vh_grid* g, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr, hipEvent_t e2 = nullptr) {
  if (e0) (void)hipEventRecord(e0, g->stream);
  enqueue_stage(g, false, g->U, g->U, g->U1);
  if (e1) (void)hipEventRecord(e1, g->stream);
  enqueue_stage(g, true, g->U1, g->U, g->U);
  if (e2) (void)hipEventRecord(e2, g->stream);
  hipLaunchKernelGGL(advance_clock_kernel, dim3(1), dim3(1), 0, g->stream, g->clock);
}

extern "C" {

const char* vh_last_error(void) { return g_err.c_str(); }
int vh_abi_version(void) { return 1; }

int vh_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

int vh_device_cu_count(int32_t* cus) {
  if (!cus) return fail(-1, "null argument%s", "");
  int dev = 0, n = 0;
  VH_HIP(hipGetDevice(&dev));
  VH_HIP(hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev));
  *cus = n;
  return 0;
}

// ---- tier 1 ----

int vh_face_flux_dev_f64(const double* dWL, const double* dWR, const int32_t* daxis, int64_t nf, double gamma,
                         double* dF, void* stream) {
  if (nf < 0) return fail(-1, "nf must be >= 0%s (got %lld)", "", nf);
  if (nf == 0) return 0;
  if (!dWL || !dWR || !daxis || !dF) return fail(-1, "null pointer%s", "");
  if ((nf + 255) / 256 > 0x7fffffffLL) return fail(-1, "too many faces%s (%lld)", "", nf);
  exact::launch_face_flux(0, nf, (hipStream_t)stream, dWL, dWR, daxis, gamma, dF);
  VH_HIP(hipGetLastError());
  return 0;
}

int vh_face_flux_dev_f32(const float* dWL, const float* dWR, const int32_t* daxis, int64_t nf, double gamma,
                         float* dF, void* stream) {
  if (nf < 0) return fail(-1, "nf must be >= 0%s (got %lld)", "", nf);
  if (nf == 0) return 0;
  if (!dWL || !dWR || !daxis || !dF) return fail(-1, "null pointer%s", "");
  if ((nf + 255) / 256 > 0x7fffffffLL) return fail(-1, "too many faces%s (%lld)", "", nf);
  exact::launch_face_flux(1, nf, (hipStream_t)stream, dWL, dWR, daxis, gamma, dF);
  VH_HIP(hipGetLastError());
  return 0;
}

}  // extern "C"

// Host-pointer form: persistent device staging per thread (grown on demand).
struct Staging {
  void* buf = nullptr;
  size_t bytes = 0;
  int device = -1;
  hipStream_t stream = nullptr;
};
static thread_local Staging g_stage;

template <typename T>
static int face_flux_host(const T* WL, const T* WR, const int32_t* ax, int64_t nf, double gamma, T* F) {
  if (nf < 0) return fail(-1, "nf must be >= 0%s (got %lld)", "", nf);
  if (nf == 0) return 0;
  if (!WL || !WR || !ax || !F) return fail(-1, "null pointer%s", "");
  for (int64_t i = 0; i < nf; ++i)
    if (ax[i] < 1 || ax[i] > 3) return fail(-1, "axis must be 1, 2 or 3%s (face %lld)", "", i);
  int dev = 0;
  VH_HIP(hipGetDevice(&dev));
  const size_t nv = (size_t)nf * 5 * sizeof(T);
  const size_t na = ((size_t)nf * sizeof(int32_t) + 255) & ~(size_t)255;
  const size_t nv_al = (nv + 255) & ~(size_t)255;
  const size_t need = 3 * nv_al + na;
  if (g_stage.device != dev || g_stage.bytes < need) {
    if (g_stage.buf && g_stage.device == dev) VH_HIP(hipFree(g_stage.buf));
    g_stage.buf = nullptr;
    VH_HIP(hipMalloc(&g_stage.buf, need));
    g_stage.bytes = need;
    if (g_stage.device != dev || !g_stage.stream) VH_HIP(hipStreamCreateWithFlags(&g_stage.stream, hipStreamNonBlocking));
    g_stage.device = dev;
  }
  char* base = (char*)g_stage.buf;
  T* dWL = (T*)base;
  T* dWR = (T*)(base + nv_al);
  T* dF = (T*)(base + 2 * nv_al);
  int32_t* dax = (int32_t*)(base + 3 * nv_al);
  hipStream_t s = g_stage.stream;
  VH_HIP(hipMemcpyAsync(dWL, WL, nv, hipMemcpyHostToDevice, s));
  VH_HIP(hipMemcpyAsync(dWR, WR, nv, hipMemcpyHostToDevice, s));
  VH_HIP(hipMemcpyAsync(dax, ax, (size_t)nf * sizeof(int32_t), hipMemcpyHostToDevice, s));
  exact::launch_face_flux(sizeof(T) == 8 ? 0 : 1, nf, s, dWL, dWR, dax, gamma, dF);
  VH_HIP(hipGetLastError());
  VH_HIP(hipMemcpyAsync(F, dF, nv, hipMemcpyDeviceToHost, s));
  VH_HIP(hipStreamSynchronize(s));
  return 0;
}

extern "C" {

int vh_face_flux_f64(const double* WL, const double* WR, const int32_t* axis, int64_t nf, double gamma,
                     double* F) {
  return face_flux_host<double>(WL, WR, axis, nf, gamma, F);
}
int vh_face_flux_f32(const float* WL, const float* WR, const int32_t* axis, int64_t nf, double gamma,
                     float* F) {
  return face_flux_host<float>(WL, WR, axis, nf, gamma, F);
}

}  // extern "C"

// ---- allocation-free host-pointer form ----
// vh_face_flux_noalloc stages WL, WR, F and the axes in `scratch`, a fixed array
// in device memory that every GPU running the library holds until the process
// ends. In host code `scratch` names a host-side stand-in, never the device
// memory: its device address is fetched once per device with
// hipGetSymbolAddress (with that device current) and only that address is
// passed to copies and kernels. Calls that overlap on one GPU share the array
// and race; each call leaves its staging behind (nothing needs to start zero).
static constexpr int64_t kNoAllocMaxFaces = VH_NOALLOC_MAX_FACES;
static constexpr size_t kNoAllocValBytes = (size_t)kNoAllocMaxFaces * 5 * sizeof(double);  // one [nf][5] array
static constexpr size_t kNoAllocBytes = 3 * kNoAllocValBytes + (size_t)kNoAllocMaxFaces * sizeof(int32_t);
static_assert(kNoAllocValBytes % 256 == 0, "staging arrays stay 256-byte aligned");
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

static constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_dev[kMaxDevices];   // device address of `scratch` per device

static int scratch_address(int dev, unsigned char** out) {
  if (dev < 0 || dev >= kMaxDevices) return fail(-1, "device index%s %lld is beyond the supported range", "", dev);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
      return fail(-2, "hipGetSymbolAddress(scratch) failed: %s (code %lld)", hipGetErrorString(e), (long long)e);
    g_scratch_dev[dev] = (unsigned char*)p;
  }
  *out = g_scratch_dev[dev];
  return 0;
}

extern "C" {

int vh_face_flux_noalloc(const void* WL, const void* WR, const int32_t* axis, int64_t nf, double gamma,
                         int32_t dtype, void* F, void* stream) {
  if (nf < 0) return fail(-1, "nf must be >= 0%s (got %lld)", "", nf);
  if (dtype != VH_F64 && dtype != VH_F32) return fail(-1, "dtype must be VH_F64 or VH_F32%s (got %lld)", "", dtype);
  if (nf > kNoAllocMaxFaces)
    return fail(-1, "vh_face_flux_noalloc: nf = %s%lld exceeds the limit of VH_NOALLOC_MAX_FACES (262144) faces "
                "per call; split the batch or use vh_face_flux_f64/_f32", "", nf);
  if (nf == 0) return 0;
  if (!WL || !WR || !axis || !F) return fail(-1, "null pointer%s", "");
  for (int64_t i = 0; i < nf; ++i)
    if (axis[i] < 1 || axis[i] > 3) return fail(-1, "axis must be 1, 2 or 3%s (face %lld)", "", i);
  int dev = 0;
  VH_HIP(hipGetDevice(&dev));
  unsigned char* base = nullptr;
  if (int rc = scratch_address(dev, &base)) return rc;
  const size_t nv = (size_t)nf * 5 * (dtype == VH_F64 ? sizeof(double) : sizeof(float));
  unsigned char* dWL = base;
  unsigned char* dWR = base + kNoAllocValBytes;
  unsigned char* dF = base + 2 * kNoAllocValBytes;
  int32_t* dax = (int32_t*)(base + 3 * kNoAllocValBytes);
  hipStream_t s = (hipStream_t)stream;
  VH_HIP(hipMemcpyAsync(dWL, WL, nv, hipMemcpyHostToDevice, s));
  VH_HIP(hipMemcpyAsync(dWR, WR, nv, hipMemcpyHostToDevice, s));
  VH_HIP(hipMemcpyAsync(dax, axis, (size_t)nf * sizeof(int32_t), hipMemcpyHostToDevice, s));
  exact::launch_face_flux(dtype == VH_F64 ? 0 : 1, nf, s, dWL, dWR, dax, gamma, dF);
  VH_HIP(hipGetLastError());
  VH_HIP(hipMemcpyAsync(F, dF, nv, hipMemcpyDeviceToHost, s));
  VH_HIP(hipStreamSynchronize(s));
  return 0;
}

int vh_buffer_alloc(void** dptr, int64_t bytes) {
  if (!dptr || bytes < 0) return fail(-1, "bad argument%s", "");
  *dptr = nullptr;
  if (bytes == 0) return 0;
  VH_HIP(hipMalloc(dptr, (size_t)bytes));
  return 0;
}
int vh_buffer_free(void* dptr) {
  if (dptr) VH_HIP(hipFree(dptr));
  return 0;
}
int vh_buffer_upload(void* dptr, const void* host, int64_t bytes) {
  if (bytes < 0 || (bytes > 0 && (!dptr || !host))) return fail(-1, "bad argument%s", "");
  if (bytes) VH_HIP(hipMemcpy(dptr, host, (size_t)bytes, hipMemcpyHostToDevice));
  return 0;
}
int vh_buffer_download(void* host, const void* dptr, int64_t bytes) {
  if (bytes < 0 || (bytes > 0 && (!dptr || !host))) return fail(-1, "bad argument%s", "");
  if (bytes) VH_HIP(hipMemcpy(host, dptr, (size_t)bytes, hipMemcpyDeviceToHost));
  return 0;
}
int vh_buffer_copy(void* ddst, const void* dsrc, int64_t bytes) {
  if (bytes < 0 || (bytes > 0 && (!ddst || !dsrc))) return fail(-1, "bad argument%s", "");
  if (bytes) VH_HIP(hipMemcpy(ddst, dsrc, (size_t)bytes, hipMemcpyDeviceToDevice));
  return 0;
}

// ---- tier 2 ----

int vh_grid_create(vh_grid** out, int32_t rank, const int64_t* dims, const double* lo, const double* hi,
                   const int32_t* bc, double gamma, int32_t dtype) {
  if (!out || !dims || !lo || !hi || !bc) return fail(-1, "null argument%s", "");
  *out = nullptr;
  if (rank < 1 || rank > 3) return fail(-1, "rank must be 1..3%s (got %lld)", "", rank);
  if (dtype != VH_F64 && dtype != VH_F32) return fail(-1, "dtype must be VH_F64 or VH_F32%s (got %lld)", "", dtype);
  if (!(gamma > 1.0)) return fail(-1, "gamma must exceed 1%s", "");
  vh_grid* g = new vh_grid();
  memset(g, 0, sizeof(*g));
  g->dtype = dtype;
  g->mode = VH_EXACT;
  g->esize = dtype == VH_F64 ? 8 : 4;
  GridDev& G = g->G;
  G.rank = rank;
  long long n = 1;
  for (int d = 0; d < 3; ++d) {
    if (d < rank) {
      if (dims[d] < 1 || dims[d] > (1 << 30)) { delete g; return fail(-1, "dims[%s] out of range (%lld)", "", dims[d]); }
      if (!(hi[d] > lo[d])) { delete g; return fail(-1, "domain needs hi > lo%s (axis %lld)", "", d); }
      G.n[d] = (int)dims[d];
      G.w[d] = (hi[d] - lo[d]) / (double)dims[d];
      for (int s = 0; s < 2; ++s) {
        const int b = bc[2 * d + s];
        if (b < 0 || b > 3) { delete g; return fail(-1, "unknown boundary code%s %lld", "", b); }
        G.bc[2 * d + s] = b;
      }
      if ((G.bc[2 * d] == BC_PERIODIC) != (G.bc[2 * d + 1] == BC_PERIODIC)) {
        delete g;
        return fail(-1, "periodic must be set on both sides of an axis%s (axis %lld)", "", d);
      }
      if (G.bc[2 * d] == BC_PERIODIC && dims[d] < 2) {
        delete g;
        return fail(-1, "a periodic axis needs at least 2 cells%s (axis %lld)", "", d);
      }
    } else {
      G.n[d] = 1;
      G.w[d] = 1.0;
      G.bc[2 * d] = G.bc[2 * d + 1] = BC_OUTFLOW;
    }
    n *= G.n[d];
  }
  G.ncell = n;
  {
    long long base = 0;
    for (int d = 0; d < 3; ++d) {
      G.bbase[2 * d] = base;
      G.bbase[2 * d + 1] = base + n / G.n[d];
      if (d < rank) base += 2 * (n / G.n[d]);
    }
  }
  G.gamma = gamma;
  G.igm1 = 1.0 / (gamma - 1.0);
  // lib/RefMesh/src/RefMesh.jl:127 and :135-142 (left folds in Float64).
  double V = G.w[0];
  for (int d = 1; d < rank; ++d) V = V * G.w[d];
  G.V = V;
  for (int a = 0; a < 3; ++a) {
    double ar = 1.0;
    for (int d = 0; d < rank; ++d)
      if (d != a) ar *= G.w[d];
    G.area[a] = ar;
  }
  if (hipGetDevice(&g->device) != hipSuccess) { delete g; return fail(-2, "no HIP device%s", ""); }
  // Default launch shape per rank and state type (measured on MI355X at 256^3 /
  // 512^2: Float64 16x16 cells, Float32 32x8, 16 z planes per workgroup); the
  // autotuner in the host package can override it per workload.
  int rc = configure(g, rank == 3 ? (g->dtype == VH_F64 ? 2 : 0) : 0, rank == 3 ? 16 : 0);
  if (rc) { vh_grid_destroy(g); return rc; }
  const size_t bytes = (size_t)5 * n * g->esize;
  if (hipMalloc(&g->U, bytes) != hipSuccess || hipMalloc(&g->U1, bytes) != hipSuccess ||
      hipMalloc(&g->clock, 4 * sizeof(double)) != hipSuccess ||
      hipMalloc(&g->totals_partials, 5 * kMaxCflBlocks * sizeof(double)) != hipSuccess) {
    vh_grid_destroy(g);
    return fail(-2, "device allocation of %s%lld bytes failed", "", (long long)(2 * bytes));
  }
  if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
    vh_grid_destroy(g);
    return fail(-2, "stream creation failed%s", "");
  }
  VH_HIP(hipMemsetAsync(g->clock, 0, 4 * sizeof(double), g->stream));
  VH_HIP(hipMemsetAsync(g->U, 0, bytes, g->stream));
  VH_HIP(hipStreamSynchronize(g->stream));
  const long long cb = (n + 255) / 256;
  g->cfl_blocks = (int)(cb < kMaxCflBlocks ? cb : kMaxCflBlocks);
  *out = g;
  return 0;
}

static void free_adj_work(vh_grid* g);

int vh_grid_destroy(vh_grid* g) {
  if (!g) return 0;
  if (g->stream) (void)hipStreamSynchronize(g->stream);
  free_adj_work(g);
  free_grav(g);
  if (g->U) (void)hipFree(g->U);
  if (g->U1) (void)hipFree(g->U1);
  if (g->partials) (void)hipFree(g->partials);
  if (g->clock) (void)hipFree(g->clock);
  if (g->totals_partials) (void)hipFree(g->totals_partials);
  if (g->pg) (void)hipFree(g->pg);
  if (g->stream) (void)hipStreamDestroy(g->stream);
  delete g;
  return 0;
}

int vh_grid_set_parent_ghost(vh_grid* g, const void* host) {
  if (!g || !host) return fail(-1, "null argument%s", "");
  if (!has_parent(g)) return fail(-1, "no side of this grid is ParentGhost%s", "");
  VH_HIP(hipSetDevice(g->device));
  long long nb = 0;
  for (int d = 0; d < g->G.rank; ++d) nb += 2 * (g->G.ncell / g->G.n[d]);
  const size_t bytes = (size_t)5 * nb * g->esize;
  VH_HIP(hipStreamSynchronize(g->stream));   // a running step may read the previous states
  if (!g->pg) VH_HIP(hipMalloc(&g->pg, bytes));
  VH_HIP(hipMemcpy(g->pg, host, bytes, hipMemcpyHostToDevice));
  g->G.pg = g->pg;
  return 0;
}

int64_t vh_grid_ncell(const vh_grid* g) { return g ? g->G.ncell : -1; }

int vh_grid_set_mode(vh_grid* g, int32_t mode) {
  if (!g) return fail(-1, "null argument%s", "");
  if (mode != VH_EXACT && mode != VH_FAST) return fail(-1, "mode must be VH_EXACT or VH_FAST%s (got %lld)", "", mode);
  VH_HIP(hipStreamSynchronize(g->stream));
  g->mode = mode;
  return 0;
}

int vh_grid_get_mode(const vh_grid* g) { return g ? g->mode : -1; }

int vh_grid_upload(vh_grid* g, const void* U) {
  if (!g || !U) return fail(-1, "null argument%s", "");
  VH_HIP(hipSetDevice(g->device));
  VH_HIP(hipMemcpyAsync(g->U, U, (size_t)5 * g->G.ncell * g->esize, hipMemcpyHostToDevice, g->stream));
  VH_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int vh_grid_download(vh_grid* g, void* U) {
  if (!g || !U) return fail(-1, "null argument%s", "");
  VH_HIP(hipSetDevice(g->device));
  VH_HIP(hipMemcpyAsync(U, g->U, (size_t)5 * g->G.ncell * g->esize, hipMemcpyDeviceToHost, g->stream));
  VH_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int vh_grid_state_device_ptr(vh_grid* g, void** dptr) {
  if (!g || !dptr) return fail(-1, "null argument%s", "");
  *dptr = g->U;
  return 0;
}

int vh_grid_compute_dt(vh_grid* g, double cfl, double* dt) {
  if (!g || !dt) return fail(-1, "null argument%s", "");
  VH_HIP(hipSetDevice(g->device));
  enqueue_cfl(g);
  hipLaunchKernelGGL(dt_finalize_kernel, dim3(1), dim3(256), 0, g->stream, g->partials, g->cfl_blocks, cfl, 0.0,
                     0, g->clock);
  VH_HIP(hipGetLastError());
  double clk[4];
  VH_HIP(hipMemcpyAsync(clk, g->clock, sizeof clk, hipMemcpyDeviceToHost, g->stream));
  VH_HIP(hipStreamSynchronize(g->stream));
  *dt = clk[1];
  return 0;
}

int vh_grid_step(vh_grid* g, double dt) {
  if (!g) return fail(-1, "null argument%s", "");
  if (int rc = need_parent_states(g)) return rc;
  VH_HIP(hipSetDevice(g->device));
  hipLaunchKernelGGL(set_dt_kernel, dim3(1), dim3(1), 0, g->stream, g->clock, dt);
  enqueue_step(g);
  VH_HIP(hipGetLastError());
  return 0;
}

int64_t vh_grid_bflux_count(const vh_grid* g) {
  if (!g) return -1;
  long long t = 0;
  for (int d = 0; d < g->G.rank; ++d) t += 2 * (g->G.ncell / g->G.n[d]);
  return t;
}

int vh_grid_step_bflux(vh_grid* g, double dt, void* dbflux) {
  if (!g || !dbflux) return fail(-1, "null argument%s", "");
  if (int rc = need_parent_states(g)) return rc;
  VH_HIP(hipSetDevice(g->device));
  hipLaunchKernelGGL(set_dt_kernel, dim3(1), dim3(1), 0, g->stream, g->clock, dt);
  g->G.bflux = dbflux;
  enqueue_step(g);
  g->G.bflux = nullptr;
  VH_HIP(hipGetLastError());
  // The register is the caller's buffer: complete it before returning.
  VH_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int vh_grid_run(vh_grid* g, double cfl, double tfinal, int64_t max_steps, double* t_io, int64_t* steps) {
  if (!g || !t_io) return fail(-1, "null argument%s", "");
  const bool clip = tfinal > 0.0;
  if (!clip && max_steps <= 0) return fail(-1, "need tfinal > 0 or max_steps > 0%s", "");
  if (int rc = need_parent_states(g)) return rc;
  VH_HIP(hipSetDevice(g->device));
  double clk[4] = {*t_io, 0.0, 0.0, 0.0};
  VH_HIP(hipMemcpyAsync(g->clock, clk, sizeof clk, hipMemcpyHostToDevice, g->stream));
  int64_t k = 0;
  enqueue_cfl(g);   // CFL partials of the current state
  int np = g->cfl_blocks;
  double t = *t_io;
  while (true) {
    if (max_steps > 0 && k >= max_steps) break;
    if (clip && !(t < tfinal * (1 - 1e-12))) break;
    if (g->grav) {   // φ(ρⁿ), held over the step (src/driver.jl:498)
      long long it;
      double rr;
      g->grav->B.rho = g->U;
      int rc = grav_solve(g->stream, g->dtype == VH_F64 ? 0 : 1, g->G, g->grav->P, g->grav->B, &it, &rr);
      if (rc) return rc;
    }
    hipLaunchKernelGGL(dt_finalize_kernel, dim3(1), dim3(256), 0, g->stream, g->partials, np, cfl, tfinal,
                       clip ? 1 : 0, g->clock);
    enqueue_step(g);
    VH_HIP(hipGetLastError());
    np = g->stage_blocks;
    ++k;
    if (clip) {
      VH_HIP(hipMemcpyAsync(clk, g->clock, sizeof clk, hipMemcpyDeviceToHost, g->stream));
      VH_HIP(hipStreamSynchronize(g->stream));
      t = clk[0];
    }
  }
  VH_HIP(hipMemcpyAsync(clk, g->clock, sizeof clk, hipMemcpyDeviceToHost, g->stream));
  VH_HIP(hipStreamSynchronize(g->stream));
  *t_io = clk[0];
  if (steps) *steps = k;
  return 0;
}

// ---- self-gravity (src/gravity.jl) ----

int vh_grid_enable_gravity(vh_grid* g, double G, int32_t phi_periodic, int64_t maxiter, double tol,
                           int32_t project_mean) {
  if (!g) return fail(-1, "null argument%s", "");
  if (phi_periodic < 0 || phi_periodic >= (1 << g->G.rank))
    return fail(-1, "phi_periodic has bits beyond the grid rank%s (%lld)", "", phi_periodic);
  if (maxiter < 0 || !(tol >= 0.0)) return fail(-1, "need maxiter >= 0 and tol >= 0%s", "");
  VH_HIP(hipSetDevice(g->device));
  VH_HIP(hipStreamSynchronize(g->stream));
  free_grav(g);
  GravWork* w = new GravWork();
  g->grav = w;
  w->P.G = G;
  w->P.gper = phi_periodic;
  w->P.project_mean = project_mean ? 1 : 0;
  w->P.maxiter = maxiter;
  w->P.tol = tol;
  const long long n = g->G.ncell;
  hipError_t e = hipMalloc(&w->B.phi, g->esize * n);
  if (e == hipSuccess) e = hipMalloc(&w->B.work, 4 * g->esize * n);
  if (e == hipSuccess) e = hipMalloc(&w->B.state, kGravStateBytes);
  if (e == hipSuccess) e = hipMalloc(&w->B.scal, 8 * sizeof(double));
  if (e == hipSuccess) e = hipMalloc(&w->B.partials, 4096 * sizeof(double));
  if (e == hipSuccess) e = hipHostMalloc(&w->B.host_state, kGravStateBytes);
  if (e == hipSuccess) e = hipMemsetAsync(w->B.phi, 0, g->esize * n, g->stream);
  if (e == hipSuccess) e = hipMemsetAsync(w->B.scal, 0, 8 * sizeof(double), g->stream);
  if (e == hipSuccess) e = hipMemsetAsync(w->B.state, 0, kGravStateBytes, g->stream);
  // Σ V accumulated cell by cell in Float64, as the reference's projections do.
  double sumv = 0.0;
  for (long long i = 0; i < n; ++i) sumv += g->G.V;
  if (e == hipSuccess) e = hipMemcpyAsync(w->B.scal + 1, &sumv, sizeof(double), hipMemcpyHostToDevice, g->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  if (e != hipSuccess) {
    free_grav(g);
    return fail(-2, "device allocation for self-gravity failed: %s", hipGetErrorString(e));
  }
  g->G.phi = w->B.phi;
  g->G.gravG = G;
  g->G.gper = phi_periodic;
  return 0;
}

int vh_grid_disable_gravity(vh_grid* g) {
  if (!g) return fail(-1, "null argument%s", "");
  VH_HIP(hipSetDevice(g->device));
  VH_HIP(hipStreamSynchronize(g->stream));
  free_grav(g);
  return 0;
}

int vh_grid_solve_poisson(vh_grid* g, int64_t* iters, double* relres) {
  if (!g || !iters || !relres) return fail(-1, "null argument%s", "");
  if (!g->grav) return fail(-1, "self-gravity is not enabled%s", "");
  VH_HIP(hipSetDevice(g->device));
  long long it = 0;
  g->grav->B.rho = g->U;
  int rc = grav_solve(g->stream, g->dtype == VH_F64 ? 0 : 1, g->G, g->grav->P, g->grav->B, &it, relres);
  *iters = it;
  return rc;
}

int vh_grid_phi_upload(vh_grid* g, const void* host) {
  if (!g || !host) return fail(-1, "null argument%s", "");
  if (!g->grav) return fail(-1, "self-gravity is not enabled%s", "");
  VH_HIP(hipSetDevice(g->device));
  VH_HIP(hipMemcpyAsync(g->grav->B.phi, host, g->esize * g->G.ncell, hipMemcpyHostToDevice, g->stream));
  VH_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int vh_grid_phi_download(vh_grid* g, void* host) {
  if (!g || !host) return fail(-1, "null argument%s", "");
  if (!g->grav) return fail(-1, "self-gravity is not enabled%s", "");
  VH_HIP(hipSetDevice(g->device));
  VH_HIP(hipMemcpyAsync(host, g->grav->B.phi, g->esize * g->G.ncell, hipMemcpyDeviceToHost, g->stream));
  VH_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int vh_grid_apply_laplacian(vh_grid* g, const void* dx, void* dout) {
  if (!g || !dx || !dout) return fail(-1, "null argument%s", "");
  if (!g->grav) return fail(-1, "self-gravity is not enabled%s", "");
  VH_HIP(hipSetDevice(g->device));
  grav_apply(g->stream, g->dtype == VH_F64 ? 0 : 1, g->G, g->grav->P.gper, dx, dout);
  VH_HIP(hipGetLastError());
  VH_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int vh_grid_gravity_accel(vh_grid* g, void* dg) {
  if (!g || !dg) return fail(-1, "null argument%s", "");
  if (!g->grav) return fail(-1, "self-gravity is not enabled%s", "");
  VH_HIP(hipSetDevice(g->device));
  grav_accel(g->stream, g->dtype == VH_F64 ? 0 : 1, g->G, g->grav->P.gper, g->grav->B.phi, dg);
  VH_HIP(hipGetLastError());
  VH_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int vh_grid_totals(vh_grid* g, double* out5) {
  if (!g || !out5) return fail(-1, "null argument%s", "");
  VH_HIP(hipSetDevice(g->device));
  const int blocks = g->cfl_blocks;
  if (g->dtype == VH_F64)
    hipLaunchKernelGGL(totals_kernel<double>, dim3(blocks), dim3(256), 0, g->stream, (const double*)g->U, g->G.ncell,
                       g->G.V, g->totals_partials);
  else
    hipLaunchKernelGGL(totals_kernel<float>, dim3(blocks), dim3(256), 0, g->stream, (const float*)g->U, g->G.ncell,
                       g->G.V, g->totals_partials);
  VH_HIP(hipGetLastError());
  double* h = new double[5 * blocks];
  hipError_t e = hipMemcpyAsync(h, g->totals_partials, sizeof(double) * 5 * blocks, hipMemcpyDeviceToHost, g->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
  if (e != hipSuccess) { delete[] h; return fail(-2, "HIP error %s in totals%s", hipGetErrorString(e)); }
  for (int f = 0; f < 5; ++f) out5[f] = 0.0;
  for (int b = 0; b < blocks; ++b)
    for (int f = 0; f < 5; ++f) out5[f] += h[5 * b + f];
  delete[] h;
  return 0;
}

int vh_grid_accumulate_flux(vh_grid* g, void* acc_host) {
  if (!g || !acc_host) return fail(-1, "null argument%s", "");
  if (int rc = need_parent_states(g)) return rc;
  VH_HIP(hipSetDevice(g->device));
  // Stage 1 with dt = −V gives invV = Tf(−V)/(Tf(1)·Tf(V)) = −1 exactly, so
  // U1 = U + acc; acc is recovered as U1 − U (rounded; inspection only).
  hipLaunchKernelGGL(set_dt_kernel, dim3(1), dim3(1), 0, g->stream, g->clock, -g->G.V);
  const long long n5 = 5 * g->G.ncell;
  const unsigned blocks = (unsigned)((n5 + 255) / 256);
  enqueue_stage(g, false, g->U, g->U, g->U1);
  if (g->dtype == VH_F64)
    hipLaunchKernelGGL(sub_kernel<double>, dim3(blocks), dim3(256), 0, g->stream, (const double*)g->U1,
                       (const double*)g->U, (double*)g->U1, n5);
  else
    hipLaunchKernelGGL(sub_kernel<float>, dim3(blocks), dim3(256), 0, g->stream, (const float*)g->U1,
                       (const float*)g->U, (float*)g->U1, n5);
  VH_HIP(hipGetLastError());
  VH_HIP(hipMemcpyAsync(acc_host, g->U1, (size_t)n5 * g->esize, hipMemcpyDeviceToHost, g->stream));
  VH_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int vh_grid_sync(vh_grid* g) {
  if (!g) return fail(-1, "null argument%s", "");
  VH_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int vh_grid_set_launch(vh_grid* g, int32_t tile_variant, int32_t zchunk) {
  if (!g) return fail(-1, "null argument%s", "");
  VH_HIP(hipSetDevice(g->device));
  VH_HIP(hipStreamSynchronize(g->stream));
  return configure(g, tile_variant, zchunk);
}

int vh_grid_tile_count(const vh_grid* g) { return g ? tile_count_for_rank(g->G.rank) : -1; }

int vh_grid_get_launch(const vh_grid* g, int32_t* tile_x, int32_t* tile_y, int32_t* zchunk, int64_t* workgroups) {
  if (!g) return fail(-1, "null argument%s", "");
  if (tile_x) *tile_x = g->tx;
  if (tile_y) *tile_y = g->ty;
  if (zchunk) *zchunk = g->G.zchunk;
  if (workgroups) *workgroups = g->stage_blocks;
  return 0;
}

int vh_grid_time_steps(vh_grid* g, double cfl, int64_t nsteps, double* ms_stage1, double* ms_stage2,
                       double* ms_step) {
  if (!g || nsteps <= 0 || nsteps > 100000) return fail(-1, "bad argument%s (nsteps %lld)", "", nsteps);
  if (int rc = need_parent_states(g)) return rc;
  VH_HIP(hipSetDevice(g->device));
  // Three events per step, recorded asynchronously; one synchronisation at the end.
  const int64_t ne = 3 * nsteps;
  hipEvent_t* ev = new hipEvent_t[ne]();
  int rc = 0;
  for (int64_t i = 0; i < ne && !rc; ++i)
    if (hipEventCreate(&ev[i]) != hipSuccess) rc = fail(-2, "hipEventCreate failed%s", "");
  if (!rc) {
    enqueue_cfl(g);
    int np = g->cfl_blocks;
    for (int64_t k = 0; k < nsteps; ++k) {
      hipLaunchKernelGGL(dt_finalize_kernel, dim3(1), dim3(256), 0, g->stream, g->partials, np, cfl, 0.0, 0,
                         g->clock);
      enqueue_step(g, ev[3 * k], ev[3 * k + 1], ev[3 * k + 2]);
      np = g->stage_blocks;
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(g->stream);
    if (e != hipSuccess) rc = fail(-2, "HIP error %s while timing%s", hipGetErrorString(e));
  }
  double s1 = 0, s2 = 0;
  float tot = 0;
  if (!rc) {
    for (int64_t k = 0; k < nsteps; ++k) {
      float a = 0, b = 0;
      (void)hipEventElapsedTime(&a, ev[3 * k], ev[3 * k + 1]);
      (void)hipEventElapsedTime(&b, ev[3 * k + 1], ev[3 * k + 2]);
      s1 += a;
      s2 += b;
    }
    (void)hipEventElapsedTime(&tot, ev[0], ev[ne - 1]);
  }
  for (int64_t i = 0; i < ne; ++i)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  delete[] ev;
  if (rc) return rc;
  if (ms_stage1) *ms_stage1 = s1 / nsteps;
  if (ms_stage2) *ms_stage2 = s2 / nsteps;
  // First stage-1 start to last stage-2 end, per step (includes the dt reduction
  // and clock kernels between steps).
  if (ms_step) *ms_step = (double)tot / nsteps;
  return 0;
}

}  // extern "C"

// ----------------------------------------------------------------------------
// reverse mode (Float64): one step's VJP and a checkpointed rollout

namespace vh {
__global__ void set_value_kernel(double* p, double v) { *p = v; }
__global__ void copy_value_kernel(const double* src, double* dst) { *dst = *src; }
}  // namespace vh

struct AdjWork {
  double* W = nullptr;
  double* gL[3] = {nullptr, nullptr, nullptr};
  double* gR[3] = {nullptr, nullptr, nullptr};
  double* lam1 = nullptr;
  double* U1 = nullptr;
  double* dt = nullptr;
  size_t bytes = 0;
  int fused = 1;   // use the single-kernel 2D periodic form where it applies
};

static void free_adj_work(vh_grid* g) {
  if (!g->adj) return;
  AdjWork* w = g->adj;
  (void)hipFree(w->W);
  for (int d = 0; d < 3; ++d) {
    (void)hipFree(w->gL[d]);
    (void)hipFree(w->gR[d]);
  }
  (void)hipFree(w->lam1);
  (void)hipFree(w->U1);
  (void)hipFree(w->dt);
  delete w;
  g->adj = nullptr;
}

static int ensure_adj_work(vh_grid* g) {
  if (g->adj) return 0;
  if (g->dtype != VH_F64) return fail(-1, "the reverse mode is Float64 only%s", "");
  AdjWork* w = new AdjWork();
  g->adj = w;
  const size_t nb = sizeof(double) * 5 * (size_t)g->G.ncell;
  VH_HIP(hipMalloc(&w->W, nb));
  VH_HIP(hipMalloc(&w->lam1, nb));
  VH_HIP(hipMalloc(&w->U1, nb));
  VH_HIP(hipMalloc(&w->dt, sizeof(double)));
  w->bytes = 3 * nb + sizeof(double);
  for (int d = 0; d < g->G.rank; ++d) {
    const size_t fb = sizeof(double) * 5 * (size_t)adj_face_slots(g->G, d);
    VH_HIP(hipMalloc(&w->gL[d], fb));
    VH_HIP(hipMalloc(&w->gR[d], fb));
    w->bytes += 2 * fb;
  }
  return 0;
}

// λ ← a1·base1 + a2·base2 + beta·(dt/V)·A'(U)ᵀλ_in, all device arrays.
static void enqueue_jat(vh_grid* g, const double* U, const double* lam_in, const double* base1, double a1,
                        const double* base2, double a2, double beta, const double* dtp, double* out) {
  AdjWork* w = g->adj;
  if (w->fused && adj_launch_jat_fused(g->stream, U, lam_in, base1, a1, base2, a2, beta, dtp, g->G, out)) return;
  adj_launch_prim(g->stream, U, w->W, g->G);
  adj_launch_faces(g->stream, w->W, lam_in, g->G, w->gL, w->gR);
  adj_launch_cells(g->stream, U, w->W, w->gL, w->gR, base1, a1, base2, a2, beta, dtp, g->G, out);
}

// VJP of one step from state U (stage-1 state U1 given, or recomputed into the
// work buffer when null): wbar → out (out may alias wbar).
static void enqueue_step_vjp(vh_grid* g, const double* U, const double* U1, const double* dtp, const double* wbar,
                             double* out) {
  AdjWork* w = g->adj;
  if (!U1) {
    if (g->mode == VH_FAST)
      fast::launch_stage(VH_F64, g->G.rank, g->variant, false, g->stage_blocks, g->stream, U, U, w->U1, dtp,
                         g->partials, g->G);
    else
      exact::launch_stage(VH_F64, g->G.rank, g->variant, false, g->stage_blocks, g->stream, U, U, w->U1, dtp,
                          g->partials, g->G);
    U1 = w->U1;
  }
  // λ1 = ½w − ½·invV·A'(U1)ᵀw
  enqueue_jat(g, U1, wbar, wbar, 0.5, nullptr, 0.0, -0.5, dtp, w->lam1);
  // ∂L/∂U = ½w + λ1 − invV·A'(U)ᵀλ1
  enqueue_jat(g, U, w->lam1, wbar, 0.5, w->lam1, 1.0, -1.0, dtp, out);
}

struct vh_adjoint {
  vh_grid* g = nullptr;
  int64_t nsteps = 0;
  int every = 20;
  int nseg = 0;
  double* dts = nullptr;
  double* ckpt = nullptr;
  double* seg = nullptr;
  double* seg1 = nullptr;
  double* w = nullptr;
  size_t bytes = 0;
  double* segL = nullptr;    // the last segment's U_k, kept from the forward run
  double* segL1 = nullptr;   // and its stage-1 states U1_k
  hipGraphExec_t graph = nullptr;   // the backward sweep (fixed buffers), captured once
  long long graph_key = -1;         // launch configuration the graph was captured with
  float fwd_ms = 0, bwd_ms = 0;
  bool have_forward = false;
};

// Reverse sweep over the segments, last first: the last segment's states come
// from the forward run; every other segment is recomputed from its checkpoint
// (U_k and U1_k), then the step VJPs run backwards through it on a->w.
static void enqueue_backward_sweep(vh_adjoint* a) {
  vh_grid* g = a->g;
  const size_t nb = sizeof(double) * 5 * (size_t)g->G.ncell;
  const size_t n5 = 5 * (size_t)g->G.ncell;
  for (int s = a->nseg - 1; s >= 0; --s) {
    const int64_t k0 = (int64_t)s * a->every;
    const int64_t k1 = std::min<int64_t>(a->nsteps, k0 + a->every);
    double* su = a->seg;
    double* su1 = a->seg1;
    if (s == a->nseg - 1) {
      su = a->segL;
      su1 = a->segL1;
    } else {
      (void)hipMemcpyAsync(a->seg, (char*)a->ckpt + nb * s, nb, hipMemcpyDeviceToDevice, g->stream);
      for (int64_t k = k0; k < k1; ++k) {
        double* Uk = a->seg + n5 * (k - k0);
        double* U1k = a->seg1 + n5 * (k - k0);
        enqueue_stage_dt(g, false, Uk, Uk, U1k, a->dts + k);
        if (k + 1 < k1) enqueue_stage_dt(g, true, U1k, Uk, a->seg + n5 * (k + 1 - k0), a->dts + k);
      }
    }
    for (int64_t k = k1 - 1; k >= k0; --k)
      enqueue_step_vjp(g, su + n5 * (k - k0), su1 + n5 * (k - k0), a->dts + k, a->w, a->w);
  }
}

extern "C" {

int vh_grid_step_vjp(vh_grid* g, const void* dU, double dt, const void* dWbar, void* dUbar) {
  if (!g || !dU || !dWbar || !dUbar) return fail(-1, "null argument%s", "");
  if (g->grav) return fail(-1, "the reverse mode does not cover self-gravity%s", "");
  if (has_parent(g)) return fail(-1, "the reverse mode does not cover ParentGhost boundaries%s", "");
  VH_HIP(hipSetDevice(g->device));
  int rc = ensure_adj_work(g);
  if (rc) return rc;
  hipLaunchKernelGGL(set_value_kernel, dim3(1), dim3(1), 0, g->stream, g->adj->dt, dt);
  enqueue_step_vjp(g, (const double*)dU, nullptr, g->adj->dt, (const double*)dWbar, (double*)dUbar);
  VH_HIP(hipGetLastError());
  VH_HIP(hipStreamSynchronize(g->stream));
  return 0;
}

int vh_grid_set_adjoint_form(vh_grid* g, int32_t fused) {
  if (!g) return fail(-1, "null argument%s", "");
  VH_HIP(hipSetDevice(g->device));
  int rc = ensure_adj_work(g);
  if (rc) return rc;
  VH_HIP(hipStreamSynchronize(g->stream));
  g->adj->fused = fused ? 1 : 0;
  return 0;
}

int vh_adjoint_destroy(vh_adjoint* a) {
  if (!a) return 0;
  if (a->g && a->g->stream) (void)hipStreamSynchronize(a->g->stream);
  (void)hipFree(a->dts);
  (void)hipFree(a->ckpt);
  (void)hipFree(a->seg);
  (void)hipFree(a->seg1);
  (void)hipFree(a->segL);
  (void)hipFree(a->segL1);
  (void)hipFree(a->w);
  if (a->graph) (void)hipGraphExecDestroy(a->graph);
  delete a;
  return 0;
}

int vh_adjoint_create(vh_adjoint** out, vh_grid* g, int64_t nsteps, int32_t every) {
  if (!out || !g) return fail(-1, "null argument%s", "");
  *out = nullptr;
  if (nsteps < 1 || every < 1) return fail(-1, "need nsteps >= 1 and every >= 1%s", "");
  if (g->dtype != VH_F64) return fail(-1, "the reverse mode is Float64 only%s", "");
  if (g->grav) return fail(-1, "the reverse mode does not cover self-gravity%s", "");
  if (has_parent(g)) return fail(-1, "the reverse mode does not cover ParentGhost boundaries%s", "");
  VH_HIP(hipSetDevice(g->device));
  int rc = ensure_adj_work(g);
  if (rc) return rc;
  vh_adjoint* a = new vh_adjoint();
  a->g = g;
  a->nsteps = nsteps;
  a->every = every > nsteps ? (int)nsteps : every;
  a->nseg = (int)((nsteps + a->every - 1) / a->every);
  const size_t nb = sizeof(double) * 5 * (size_t)g->G.ncell;
  hipError_t e = hipMalloc(&a->dts, sizeof(double) * nsteps);
  if (e == hipSuccess) e = hipMalloc(&a->ckpt, nb * a->nseg);
  if (e == hipSuccess) e = hipMalloc(&a->seg, nb * a->every);
  if (e == hipSuccess) e = hipMalloc(&a->seg1, nb * a->every);
  if (e == hipSuccess) e = hipMalloc(&a->segL, nb * a->every);
  if (e == hipSuccess) e = hipMalloc(&a->segL1, nb * a->every);
  if (e == hipSuccess) e = hipMalloc(&a->w, nb);
  if (e != hipSuccess) {
    vh_adjoint_destroy(a);
    return fail(-2, "device allocation for the rollout failed: %s", hipGetErrorString(e));
  }
  a->bytes = sizeof(double) * nsteps + nb * (a->nseg + 4 * (size_t)a->every + 1) + g->adj->bytes;
  *out = a;
  return 0;
}

int vh_adjoint_forward(vh_adjoint* a, double cfl) {
  if (!a) return fail(-1, "null argument%s", "");
  vh_grid* g = a->g;
  if (g->grav) return fail(-1, "the reverse mode does not cover self-gravity%s", "");
  VH_HIP(hipSetDevice(g->device));
  const size_t nb = sizeof(double) * 5 * (size_t)g->G.ncell;
  hipEvent_t e0, e1;
  VH_HIP(hipEventCreate(&e0));
  VH_HIP(hipEventCreate(&e1));
  VH_HIP(hipEventRecord(e0, g->stream));
  VH_HIP(hipMemcpyAsync(a->ckpt, g->U, nb, hipMemcpyDeviceToDevice, g->stream));
  enqueue_cfl(g);
  int np = g->cfl_blocks;
  // The last segment's states are kept as the forward run produces them (the
  // backward pass starts there and needs no recompute for it).
  const size_t n5 = 5 * (size_t)g->G.ncell;
  const int64_t klast = (int64_t)(a->nseg - 1) * a->every;
  for (int64_t k = 0; k < a->nsteps; ++k) {
    hipLaunchKernelGGL(dt_finalize_kernel, dim3(1), dim3(256), 0, g->stream, g->partials, np, cfl, 0.0, 0, g->clock);
    hipLaunchKernelGGL(copy_value_kernel, dim3(1), dim3(1), 0, g->stream, g->clock + 1, a->dts + k);
    if (k < klast) {
      enqueue_step(g);
    } else {
      const int64_t j = k - klast;
      if (j == 0) VH_HIP(hipMemcpyAsync(a->segL, g->U, nb, hipMemcpyDeviceToDevice, g->stream));
      double* Uk = a->segL + n5 * j;
      double* U1k = a->segL1 + n5 * j;
      double* Un = k + 1 < a->nsteps ? a->segL + n5 * (j + 1) : (double*)g->U;
      enqueue_stage(g, false, Uk, Uk, U1k);
      enqueue_stage(g, true, U1k, Uk, Un);
      hipLaunchKernelGGL(advance_clock_kernel, dim3(1), dim3(1), 0, g->stream, g->clock);
    }
    np = g->stage_blocks;
    if ((k + 1) % a->every == 0 && k + 1 < a->nsteps)
      VH_HIP(hipMemcpyAsync((char*)a->ckpt + nb * ((k + 1) / a->every), g->U, nb, hipMemcpyDeviceToDevice, g->stream));
  }
  VH_HIP(hipEventRecord(e1, g->stream));
  VH_HIP(hipGetLastError());
  VH_HIP(hipEventSynchronize(e1));
  VH_HIP(hipEventElapsedTime(&a->fwd_ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  a->have_forward = true;
  return 0;
}

int vh_adjoint_backward(vh_adjoint* a, const void* dWbar, void* dUbar) {
  if (!a || !dWbar || !dUbar) return fail(-1, "null argument%s", "");
  if (!a->have_forward) return fail(-1, "run vh_adjoint_forward first%s", "");
  vh_grid* g = a->g;
  VH_HIP(hipSetDevice(g->device));
  const size_t nb = sizeof(double) * 5 * (size_t)g->G.ncell;
  const size_t n5 = 5 * (size_t)g->G.ncell;
  hipEvent_t e0, e1;
  VH_HIP(hipEventCreate(&e0));
  VH_HIP(hipEventCreate(&e1));
  VH_HIP(hipEventRecord(e0, g->stream));
  VH_HIP(hipMemcpyAsync(a->w, dWbar, nb, hipMemcpyDeviceToDevice, g->stream));
  const long long key = (((long long)g->mode * 64 + g->variant) * 2 + g->adj->fused) * (1LL << 32) + g->stage_blocks;
  if (a->graph && a->graph_key != key) {
    (void)hipGraphExecDestroy(a->graph);
    a->graph = nullptr;
  }
  if (!a->graph) {
    // The sweep reads and writes only the rollout's own buffers, so one capture
    // serves every later call: launch gaps between its ~4 kernels per step go away.
    hipGraph_t gr = nullptr;
    VH_HIP(hipStreamBeginCapture(g->stream, hipStreamCaptureModeThreadLocal));
    enqueue_backward_sweep(a);
    VH_HIP(hipStreamEndCapture(g->stream, &gr));
    hipError_t e = hipGraphInstantiate(&a->graph, gr, nullptr, nullptr, 0);
    (void)hipGraphDestroy(gr);
    if (e != hipSuccess) {
      a->graph = nullptr;
      return fail(-2, "graph instantiation failed: %s", hipGetErrorString(e));
    }
    a->graph_key = key;
  }
  VH_HIP(hipGraphLaunch(a->graph, g->stream));
  VH_HIP(hipMemcpyAsync(dUbar, a->w, nb, hipMemcpyDeviceToDevice, g->stream));
  VH_HIP(hipEventRecord(e1, g->stream));
  VH_HIP(hipGetLastError());
  VH_HIP(hipEventSynchronize(e1));
  VH_HIP(hipEventElapsedTime(&a->bwd_ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return 0;
}

int vh_adjoint_info(const vh_adjoint* a, double* ms_forward, double* ms_backward, int64_t* device_bytes,
                    double* dts_host) {
  if (!a) return fail(-1, "null argument%s", "");
  if (ms_forward) *ms_forward = a->fwd_ms;
  if (ms_backward) *ms_backward = a->bwd_ms;
  if (device_bytes) *device_bytes = (int64_t)a->bytes;
  if (dts_host) {
    if (!a->have_forward) return fail(-1, "run vh_adjoint_forward first%s", "");
    VH_HIP(hipMemcpy(dts_host, a->dts, sizeof(double) * a->nsteps, hipMemcpyDeviceToHost));
  }
  return 0;
}

}  // extern "C"
