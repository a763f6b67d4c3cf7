// Shared device helpers and argument checks of the h3d HIP library.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <atomic>

#include "h3d.h"

#define H3D_TRY(expr)                                  \
  do {                                                 \
    hipError_t h3d_e_ = (expr);                        \
    if (h3d_e_ != hipSuccess) return (int)h3d_e_;      \
  } while (0)

// Status of the last launch on this thread, as an h3d return code.
static inline int h3d_launch_status() {
  hipError_t e = hipGetLastError();
  return e == hipSuccess ? H3D_OK : (int)e;
}

static inline hipStream_t h3d_stream(void *s) {
  return reinterpret_cast<hipStream_t>(s);
}

// A grid the kernels can address: positive sizes, strides that hold a row
// and a plane, and at least one interior point per axis.
static inline bool h3d_grid_ok(const h3d_grid *g) {
  if (!g) return false;
  if (g->nx < 3 || g->ny < 3 || g->nz < 3) return false;
  if (g->sy < g->nx) return false;
  if (g->sz < (long long)g->sy * g->ny) return false;
  // within-plane byte offsets are 32-bit in the kernels (8-byte elements)
  if (g->sz > (1LL << 28)) return false;
  return true;
}

// Timing events and a captured graph of a run, released on every return
// path (early returns after a failed launch included).
struct h3d_run_res {
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipGraph_t gr = nullptr;
  hipGraphExec_t ge = nullptr;
  h3d_run_res() = default;
  h3d_run_res(const h3d_run_res &) = delete;
  h3d_run_res &operator=(const h3d_run_res &) = delete;
  ~h3d_run_res() {
    if (ge) (void)hipGraphExecDestroy(ge);
    if (gr) (void)hipGraphDestroy(gr);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
  }
};

// Lets kernel `fn` use up to the CU's 160 KB of LDS as dynamic shared
// memory (above the 64 KB default), once per kernel and device: bit d of
// *done records device d (devices past 63 set the attribute every call).
static inline int h3d_allow_big_lds(const void *fn, std::atomic<unsigned long long> *done) {
  int dev = 0;
  H3D_TRY(hipGetDevice(&dev));
  const unsigned long long bit = dev < 64 ? 1ull << dev : 0ull;
  if (bit && (done->load(std::memory_order_relaxed) & bit)) return H3D_OK;
  H3D_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
  done->fetch_or(bit, std::memory_order_relaxed);
  return H3D_OK;
}

static inline size_t h3d_grid_elems(const h3d_grid *g) {
  return (size_t)g->sz * (size_t)g->nz;
}

// XCD-aware tile order.  Workgroups are dealt round-robin to the 8 XCDs
// (workgroup b runs on XCD b % 8), each with its own L2.  Remapping so
// that XCD x receives a contiguous run of logical tiles keeps neighbouring
// tiles, which share halo rows, in one L2.  XCD x gets q+1 workgroups if
// x < r (nb = 8q + r), else q; its run starts after those of XCDs 0..x-1,
// so the map is a bijection of [0, nb) for every nb.
__device__ __forceinline__ unsigned h3d_xcd_tile(unsigned b, unsigned nb) {
  const unsigned nxcd = 8, q = nb / nxcd, r = nb % nxcd;
  const unsigned x = b % nxcd, p = b / nxcd;
  return x * q + (x < r ? x : r) + p;
}

// Whole-wavefront lane shifts through DPP (wave_shr:1 / wave_shl:1): lane
// l receives lane l-1 (shr) or l+1 (shl); the lane with no source keeps
// `edge`.  64-bit values move as two dwords.
__device__ __forceinline__ double h3d_from_left(double v, double edge) {
  int2 s = *reinterpret_cast<int2 *>(&v), o = *reinterpret_cast<int2 *>(&edge), r;
  r.x = __builtin_amdgcn_update_dpp(o.x, s.x, 0x138, 0xf, 0xf, false);
  r.y = __builtin_amdgcn_update_dpp(o.y, s.y, 0x138, 0xf, 0xf, false);
  return *reinterpret_cast<double *>(&r);
}
__device__ __forceinline__ double h3d_from_right(double v, double edge) {
  int2 s = *reinterpret_cast<int2 *>(&v), o = *reinterpret_cast<int2 *>(&edge), r;
  r.x = __builtin_amdgcn_update_dpp(o.x, s.x, 0x130, 0xf, 0xf, false);
  r.y = __builtin_amdgcn_update_dpp(o.y, s.y, 0x130, 0xf, 0xf, false);
  return *reinterpret_cast<double *>(&r);
}
__device__ __forceinline__ float h3d_from_left(float v, float edge) {
  int r = __builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v),
                                      0x138, 0xf, 0xf, false);
  return __int_as_float(r);
}
__device__ __forceinline__ float h3d_from_right(float v, float edge) {
  int r = __builtin_amdgcn_update_dpp(__float_as_int(edge), __float_as_int(v),
                                      0x130, 0xf, 0xf, false);
  return __int_as_float(r);
}

// Library-internal entry points shared between translation units.
int h3d_default_tune_f64(const h3d_grid *g, int k_start, int k_end,
                         h3d_tune *t);
// FP64 pass shape geometry: resident workgroups per CU, tiles per z chunk.
int h3d_pass_f64_geometry(const h3d_grid *g, int kblock, int rows, int waves, int rhs,
                          int *wg_per_cu, long long *tiles);

// Passes captured in one hipGraph by the run functions (even, so that a
// replay keeps the ping-pong parity); longer runs replay it.
#define H3D_GRAPH_PASSES 1024

// FP32 pass shapes: default launch shape of a depth, a shape's geometry
// (resident workgroups per CU, tiles per z chunk) and its best z chunks.
int h3d_default_tune_f32(const h3d_grid *g, int kblock, int mode, int bf16, h3d_tune *t);
int h3d_pass_f32_geometry(const h3d_grid *g, int kblock, int rows, int waves, int mode,
                          int bf16, int *wg_per_cu, long long *tiles);
int h3d_pass_f32_zchunks(const h3d_grid *g, int kblock, int rows, int waves, int mode,
                         int bf16, int *zc, int n, int *got);

// The autotuner's shape and device time per pass for an FP64 problem on
// grid g (mode 1: with rhs) at pass depth kblock, from a search at that
// fixed depth or a free search that chose it; H3D_ELIMIT when it has none.
int h3d_tuned_f64(const h3d_grid *g, int mode, int kblock, h3d_tune *t, float *ms_per_pass);

// Up to n z-chunk lengths (best first) for a pass of `kblock` steps over
// `planes` planes whose launch has `tiles` workgroups per z chunk and
// `wg_per_cu` resident workgroups per CU; *got receives how many were
// written.
int h3d_round_zchunks(int planes, int kblock, long long tiles, int wg_per_cu, int *zc, int n,
                      int *got);
