// FP64 global reductions: checksum/L2/max diagnostics, the error metrics
// of the reference's compare.py / verify.sh, and the time-step bound.
//
// Every reduction is two deterministic passes: a fixed grid of workgroups
// writes one partial record each (per-lane Neumaier-compensated sums,
// then a lane-shuffle butterfly inside each wavefront and the wavefront
// results in order), and a single workgroup folds the partials in index
// order.  The grid depends only on the array shape, so repeated
// calls give bit-identical results.  Inputs are read once: these kernels
// are HBM-bound like the stencil.
#include <math.h>

#include <mutex>

#include "h3d_internal.h"

namespace {

constexpr int RB = 256;      // threads per workgroup
constexpr int MAXPART = 1024; // partial records

struct Ksum {  // Neumaier pair
  double s, c;
  __device__ void add(double x) {
#pragma clang fp contract(off)
    double t = s + x;
    if (fabs(s) >= fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
  }
  __device__ void merge(const Ksum &o) {
    add(o.s);
    c += o.c;
  }
  __device__ double value() const { return s + c; }
};

// ---- record types -------------------------------------------------------

struct Stats {
  Ksum sum, sq;
  double mx, bad;
  __device__ void init() { sum = {0, 0}; sq = {0, 0}; mx = 0; bad = 0; }
  __device__ void take(double v) {
#pragma clang fp contract(off)
    if (!isfinite(v)) { bad += 1.0; return; }
    sum.add(v);
    sq.add(v * v);
    mx = fmax(mx, fabs(v));
  }
  __device__ void merge(const Stats &o) {
    sum.merge(o.sum);
    sq.merge(o.sq);
    mx = fmax(mx, o.mx);
    bad += o.bad;
  }
};

// (value, index) ordered like numpy's argmax over |d|: NaN is the largest,
// ties go to the lower index.
__device__ inline bool first_is_max(double a, double ia, double b, double ib) {
  const bool na = isnan(a), nb = isnan(b);
  if (na && nb) return ia <= ib;
  if (na) return true;
  if (nb) return false;
  if (a > b) return true;
  if (a < b) return false;
  return ia <= ib;
}

struct Cmp {
  Ksum sabs, sq, rsq;
  double mx, imx, mrel, nnan, ninf, nfail, relnan;
  __device__ void init() {
    sabs = {0, 0}; sq = {0, 0}; rsq = {0, 0};
    mx = -1.0; imx = 1e300; mrel = 0; nnan = 0; ninf = 0; nfail = 0; relnan = 0;
  }
  __device__ void take(double r, double c, double idx, double atol, double rtol) {
    // numpy's unfused arithmetic: |d| <= atol + rtol*|ref| rounds the
    // product before the sum, and d*d, r*r are rounded before they are
    // summed (the file is also built without contraction)
#pragma clang fp contract(off)
    if (isnan(c)) nnan += 1.0;
    else if (isinf(c)) ninf += 1.0;
    const double d = c - r, ad = fabs(d), ar = fabs(r);
    if (first_is_max(ad, idx, mx, imx)) { mx = ad; imx = idx; }
    sabs.add(ad);
    sq.add(d * d);
    rsq.add(r * r);
    if (ar > 1e-12) {
      const double rel = ad / fmax(ar, 1e-30);
      if (isnan(rel)) relnan = 1.0;
      else mrel = fmax(mrel, rel);
    }
    if (!isfinite(c) || !(ad <= atol + rtol * ar)) nfail += 1.0;
  }
  __device__ void merge(const Cmp &o) {
    sabs.merge(o.sabs);
    sq.merge(o.sq);
    rsq.merge(o.rsq);
    if (o.imx < 1e299 && first_is_max(o.mx, o.imx, mx, imx)) { mx = o.mx; imx = o.imx; }
    mrel = fmax(mrel, o.mrel);
    nnan += o.nnan; ninf += o.ninf; nfail += o.nfail;
    relnan = fmax(relnan, o.relnan);
  }
};

struct Fmax {
  double m;
  __device__ void init() { m = 0.0; }
  __device__ void merge(const Fmax &o) { m = fmax(m, o.m); }
};

// A record moved between lanes one dword at a time (records are plain
// doubles).
template <typename R>
__device__ R shfl_xor_record(const R &r, int mask) {
  static_assert(sizeof(R) % sizeof(int) == 0, "record of dwords");
  R o;
  const int *src = reinterpret_cast<const int *>(&r);
  int *dst = reinterpret_cast<int *>(&o);
#pragma unroll
  for (int k = 0; k < (int)(sizeof(R) / sizeof(int)); ++k)
    dst[k] = __shfl_xor(src[k], mask, 64);
  return o;
}

// Workgroup fold: a butterfly over the 64 lanes of each wavefront (lane
// shuffles, fixed pairing, so the result is deterministic), then the RB/64
// wavefront results merged in wavefront order.  The result is in sh[0].
template <typename R>
__device__ void block_fold(R &acc, R *sh) {
#pragma unroll
  for (int m = 32; m > 0; m >>= 1) {
    const R o = shfl_xor_record(acc, m);
    acc.merge(o);
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) sh[wave] = acc;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < RB / 64; ++w) sh[0].merge(sh[w]);
  __syncthreads();
}

template <typename R>
__global__ void __launch_bounds__(RB) fold_partials(R *part, int n) {
  __shared__ R sh[RB / 64];
  R acc;
  acc.init();
  // lane t folds partials t, t+RB, ... in order, then a fixed tree
  for (int p = threadIdx.x; p < n; p += RB) acc.merge(part[p]);
  block_fold(acc, sh);
  if (threadIdx.x == 0) part[0] = sh[0];
}

// ---- pass-1 kernels -----------------------------------------------------

template <typename T>
__global__ void __launch_bounds__(RB)
stats_kernel(const T *__restrict__ u, int nx, int ny, int nz, int sy,
             long long sz, Stats *part) {
  __shared__ Stats sh[RB / 64];
  Stats acc;
  acc.init();
  const long long rows = (long long)ny * nz;
  for (long long r = blockIdx.x; r < rows; r += gridDim.x) {
    const int k = (int)(r / ny), j = (int)(r % ny);
    const T *row = u + (long long)k * sz + (long long)j * sy;
    for (int i = threadIdx.x; i < nx; i += RB) acc.take((double)row[i]);
  }
  block_fold(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

template <typename T>
__global__ void __launch_bounds__(RB)
cmp_kernel(const double *__restrict__ ref, const T *__restrict__ cand, int nx,
           int ny, int nz, int sy, long long sz, double atol, double rtol,
           Cmp *part) {
  __shared__ Cmp sh[RB / 64];
  Cmp acc;
  acc.init();
  const long long rows = (long long)ny * nz;
  for (long long r = blockIdx.x; r < rows; r += gridDim.x) {
    const int k = (int)(r / ny), j = (int)(r % ny);
    const long long base = (long long)k * sz + (long long)j * sy;
    const double flat0 = (double)r * nx;
    for (int i = threadIdx.x; i < nx; i += RB)
      acc.take(ref[base + i], (double)cand[base + i], flat0 + i, atol, rtol);
  }
  block_fold(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void __launch_bounds__(RB)
fmax_kernel(const double *__restrict__ kp, int nx, int ny, int nz, int sy,
            long long sz, Fmax *part) {
  __shared__ Fmax sh[RB / 64];
  Fmax acc;
  acc.init();
  const long long rows = (long long)(ny - 2) * (nz - 2);
  for (long long r = blockIdx.x; r < rows; r += gridDim.x) {
    const int k = 1 + (int)(r / (ny - 2)), j = 1 + (int)(r % (ny - 2));
    const long long base = (long long)k * sz + (long long)j * sy;
    for (int i = 1 + threadIdx.x; i < nx - 1; i += RB) {
      const long long c = base + i;
      const double kc = kp[c];
      const double s = (kp[c + 1] + kc) + (kc + kp[c - 1]) + (kp[c + sy] + kc) +
                       (kc + kp[c - sy]) + (kp[c + sz] + kc) + (kc + kp[c - sz]);
      acc.m = fmax(acc.m, s);
    }
  }
  block_fold(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

int nparts_for(long long rows) {
  return (int)(rows < MAXPART ? (rows > 0 ? rows : 1) : MAXPART);
}

// Runs pass 1 (launcher `l1`) and pass 2, copies the folded record home.
template <typename R, typename L1>
int reduce_to_host(R *host, int parts, hipStream_t s, L1 l1) {
  void *ws = nullptr;
  H3D_TRY(hipMallocAsync(&ws, sizeof(R) * parts, s));
  R *part = static_cast<R *>(ws);
  int rc = l1(part);
  if (rc == H3D_OK) {
    hipLaunchKernelGGL((fold_partials<R>), dim3(1), dim3(RB), 0, s, part, parts);
    rc = h3d_launch_status();
  }
  if (rc == H3D_OK) {
    hipError_t e = hipMemcpyAsync(host, part, sizeof(R), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e != hipSuccess) rc = (int)e;
  }
  hipError_t ef = hipFreeAsync(ws, s);
  if (rc == H3D_OK && ef != hipSuccess) rc = (int)ef;
  return rc;
}

template <typename T>
int stats_impl(const T *u, const h3d_grid *g, double *out, void *stream) {
  if (!u || !out || !h3d_grid_ok(g)) return H3D_EARG;
  hipStream_t s = h3d_stream(stream);
  const int parts = nparts_for((long long)g->ny * g->nz);
  Stats r;
  int rc = reduce_to_host(&r, parts, s, [&](Stats *part) {
    hipLaunchKernelGGL((stats_kernel<T>), dim3(parts), dim3(RB), 0, s, u,
                       g->nx, g->ny, g->nz, g->sy, g->sz, part);
    return h3d_launch_status();
  });
  if (rc != H3D_OK) return rc;
  out[0] = r.sum.s + r.sum.c;
  out[1] = r.sq.s + r.sq.c;
  out[2] = r.mx;
  out[3] = r.bad;
  return H3D_OK;
}

template <typename T>
int cmp_impl(const double *ref, const T *cand, const h3d_grid *g, double atol,
             double rtol, double *out, void *stream) {
  if (!ref || !cand || !out || !h3d_grid_ok(g)) return H3D_EARG;
  hipStream_t s = h3d_stream(stream);
  const int parts = nparts_for((long long)g->ny * g->nz);
  Cmp r;
  int rc = reduce_to_host(&r, parts, s, [&](Cmp *part) {
    hipLaunchKernelGGL((cmp_kernel<T>), dim3(parts), dim3(RB), 0, s, ref, cand,
                       g->nx, g->ny, g->nz, g->sy, g->sz, atol, rtol, part);
    return h3d_launch_status();
  });
  if (rc != H3D_OK) return rc;
  const double n = (double)g->nx * g->ny * g->nz;
  const double sq = r.sq.s + r.sq.c, rsq = r.rsq.s + r.rsq.c;
  const double l2d = sqrt(sq), l2r = sqrt(rsq);
  out[0] = r.mx;
  out[1] = (r.sabs.s + r.sabs.c) / n;
  out[2] = sqrt(sq / n);
  out[3] = l2r > 0.0 ? l2d / l2r : l2d;
  out[4] = r.relnan > 0 ? NAN : r.mrel;
  out[5] = r.imx;
  out[6] = r.nnan;
  out[7] = r.ninf;
  out[8] = r.nfail;
  return H3D_OK;
}

}  // namespace

extern "C" int h3d_stats_f64(const double *u, const h3d_grid *g, double *out,
                             void *stream) {
  return stats_impl<double>(u, g, out, stream);
}
extern "C" int h3d_stats_f32(const float *u, const h3d_grid *g, double *out,
                             void *stream) {
  return stats_impl<float>(u, g, out, stream);
}
extern "C" int h3d_compare_f64(const double *ref, const double *cand,
                               const h3d_grid *g, double atol, double rtol,
                               double *out, void *stream) {
  return cmp_impl<double>(ref, cand, g, atol, rtol, out, stream);
}
extern "C" int h3d_compare_f32(const double *ref, const float *cand,
                               const h3d_grid *g, double atol, double rtol,
                               double *out, void *stream) {
  return cmp_impl<float>(ref, cand, g, atol, rtol, out, stream);
}
// ---- h3d_stats_f64 without allocation ----------------------------------
//
// The same two passes as h3d_stats_f64, with the partial records in a
// fixed array of the library's own device memory (one copy per GPU, kept
// until the process ends) instead of a stream-ordered allocation, and the
// four results written to the caller's device memory, ordered on the
// stream, with no host synchronisation.  For callers that cannot allocate
// device memory (code under graph capture, loops that forbid hipMalloc).
// The partition into records is h3d_stats_f64's, so the results carry the
// same bits.  Every record a call folds is written by that call first, so
// nothing needs clearing between calls; calls that overlap on one GPU
// share the array and race.

namespace {

constexpr int NOALLOC_PARTS = MAXPART;  // records: 1024 x 48 bytes
__device__ __align__(256) Stats noalloc_scratch[NOALLOC_PARTS];

// Device address of this device's copy of noalloc_scratch, looked up once
// per device with that device current (the array's name is a host-side
// handle, not an address a kernel may use).
std::mutex noalloc_mu;
Stats *noalloc_ptr[64] = {};

int noalloc_part(Stats **out) {
  int dev = 0;
  H3D_TRY(hipGetDevice(&dev));
  if (dev < 0 || dev >= 64) return H3D_ELIMIT;
  std::lock_guard<std::mutex> lk(noalloc_mu);
  if (!noalloc_ptr[dev]) {
    void *p = nullptr;
    H3D_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)));
    noalloc_ptr[dev] = static_cast<Stats *>(p);
  }
  *out = noalloc_ptr[dev];
  return H3D_OK;
}

__global__ void stats_emit(const Stats *part, double *out) {
  if (threadIdx.x == 0) {
    const Stats r = part[0];
    out[0] = r.sum.s + r.sum.c;
    out[1] = r.sq.s + r.sq.c;
    out[2] = r.mx;
    out[3] = r.bad;
  }
}

}  // namespace

extern "C" int h3d_stats_f64_noalloc(const double *u, const h3d_grid *g, double *out_dev,
                                     void *stream) {
  if (!u || !out_dev || !h3d_grid_ok(g)) return H3D_EARG;
  const int parts = nparts_for((long long)g->ny * g->nz);
  if (parts > NOALLOC_PARTS) return H3D_ELIMIT;
  Stats *part = nullptr;
  const int rc = noalloc_part(&part);
  if (rc != H3D_OK) return rc;
  hipStream_t s = h3d_stream(stream);
  hipLaunchKernelGGL((stats_kernel<double>), dim3(parts), dim3(RB), 0, s, u, g->nx, g->ny,
                     g->nz, g->sy, g->sz, part);
  int lrc = h3d_launch_status();
  if (lrc != H3D_OK) return lrc;
  hipLaunchKernelGGL((fold_partials<Stats>), dim3(1), dim3(RB), 0, s, part, parts);
  lrc = h3d_launch_status();
  if (lrc != H3D_OK) return lrc;
  hipLaunchKernelGGL(stats_emit, dim3(1), dim3(64), 0, s, part, out_dev);
  return h3d_launch_status();
}

extern "C" int h3d_face_sum_max_f64(const double *kappa, const h3d_grid *g,
                                    double *out, void *stream) {
  if (!kappa || !out || !h3d_grid_ok(g)) return H3D_EARG;
  hipStream_t s = h3d_stream(stream);
  const int parts = nparts_for((long long)(g->ny - 2) * (g->nz - 2));
  Fmax r;
  int rc = reduce_to_host(&r, parts, s, [&](Fmax *part) {
    hipLaunchKernelGGL(fmax_kernel, dim3(parts), dim3(RB), 0, s, kappa, g->nx,
                       g->ny, g->nz, g->sy, g->sz, part);
    return h3d_launch_status();
  });
  if (rc != H3D_OK) return rc;
  *out = r.m;
  return H3D_OK;
}
