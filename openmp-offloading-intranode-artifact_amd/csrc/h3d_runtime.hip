// Device memory, copies, streams and the HBM bandwidth probe.
//
// The memory helpers stand where the reference calls omp_target_alloc /
// omp_target_memcpy / omp_target_free (OpenMP_Single_Source/1-openmp.c:
// 166-179, 214, 231-232) and hipMalloc / hipMemcpy in its HIP twin
// (HIP/1-hip.cpp:217-225).  Unlike the reference they return a status
// instead of exiting.
#include <string.h>

#include "h3d_internal.h"

extern "C" int h3d_device_count(int *count) {
  if (!count) return H3D_EARG;
  int n = 0;
  hipError_t e = hipGetDeviceCount(&n);
  if (e == hipErrorNoDevice) n = 0;
  else if (e != hipSuccess) return (int)e;
  *count = n;
  return H3D_OK;
}

static int set_dev(int device_id) {
  int n = 0;
  int rc = h3d_device_count(&n);
  if (rc != H3D_OK) return rc;
  if (device_id < 0 || device_id >= n) return H3D_ENODEV;
  H3D_TRY(hipSetDevice(device_id));
  return H3D_OK;
}

extern "C" int h3d_target_alloc(void **ptr, size_t bytes, int device_id) {
  if (!ptr) return H3D_EARG;
  *ptr = nullptr;
  int rc = set_dev(device_id);
  if (rc != H3D_OK) return rc;
  if (bytes == 0) return H3D_OK;
  H3D_TRY(hipMalloc(ptr, bytes));
  return H3D_OK;
}

extern "C" int h3d_target_free(void *ptr, int device_id) {
  int rc = set_dev(device_id);
  if (rc != H3D_OK) return rc;
  if (ptr) H3D_TRY(hipFree(ptr));
  return H3D_OK;
}

extern "C" int h3d_target_memcpy(void *dst, const void *src, size_t bytes,
                                 int dst_device, int src_device) {
  if ((!dst || !src) && bytes) return H3D_EARG;
  if (bytes == 0) return H3D_OK;
  hipMemcpyKind kind;
  if (dst_device < 0 && src_device < 0) {
    memcpy(dst, src, bytes);
    return H3D_OK;
  } else if (dst_device < 0) {
    kind = hipMemcpyDeviceToHost;
  } else if (src_device < 0) {
    kind = hipMemcpyHostToDevice;
  } else {
    kind = hipMemcpyDeviceToDevice;
  }
  int rc = set_dev(dst_device >= 0 ? dst_device : src_device);
  if (rc != H3D_OK) return rc;
  if (kind == hipMemcpyDeviceToDevice && dst_device != src_device) {
    H3D_TRY(hipMemcpyPeer(dst, dst_device, src, src_device, bytes));
    H3D_TRY(hipDeviceSynchronize());
    return H3D_OK;
  }
  H3D_TRY(hipMemcpy(dst, src, bytes, kind));
  // hipMemcpy may return before the copy has landed (device-to-device
  // copies, and the DMA leg of a staged copy from pageable memory); wait
  // for it, so that work the caller queues next on any stream (including
  // non-blocking ones, which do not order behind the null stream) sees the
  // copied data
  H3D_TRY(hipStreamSynchronize(nullptr));
  return H3D_OK;
}

extern "C" int h3d_device_synchronize(int device_id) {
  int rc = set_dev(device_id);
  if (rc != H3D_OK) return rc;
  H3D_TRY(hipDeviceSynchronize());
  return H3D_OK;
}

extern "C" int h3d_device_arch(int device_id, char *buf, int buflen) {
  if (!buf || buflen < 1) return H3D_EARG;
  int rc = set_dev(device_id);
  if (rc != H3D_OK) return rc;
  hipDeviceProp_t p;
  H3D_TRY(hipGetDeviceProperties(&p, device_id));
  strncpy(buf, p.gcnArchName, (size_t)buflen - 1);
  buf[buflen - 1] = 0;
  return H3D_OK;
}

extern "C" int h3d_stream_create(void **stream, int device_id) {
  if (!stream) return H3D_EARG;
  int rc = set_dev(device_id);
  if (rc != H3D_OK) return rc;
  hipStream_t s;
  H3D_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return H3D_OK;
}

extern "C" int h3d_stream_destroy(void *stream) {
  if (!stream) return H3D_EARG;
  H3D_TRY(hipStreamDestroy(h3d_stream(stream)));
  return H3D_OK;
}

extern "C" int h3d_stream_synchronize(void *stream) {
  H3D_TRY(hipStreamSynchronize(h3d_stream(stream)));
  return H3D_OK;
}

extern "C" const char *h3d_error_string(int code) {
  switch (code) {
    case H3D_OK: return "success";
    case H3D_EARG: return "invalid argument";
    case H3D_ELIMIT: return "supported limit exceeded";
    case H3D_ENODEV: return "no such device";
    case H3D_EHOST: return "host-side failure";
    default:
      if (code > 0) return hipGetErrorString((hipError_t)code);
      return "unknown error";
  }
}

// ---------------------------------------------------------------------------
// HBM bandwidth probe.  Arrays far larger than the 256 MB Infinity Cache;
// each workgroup streams one contiguous chunk with four 16-byte loads in
// flight per lane and nontemporal loads and stores (the fastest of the
// copy forms measured on MI355X).

namespace {

typedef double v2d __attribute__((ext_vector_type(2)));
constexpr int PROBE_UNROLL = 4;

__global__ void __launch_bounds__(256)
probe_copy(v2d *__restrict__ a, const v2d *__restrict__ b, size_t n) {
  const size_t per = (n + gridDim.x - 1) / gridDim.x;
  const size_t lo = (size_t)blockIdx.x * per;
  const size_t hi = lo + per < n ? lo + per : n;
  for (size_t t = lo + threadIdx.x; t < hi; t += PROBE_UNROLL * 256) {
    v2d x[PROBE_UNROLL];
#pragma unroll
    for (int u = 0; u < PROBE_UNROLL; ++u) {
      const size_t q = t + u * 256;
      x[u] = q < hi ? __builtin_nontemporal_load(&b[q]) : v2d{0, 0};
    }
#pragma unroll
    for (int u = 0; u < PROBE_UNROLL; ++u) {
      const size_t q = t + u * 256;
      if (q < hi) __builtin_nontemporal_store(x[u], &a[q]);
    }
  }
}

__global__ void __launch_bounds__(256)
probe_triad(v2d *__restrict__ a, const v2d *__restrict__ b,
            const v2d *__restrict__ c, double s, size_t n) {
  const size_t per = (n + gridDim.x - 1) / gridDim.x;
  const size_t lo = (size_t)blockIdx.x * per;
  const size_t hi = lo + per < n ? lo + per : n;
  for (size_t t = lo + threadIdx.x; t < hi; t += PROBE_UNROLL * 256) {
    v2d x[PROBE_UNROLL], y[PROBE_UNROLL];
#pragma unroll
    for (int u = 0; u < PROBE_UNROLL; ++u) {
      const size_t q = t + u * 256;
      x[u] = q < hi ? __builtin_nontemporal_load(&b[q]) : v2d{0, 0};
      y[u] = q < hi ? __builtin_nontemporal_load(&c[q]) : v2d{0, 0};
    }
#pragma unroll
    for (int u = 0; u < PROBE_UNROLL; ++u) {
      const size_t q = t + u * 256;
      if (q < hi) __builtin_nontemporal_store(x[u] + s * y[u], &a[q]);
    }
  }
}

__global__ void probe_fill(v2d *a, size_t n, double v) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride)
    a[t] = v2d{v, v};
}

}  // namespace

extern "C" int h3d_stream_probe(size_t bytes, int reps, int device_id,
                                double *copy_gbs, double *triad_gbs) {
  if (bytes < (1u << 20) || reps < 1) return H3D_EARG;
  int rc = set_dev(device_id);
  if (rc != H3D_OK) return rc;
  const size_t n = bytes / sizeof(v2d);
  v2d *a = nullptr, *b = nullptr, *c = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  double best_copy = 0, best_triad = 0;
  hipError_t e = hipMalloc(&a, n * sizeof(v2d));
  if (e == hipSuccess) e = hipMalloc(&b, n * sizeof(v2d));
  if (e == hipSuccess) e = hipMalloc(&c, n * sizeof(v2d));
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  int dev_cus = 256;
  if (e == hipSuccess) {
    hipDeviceProp_t p;
    if (hipGetDeviceProperties(&p, device_id) == hipSuccess) dev_cus = p.multiProcessorCount;
  }
  const unsigned blocks = (unsigned)(dev_cus * 32);
  if (e == hipSuccess) {
    hipLaunchKernelGGL(probe_fill, dim3(blocks), dim3(256), 0, 0, b, n, 1.0);
    hipLaunchKernelGGL(probe_fill, dim3(blocks), dim3(256), 0, 0, c, n, 2.0);
    e = hipGetLastError();
  }
  for (int r = 0; e == hipSuccess && r <= reps; ++r) {
    float ms = 0;
    (void)hipEventRecord(e0, 0);
    hipLaunchKernelGGL(probe_copy, dim3(blocks), dim3(256), 0, 0, a, b, n);
    (void)hipEventRecord(e1, 0);
    e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess && r > 0) {  // first launch is a warm-up
      double gbs = 2.0 * n * sizeof(v2d) / (ms * 1e-3) / 1e9;
      if (gbs > best_copy) best_copy = gbs;
    }
    if (e != hipSuccess) break;
    (void)hipEventRecord(e0, 0);
    hipLaunchKernelGGL(probe_triad, dim3(blocks), dim3(256), 0, 0, a, b, c, 3.0, n);
    (void)hipEventRecord(e1, 0);
    e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess && r > 0) {
      double gbs = 3.0 * n * sizeof(v2d) / (ms * 1e-3) / 1e9;
      if (gbs > best_triad) best_triad = gbs;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (a) (void)hipFree(a);
  if (b) (void)hipFree(b);
  if (c) (void)hipFree(c);
  if (e != hipSuccess) return (int)e;
  if (copy_gbs) *copy_gbs = best_copy;
  if (triad_gbs) *triad_gbs = best_triad;
  return H3D_OK;
}

// Host <-> device copy rates: the quantities of the reference's transfer
// benchmarks (Bandwidth_Measurements/AMD-MI250X/Synchronous-test-H-to-D-
// Measurements/hip-1.hip:56-92 and its D-to-H and asynchronous twins),
// timed the same way, with the host clock around a blocking copy after two
// warm-up copies.  Best of `reps`.  out[0] H2D GB/s and out[1] D2H GB/s
// for `bytes`; out[2] H2D and out[3] D2H microseconds of a 4 KiB copy
// (the fixed cost of one transfer).  pinned != 0 uses hipHostMalloc'ed
// host memory, else malloc'ed (pageable) memory.
#include <chrono>
#include <stdlib.h>

extern "C" int h3d_transfer_probe(size_t bytes, int reps, int pinned, int device_id,
                                  double *out) {
  if (bytes < 4096 || reps < 1 || !out) return H3D_EARG;
  int rc = set_dev(device_id);
  if (rc != H3D_OK) return rc;
  void *h = nullptr, *d = nullptr;
  hipError_t e = pinned ? hipHostMalloc(&h, bytes, hipHostMallocDefault)
                        : (h = malloc(bytes)) ? hipSuccess : hipErrorOutOfMemory;
  if (e == hipSuccess) e = hipMalloc(&d, bytes);
  if (e == hipSuccess) memset(h, 0x3c, bytes);
  using clk = std::chrono::steady_clock;
  auto best = [&](hipMemcpyKind kind, size_t nb, int n, double *sec) {
    void *dst = kind == hipMemcpyHostToDevice ? d : h;
    const void *src = kind == hipMemcpyHostToDevice ? h : d;
    for (int w = 0; w < 2 && e == hipSuccess; ++w) e = hipMemcpy(dst, src, nb, kind);
    double b = 1e30;
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
      const auto t0 = clk::now();
      for (int i = 0; i < n && e == hipSuccess; ++i) e = hipMemcpy(dst, src, nb, kind);
      const double s = std::chrono::duration<double>(clk::now() - t0).count() / n;
      if (s < b) b = s;
    }
    *sec = b;
  };
  double s[4] = {0, 0, 0, 0};
  best(hipMemcpyHostToDevice, bytes, 1, &s[0]);
  best(hipMemcpyDeviceToHost, bytes, 1, &s[1]);
  best(hipMemcpyHostToDevice, 4096, 100, &s[2]);
  best(hipMemcpyDeviceToHost, 4096, 100, &s[3]);
  if (d) (void)hipFree(d);
  if (h) {
    if (pinned) (void)hipHostFree(h);
    else free(h);
  }
  if (e != hipSuccess) return (int)e;
  out[0] = bytes / s[0] / 1e9;
  out[1] = bytes / s[1] / 1e9;
  out[2] = s[2] * 1e6;
  out[3] = s[3] * 1e6;
  return H3D_OK;
}

// One host<->device transfer measured as the reference's transfer
// benchmarks do (Bandwidth_Measurements/AMD-MI250X/Synchronous-test-H-to-D-
// Measurements/hip-1.hip:67-92 and its D-to-H twin; Asynchronous-test-H-to-
// D-Measurements/hip-1.hip:79-98): two untimed warm-up copies, then one
// timed copy -- the host clock around a blocking hipMemcpy (async = 0), or
// HIP events on a stream around hipMemcpyAsync (async = 1) -- then a check
// of the data that arrived.  The reference fills its buffer with 1.0 and
// spot-checks 1000 of the first 1000 elements; here every 64-bit word i
// carries its own value (a multiplicative hash of i), so a block landing
// at a wrong offset fails too, and the check covers every word up to
// 1 GiB and above that the first and last 4096 words plus 4096 spread
// through the buffer.  direction 0 = H2D (the device copy is read back to
// check it), 1 = D2H (the device buffer is filled by an H2D copy first).
// out[0] GB/s (1e9 B/s), out[1] seconds of the timed copy, out[2] words
// that did not arrive intact (0 when verified), out[3] words checked.
#include <vector>

static inline uint64_t xfer_word(size_t i) {
  return (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x243F6A8885A308D3ull;
}

extern "C" int h3d_transfer_bench(size_t bytes, int pinned, int async, int direction,
                                  int device_id, double *out) {
  if (bytes < 8 || bytes % 8 || !out || (direction != 0 && direction != 1) ||
      (async != 0 && async != 1))
    return H3D_EARG;
  int rc = set_dev(device_id);
  if (rc != H3D_OK) return rc;
  const size_t words = bytes / 8;
  uint64_t *h = nullptr, *d = nullptr;
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = pinned ? hipHostMalloc((void **)&h, bytes, hipHostMallocDefault)
                        : (h = (uint64_t *)malloc(bytes)) ? hipSuccess : hipErrorOutOfMemory;
  if (e == hipSuccess) e = hipMalloc((void **)&d, bytes);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) {
    for (size_t i = 0; i < words; ++i) h[i] = xfer_word(i);
    if (direction == 1) {  // device holds the pattern, host is cleared
      e = hipMemcpy(d, h, bytes, hipMemcpyHostToDevice);
      if (e == hipSuccess) memset(h, 0, bytes);
    }
  }
  void *dst = direction == 0 ? (void *)d : (void *)h;
  const void *src = direction == 0 ? (const void *)h : (const void *)d;
  const hipMemcpyKind kind = direction == 0 ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost;
  auto copy = [&]() {
    if (!async) return hipMemcpy(dst, src, bytes, kind);
    hipError_t r = hipMemcpyAsync(dst, src, bytes, kind, s);
    return r == hipSuccess ? hipStreamSynchronize(s) : r;
  };
  for (int w = 0; w < 2 && e == hipSuccess; ++w) e = copy();
  double sec = 0.0;
  if (e == hipSuccess && !async) {
    const auto t0 = std::chrono::steady_clock::now();
    e = hipMemcpy(dst, src, bytes, kind);
    sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  } else if (e == hipSuccess) {
    float ms = 0.f;
    e = hipEventRecord(e0, s);
    if (e == hipSuccess) e = hipMemcpyAsync(dst, src, bytes, kind, s);
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    sec = ms * 1e-3;
  }
  // the check: which words to look at, and where their values are
  std::vector<size_t> idx;
  const bool all = bytes <= (1ull << 30);
  if (e == hipSuccess && !all) {
    const size_t edge = words < 4096 ? words : 4096;
    for (size_t i = 0; i < edge; ++i) idx.push_back(i);
    for (size_t i = words - edge; i < words; ++i) idx.push_back(i);
    for (size_t i = 0; i < 4096; ++i) idx.push_back((size_t)(xfer_word(i) % words));
  }
  double bad = 0, checked = 0;
  if (e == hipSuccess && direction == 1) {  // the host buffer received the data
    if (all) {
      for (size_t i = 0; i < words; ++i) bad += h[i] != xfer_word(i);
      checked = (double)words;
    } else {
      for (size_t i : idx) bad += h[i] != xfer_word(i);
      checked = (double)idx.size();
    }
  } else if (e == hipSuccess) {  // read the device copy back
    if (all) {
      std::vector<uint64_t> back(words);
      e = hipMemcpy(back.data(), d, bytes, hipMemcpyDeviceToHost);
      for (size_t i = 0; e == hipSuccess && i < words; ++i) bad += back[i] != xfer_word(i);
      checked = (double)words;
    } else {
      for (size_t i : idx) {
        uint64_t v = 0;
        e = hipMemcpy(&v, d + i, 8, hipMemcpyDeviceToHost);
        if (e != hipSuccess) break;
        bad += v != xfer_word(i);
      }
      checked = (double)idx.size();
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s) (void)hipStreamDestroy(s);
  if (d) (void)hipFree(d);
  if (h) {
    if (pinned) (void)hipHostFree(h);
    else free(h);
  }
  if (e != hipSuccess) return (int)e;
  out[0] = sec > 0 ? bytes / sec / 1e9 : 0.0;
  out[1] = sec;
  out[2] = bad;
  out[3] = checked;
  return H3D_OK;
}
