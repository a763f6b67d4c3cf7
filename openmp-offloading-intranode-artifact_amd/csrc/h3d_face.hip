// Face pack/unpack: one whole XY plane k (x/y ring included) to/from a
// dense [j][i] buffer of nx*ny values; band pack/unpack: `count`
// consecutive planes to/from a dense [k][j][i] buffer (the exchange unit of
// slabs that hold several ghost planes per side).  This is the unit the reference's
// z-slab variants move between devices each step (2-openmp.c:321,
// 4-openmp-stream-omp-p2p.c:609-616: an XY plane is contiguous in the
// [k][j][i] layout, so the reference transfers it with one copy).  Here the
// plane may live in a pitched array, so the copy is a kernel that strips
// or restores the row padding; with sy == nx it degenerates to a plain
// contiguous copy done with 16-byte vector accesses.  Pure copies: the
// result is bit-identical to the source.
#include "h3d_internal.h"

namespace {

template <typename T>
__global__ void __launch_bounds__(256)
face_copy(T *__restrict__ dst, const T *__restrict__ src, int nx, int ny,
          int dst_sy, int src_sy, long long dst_sz, long long src_sz) {
  // blockIdx.z is the plane of a band, blockIdx.y walks rows, x threads
  // walk a row.
  const T *sp = src + (long long)blockIdx.z * src_sz;
  T *dp = dst + (long long)blockIdx.z * dst_sz;
  for (int j = blockIdx.y; j < ny; j += gridDim.y) {
    const T *s = sp + (long long)j * src_sy;
    T *d = dp + (long long)j * dst_sy;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < nx;
         i += gridDim.x * blockDim.x)
      d[i] = s[i];
  }
}

// Contiguous case: n bytes, 16-byte vectors plus a scalar tail.
__global__ void __launch_bounds__(256)
flat_copy(uint4 *__restrict__ dst, const uint4 *__restrict__ src, size_t nvec,
          unsigned char *__restrict__ dtail, const unsigned char *__restrict__ stail,
          int ntail) {
  size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t v = t; v < nvec; v += stride) dst[v] = src[v];
  if (t < (size_t)ntail) dtail[t] = stail[t];
}

// Planes k .. k+count-1 of u to/from a dense [k][j][i] buffer.
template <typename T>
int plane_copy(T *dst, const T *src, const h3d_grid *g, int k, int count, bool pack,
               hipStream_t s) {
  if (!dst || !src || !h3d_grid_ok(g) || k < 0 || count < 1 || count > 65535 ||
      k > g->nz - count)
    return H3D_EARG;
  const long long plane = (long long)k * g->sz;
  const long long face = (long long)g->nx * g->ny;
  T *d = pack ? dst : dst + plane;
  const T *sr = pack ? src + plane : src;
  if (g->sy == g->nx && (count == 1 || g->sz == face)) {
    const size_t bytes = (size_t)face * count * sizeof(T);
    const bool aligned = ((uintptr_t)d % 16 == 0) && ((uintptr_t)sr % 16 == 0);
    if (aligned) {
      const size_t nvec = bytes / 16;
      const int ntail = (int)(bytes % 16);
      const unsigned blocks = (unsigned)((nvec + 255) / 256 < 4096 ? (nvec + 255) / 256 : 4096);
      hipLaunchKernelGGL(flat_copy, dim3(blocks > 0 ? blocks : 1), dim3(256), 0,
                         s, reinterpret_cast<uint4 *>(d),
                         reinterpret_cast<const uint4 *>(sr), nvec,
                         reinterpret_cast<unsigned char *>(d) + nvec * 16,
                         reinterpret_cast<const unsigned char *>(sr) + nvec * 16,
                         ntail);
      return h3d_launch_status();
    }
  }
  const int dsy = pack ? g->nx : g->sy;
  const int ssy = pack ? g->sy : g->nx;
  const long long dsz = pack ? face : g->sz;
  const long long ssz = pack ? g->sz : face;
  dim3 grid((unsigned)((g->nx + 255) / 256), (unsigned)min(g->ny, 4096), (unsigned)count);
  hipLaunchKernelGGL((face_copy<T>), grid, dim3(256), 0, s, d, sr, g->nx,
                     g->ny, dsy, ssy, dsz, ssz);
  return h3d_launch_status();
}

}  // namespace

extern "C" int h3d_pack_face_f64(double *buf, const double *u,
                                 const h3d_grid *g, int k, void *stream) {
  return plane_copy<double>(buf, u, g, k, 1, true, h3d_stream(stream));
}
extern "C" int h3d_unpack_face_f64(double *u, const double *buf,
                                   const h3d_grid *g, int k, void *stream) {
  return plane_copy<double>(u, buf, g, k, 1, false, h3d_stream(stream));
}
extern "C" int h3d_pack_face_f32(float *buf, const float *u, const h3d_grid *g,
                                 int k, void *stream) {
  return plane_copy<float>(buf, u, g, k, 1, true, h3d_stream(stream));
}
extern "C" int h3d_unpack_face_f32(float *u, const float *buf,
                                   const h3d_grid *g, int k, void *stream) {
  return plane_copy<float>(u, buf, g, k, 1, false, h3d_stream(stream));
}
extern "C" int h3d_pack_band_f64(double *buf, const double *u, const h3d_grid *g,
                                 int k, int count, void *stream) {
  return plane_copy<double>(buf, u, g, k, count, true, h3d_stream(stream));
}
extern "C" int h3d_unpack_band_f64(double *u, const double *buf, const h3d_grid *g,
                                   int k, int count, void *stream) {
  return plane_copy<double>(u, buf, g, k, count, false, h3d_stream(stream));
}
extern "C" int h3d_pack_band_f32(float *buf, const float *u, const h3d_grid *g, int k,
                                 int count, void *stream) {
  return plane_copy<float>(buf, u, g, k, count, true, h3d_stream(stream));
}
extern "C" int h3d_unpack_band_f32(float *u, const float *buf, const h3d_grid *g, int k,
                                   int count, void *stream) {
  return plane_copy<float>(u, buf, g, k, count, false, h3d_stream(stream));
}
extern "C" int h3d_pack_band_u16(uint16_t *buf, const uint16_t *u, const h3d_grid *g, int k,
                                 int count, void *stream) {
  return plane_copy<uint16_t>(buf, u, g, k, count, true, h3d_stream(stream));
}
extern "C" int h3d_unpack_band_u16(uint16_t *u, const uint16_t *buf, const h3d_grid *g,
                                   int k, int count, void *stream) {
  return plane_copy<uint16_t>(u, buf, g, k, count, false, h3d_stream(stream));
}
