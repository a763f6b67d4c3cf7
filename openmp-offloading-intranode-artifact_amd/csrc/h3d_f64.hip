// FP64 explicit 7-point variable-coefficient heat update for gfx950.
//
// Computes, for every interior point of planes k_start..k_end,
//   u_new = (u_old + rhs) + factor * ((((((A - B) + C) - D) + E) - F)
//   A = (k[i+1] + k) * (u[i+1] - u)      B = (k + k[i-1]) * (u - u[i-1])
//   C, D: the same along j;               E, F: the same along k,
// the expression and association order of heat3D_variable_coeff.c:174-181
// (and 1-openmp.c:74-82).  This file is compiled with -ffp-contract=off so
// that no multiply-add is fused: the result is bit-identical to the
// reference's CPU and GPU builds, which use the same flag.
//
// Kernel design (one wavefront = one 64-point x-segment of RY rows):
//  - each lane owns one x position and RY consecutive y rows, and marches
//    along z over a chunk of planes;
//  - z neighbours stay in registers (planes k-1, k, k+1 per row);
//  - x neighbours come from the adjacent lanes through DPP wave shifts,
//    with lanes 0 and 63 loading the two x-halo points;
//  - y neighbours inside the RY rows are registers; the rows above and
//    below the strip are loaded directly (they are a neighbouring
//    wavefront's rows, so they hit in L1/L2);
//  - the next plane's loads are issued before the current plane's
//    arithmetic, so one plane of loads is always in flight per wave;
//  - no LDS and no barriers: waves are independent, so occupancy alone
//    hides latency;
//  - workgroups are remapped so that each XCD walks a contiguous range of
//    tiles (neighbouring tiles share halo rows in that XCD's L2).
// Every HBM byte of u_old and kappa is read once per step from the
// memory side (halo re-reads hit the caches), and u_new written once:
// 24 B per lattice update (32 B with rhs).
#include "h3d_internal.h"

namespace {

template <int RY, bool RHS>
__global__ void __launch_bounds__(1024)
h3d_march_f64(double *__restrict__ un, const double *__restrict__ uo,
              const double *__restrict__ rhs, const double *__restrict__ kap,
              int nx, int ny, int sy, long long sz, double factor,
              int k_start, int k_end, int zchunk, int tiles_x, int tiles_y) {
#pragma clang fp contract(off)
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int waves = blockDim.x >> 6;
  unsigned t = h3d_xcd_tile(blockIdx.x, gridDim.x);
  const int tx = (int)(t % (unsigned)tiles_x);
  t /= (unsigned)tiles_x;
  const int ty = (int)(t % (unsigned)tiles_y);
  const int tz = (int)(t / (unsigned)tiles_y);

  const int i = 1 + tx * 64 + lane;
  const int j0 = 1 + (ty * waves + wave) * RY;
  if (j0 > ny - 2) return;  // wave-uniform; the kernel has no barriers
  const int kb = k_start + tz * zchunk;
  if (kb > k_end) return;
  const int ke = min(k_end, kb + zchunk - 1);

  const bool lane_ld = i <= nx - 1;   // this lane's column exists
  const bool lane_cmp = i <= nx - 2;  // and is interior
  // x halo: lane 0 holds the point left of the segment, lane 63 the point
  // right of it (needed only when lane 63 itself is interior).
  const bool xh_ld = (lane == 0) || (lane == 63 && lane_cmp);
  const int xh_di = (lane == 0) ? -1 : 1;

  int off[RY];
  bool row_ld[RY], row_cmp[RY];
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int j = j0 + r;
    off[r] = j * sy + i;
    row_ld[r] = j <= ny - 1;
    row_cmp[r] = j <= ny - 2;
  }
  const int lo_off = (j0 - 1) * sy + i;
  const int hi_off = (j0 + RY) * sy + i;
  const bool hi_ld = j0 + RY <= ny - 1;

  auto ld = [&](const double *p, int o, bool pred) -> double {
    return pred ? p[o] : 0.0;
  };

  double um[RY], uc[RY], up[RY], km[RY], kc[RY], kq[RY];
  double uxh[RY], kxh[RY], ulo, uhi, klo, khi;
  double rc[RY];

  {
    const double *u0 = uo + (long long)(kb - 1) * sz;
    const double *k0 = kap + (long long)(kb - 1) * sz;
    const double *u1 = u0 + sz, *k1 = k0 + sz;
    const double *u2 = u1 + sz, *k2 = k1 + sz;
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const bool p = lane_ld && row_ld[r];
      um[r] = ld(u0, off[r], p);
      km[r] = ld(k0, off[r], p);
      uc[r] = ld(u1, off[r], p);
      kc[r] = ld(k1, off[r], p);
      up[r] = ld(u2, off[r], p);
      kq[r] = ld(k2, off[r], p);
      uxh[r] = ld(u1, off[r] + xh_di, xh_ld && row_cmp[r]);
      kxh[r] = ld(k1, off[r] + xh_di, xh_ld && row_cmp[r]);
      if (RHS) rc[r] = ld(rhs + (long long)kb * sz, off[r], p);
    }
    ulo = ld(u1, lo_off, lane_ld);
    klo = ld(k1, lo_off, lane_ld);
    uhi = ld(u1, hi_off, lane_ld && hi_ld);
    khi = ld(k1, hi_off, lane_ld && hi_ld);
  }

  for (int k = kb; k <= ke; ++k) {
    // Prefetch: halos of plane k+1 and rows of plane k+2.
    double nu[RY], nk[RY], nuxh[RY], nkxh[RY], nulo, nuhi, nklo, nkhi, nr[RY];
    const bool more = k < ke;
    {
      const double *u1 = uo + (long long)(k + 1) * sz;
      const double *k1 = kap + (long long)(k + 1) * sz;
      const double *u2 = u1 + sz, *k2 = k1 + sz;
#pragma unroll
      for (int r = 0; r < RY; ++r) {
        const bool p = more && lane_ld && row_ld[r];
        nu[r] = ld(u2, off[r], p);
        nk[r] = ld(k2, off[r], p);
        nuxh[r] = ld(u1, off[r] + xh_di, more && xh_ld && row_cmp[r]);
        nkxh[r] = ld(k1, off[r] + xh_di, more && xh_ld && row_cmp[r]);
        if (RHS) nr[r] = ld(rhs + (long long)(k + 1) * sz, off[r], p);
      }
      nulo = ld(u1, lo_off, more && lane_ld);
      nklo = ld(k1, lo_off, more && lane_ld);
      nuhi = ld(u1, hi_off, more && lane_ld && hi_ld);
      nkhi = ld(k1, hi_off, more && lane_ld && hi_ld);
    }

    double *out = un + (long long)k * sz;
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const double u = uc[r], kk = kc[r];
      const double uim = h3d_from_left(u, uxh[r]);
      const double uip = h3d_from_right(u, uxh[r]);
      const double kim = h3d_from_left(kk, kxh[r]);
      const double kip = h3d_from_right(kk, kxh[r]);
      const double ujm = (r == 0) ? ulo : uc[r - 1];
      const double ujp = (r == RY - 1) ? uhi : uc[r + 1];
      const double kjm = (r == 0) ? klo : kc[r - 1];
      const double kjp = (r == RY - 1) ? khi : kc[r + 1];
      const double s = (kip + kk) * (uip - u) - (kk + kim) * (u - uim) +
                       (kjp + kk) * (ujp - u) - (kk + kjm) * (u - ujm) +
                       (kq[r] + kk) * (up[r] - u) - (kk + km[r]) * (u - um[r]);
      double res;
      if (RHS) res = (u + rc[r]) + factor * s;
      else res = u + factor * s;
      if (lane_cmp && row_cmp[r]) out[off[r]] = res;
    }

#pragma unroll
    for (int r = 0; r < RY; ++r) {
      um[r] = uc[r];
      km[r] = kc[r];
      uc[r] = up[r];
      kc[r] = kq[r];
      up[r] = nu[r];
      kq[r] = nk[r];
      uxh[r] = nuxh[r];
      kxh[r] = nkxh[r];
      if (RHS) rc[r] = nr[r];
    }
    ulo = nulo;
    uhi = nuhi;
    klo = nklo;
    khi = nkhi;
  }
}

template <int RY>
int launch_march(double *un, const double *uo, const double *rhs,
                 const double *kap, const h3d_grid *g, double factor, int ks,
                 int ke, int waves, int zchunk, hipStream_t s) {
  const int tiles_x = (g->nx - 2 + 63) / 64;
  const int rows_wg = RY * waves;
  const int tiles_y = (g->ny - 2 + rows_wg - 1) / rows_wg;
  const int planes = ke - ks + 1;
  const int chunks = (planes + zchunk - 1) / zchunk;
  const long long nblocks = (long long)tiles_x * tiles_y * chunks;
  if (nblocks > 0x7fffffffLL) return H3D_ELIMIT;
  dim3 grid((unsigned)nblocks), block(64 * waves);
  if (rhs)
    hipLaunchKernelGGL((h3d_march_f64<RY, true>), grid, block, 0, s, un, uo,
                       rhs, kap, g->nx, g->ny, g->sy, g->sz, factor, ks, ke,
                       zchunk, tiles_x, tiles_y);
  else
    hipLaunchKernelGGL((h3d_march_f64<RY, false>), grid, block, 0, s, un, uo,
                       rhs, kap, g->nx, g->ny, g->sy, g->sz, factor, ks, ke,
                       zchunk, tiles_x, tiles_y);
  return h3d_launch_status();
}

bool tune_ok(const h3d_tune &t) {
  if (!(t.rows == 1 || t.rows == 2 || t.rows == 4 || t.rows == 8)) return false;
  if (t.waves < 1 || t.waves > 16) return false;
  if (t.zchunk < 1) return false;
  return true;
}

}  // namespace

int h3d_default_tune_f64(const h3d_grid *g, int k_start, int k_end,
                         h3d_tune *t) {
  t->rows = 2;
  t->waves = 4;
  t->kblock = 1;
  // Enough workgroups to fill 256 CUs several times over, while keeping
  // the re-read chunk edge planes (2 per chunk) a small fraction.
  const int planes = k_end - k_start + 1;
  const long long tiles = (long long)((g->nx - 2 + 63) / 64) *
                          ((g->ny - 2 + t->rows * t->waves - 1) /
                           (t->rows * t->waves));
  int zc = planes;
  while (zc > 32 && tiles * ((planes + zc - 1) / zc) < 4096) zc = (zc + 1) / 2;
  t->zchunk = zc;
  return H3D_OK;
}

extern "C" int h3d_step_f64(double *u_new, const double *u_old,
                            const double *rhs, const double *kappa,
                            const h3d_grid *g, double factor, int k_start,
                            int k_end, const h3d_tune *tune, void *stream) {
  if (!u_new || !u_old || !kappa || !h3d_grid_ok(g)) return H3D_EARG;
  if (u_new == u_old) return H3D_EARG;
  if (k_start < 1 || k_end > g->nz - 2) return H3D_EARG;
  if (k_end < k_start) return H3D_OK;  // empty range: nothing to do
  h3d_tune t;
  h3d_default_tune_f64(g, k_start, k_end, &t);
  if (tune) {
    if (tune->rows) t.rows = tune->rows;
    if (tune->waves) t.waves = tune->waves;
    if (tune->zchunk) t.zchunk = tune->zchunk;
  }
  if (!tune_ok(t)) return H3D_EARG;
  hipStream_t s = h3d_stream(stream);
  switch (t.rows) {
    case 1: return launch_march<1>(u_new, u_old, rhs, kappa, g, factor, k_start, k_end, t.waves, t.zchunk, s);
    case 2: return launch_march<2>(u_new, u_old, rhs, kappa, g, factor, k_start, k_end, t.waves, t.zchunk, s);
    case 4: return launch_march<4>(u_new, u_old, rhs, kappa, g, factor, k_start, k_end, t.waves, t.zchunk, s);
    default: return launch_march<8>(u_new, u_old, rhs, kappa, g, factor, k_start, k_end, t.waves, t.zchunk, s);
  }
}

// ---------------------------------------------------------------------------
// The reference kernel signature (dense layout, device's null stream).

static int dense_step(double *u_new, const double *u_old, const double *rhs,
                      const double *kappa, int nz, int ny, int nx,
                      double factor, int k_start, int k_end, int device_id) {
  if (nx < 3 || ny < 3 || nz < 3) return H3D_EARG;
  int ndev = 0;
  H3D_TRY(hipGetDeviceCount(&ndev));
  if (device_id < 0 || device_id >= ndev) return H3D_ENODEV;
  H3D_TRY(hipSetDevice(device_id));
  h3d_grid g{nx, ny, nz, nx, (long long)nx * ny};
  return h3d_step_f64(u_new, u_old, rhs, kappa, &g, factor, k_start, k_end,
                      nullptr, nullptr);
}

extern "C" int h3d_heat_update(double *u_new, const double *u_old,
                               const double *rhs, const double *kappa, int nz,
                               int ny, int nx, double factor, int device_id) {
  return dense_step(u_new, u_old, rhs, kappa, nz, ny, nx, factor, 1, nz - 2,
                    device_id);
}

extern "C" int h3d_heat_update_interior(double *u_new, const double *u_old,
                                        const double *rhs, const double *kappa,
                                        int nz, int ny, int nx, double factor,
                                        int k_start, int k_end,
                                        int device_id) {
  return dense_step(u_new, u_old, rhs, kappa, nz, ny, nx, factor, k_start,
                    k_end, device_id);
}

extern "C" int h3d_heat_update_boundary_bottom(double *u_new,
                                               const double *u_old,
                                               const double *rhs,
                                               const double *kappa, int nz,
                                               int ny, int nx, double factor,
                                               int device_id) {
  return dense_step(u_new, u_old, rhs, kappa, nz, ny, nx, factor, 1, 1,
                    device_id);
}

extern "C" int h3d_heat_update_boundary_top(double *u_new, const double *u_old,
                                            const double *rhs,
                                            const double *kappa, int nz,
                                            int ny, int nx, double factor,
                                            int device_id) {
  return dense_step(u_new, u_old, rhs, kappa, nz, ny, nx, factor, nz - 2,
                    nz - 2, device_id);
}

// ---------------------------------------------------------------------------
// Time loop with ping-pong buffers, optionally replayed from a hipGraph.

extern "C" int h3d_run_f64(double *u_a, double *u_b, const double *rhs,
                           const double *kappa, const h3d_grid *g,
                           double factor, int steps, const h3d_tune *tune,
                           int graph, void *stream, float *ms,
                           int *final_in_b) {
  if (!u_a || !u_b || !kappa || !h3d_grid_ok(g) || steps < 0) return H3D_EARG;
  hipStream_t s = h3d_stream(stream);
  if (graph && !s) return H3D_EARG;  // capture needs a created stream
  hipEvent_t e0, e1;
  H3D_TRY(hipEventCreate(&e0));
  H3D_TRY(hipEventCreate(&e1));
  int rc = H3D_OK;
  hipGraph_t gr = nullptr;
  hipGraphExec_t ge = nullptr;
  if (graph && steps >= 2) {
    H3D_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int r1 = h3d_step_f64(u_b, u_a, rhs, kappa, g, factor, 1, g->nz - 2, tune, s);
    int r2 = h3d_step_f64(u_a, u_b, rhs, kappa, g, factor, 1, g->nz - 2, tune, s);
    hipError_t ec = hipStreamEndCapture(s, &gr);
    if (r1 != H3D_OK) return r1;
    if (r2 != H3D_OK) return r2;
    H3D_TRY(ec);
    H3D_TRY(hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0));
  }
  H3D_TRY(hipEventRecord(e0, s));
  int done = 0;
  if (ge) {
    for (; done + 2 <= steps; done += 2) {
      hipError_t el = hipGraphLaunch(ge, s);
      if (el != hipSuccess) { rc = (int)el; break; }
    }
  }
  for (; rc == H3D_OK && done < steps; ++done) {
    const bool even = (done % 2) == 0;
    rc = h3d_step_f64(even ? u_b : u_a, even ? u_a : u_b, rhs, kappa, g,
                      factor, 1, g->nz - 2, tune, s);
  }
  if (rc == H3D_OK) {
    H3D_TRY(hipEventRecord(e1, s));
    H3D_TRY(hipEventSynchronize(e1));
    float t = 0.f;
    H3D_TRY(hipEventElapsedTime(&t, e0, e1));
    if (ms) *ms = t;
    if (final_in_b) *final_in_b = steps % 2;
  }
  if (ge) (void)hipGraphExecDestroy(ge);
  if (gr) (void)hipGraphDestroy(gr);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}
