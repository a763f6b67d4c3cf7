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
//    tiles (neighbouring tiles share halo rows in that XCD's L2);
//  - when every row of u_new starts at the same offset within a 128-byte
//    line, the tiles' x origin is placed on a line boundary, so a wave
//    writes whole lines (rows of the reference's dense layout are 8-byte
//    shifted from the interior: n = 1024 runs 27% faster aligned).
// Every HBM byte of u_old and kappa is read once per step from the
// memory side (halo re-reads hit the caches), and u_new written once:
// 24 B per lattice update (32 B with rhs).
#include "h3d_tiles.h"

namespace {

template <int RY, bool RHS>
__global__ void __launch_bounds__(1024)
h3d_march_f64(double *__restrict__ un, const double *__restrict__ uo,
              const double *__restrict__ rhs, const double *__restrict__ kap,
              int nx, int ny, int sy, long long sz, double factor,
              int k_start, int k_end, int zchunk, int x0, int tiles_x,
              int tiles_y) {
  unsigned t = h3d_xcd_tile(blockIdx.x, gridDim.x);
  const int tx = (int)(t % (unsigned)tiles_x);
  t /= (unsigned)tiles_x;
  const int ty = (int)(t % (unsigned)tiles_y);
  const int tz = (int)(t / (unsigned)tiles_y);
  march_f64_tile<RY, RHS>(un, uo, rhs, kap, nx, ny, sy, sz, factor, k_start,
                          k_end, zchunk, x0, tx, ty, tz);
}

// x origin of the tiles (lane 0 of tile 0), at most 1 so that tile 0
// covers x = 1: the x whose element of every row of u_new starts a 128-byte
// line when the row stride is a whole number of lines, else 1.
int tile_x0(const double *un, const h3d_grid *g) {
  const uintptr_t a = (uintptr_t)un;
  if (a % 8 != 0 || ((size_t)g->sy * 8) % 128 != 0) return 1;
  const int x = (int)(((128 - a % 128) % 128) / 8);  // 0..15
  return x <= 1 ? x : x - 64;
}

template <int RY>
int launch_march(double *un, const double *uo, const double *rhs,
                 const double *kap, const h3d_grid *g, double factor, int ks,
                 int ke, int waves, int zchunk, hipStream_t s) {
  const int x0 = tile_x0(un, g);
  // tiles cover x0 .. x0 + 64*tiles_x - 1, which must reach nx - 2
  const int tiles_x = (g->nx - 1 - x0 + 63) / 64;
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
                       zchunk, x0, tiles_x, tiles_y);
  else
    hipLaunchKernelGGL((h3d_march_f64<RY, false>), grid, block, 0, s, un, uo,
                       rhs, kap, g->nx, g->ny, g->sy, g->sz, factor, ks, ke,
                       zchunk, x0, tiles_x, tiles_y);
  return h3d_launch_status();
}

// Temporally blocked FP64 passes: KB steps per HBM pass (pass_f64_tile;
// the FP32 pass's tiling, h3d_f32.hip), one (tile, z chunk) per workgroup
// with the chunk slowest and an XCD-aware order.
template <int KB, int RY, int W, bool RHS>
__global__ void __launch_bounds__(W * 64)
h3d_pass_f64_kernel(double *__restrict__ uout, const double *__restrict__ uin,
                    const double *__restrict__ rhs, const double *__restrict__ kf,
                    int nx, int ny, int nz, int sy, long long sz, double factor,
                    int ks, int ke, int zchunk, int tiles_x, int tiles_y) {
  extern __shared__ double lds64[];
  const unsigned b = h3d_xcd_tile(blockIdx.x, gridDim.x);
  const int tx = (int)(b % (unsigned)tiles_x);
  const int ty = (int)((b / (unsigned)tiles_x) % (unsigned)tiles_y);
  const int tz = (int)(b / (unsigned)tiles_x / (unsigned)tiles_y);
  const int zb = ks + tz * zchunk;
  pass_f64_tile<KB, RY, W, RHS>(lds64, uout, uin, rhs, kf, nx, ny, nz, sy, sz, factor,
                                zb, min(ke, zb + zchunk - 1), tx, ty);
}

// With `occ` set nothing is launched: the shape's resident workgroups per
// CU go to *occ and its tiles per z chunk to *tiles.
template <int KB, int RY, int W, bool RHS>
int launch_pass_one(double *uo, const double *ui, const double *rhs, const double *kf,
                    const h3d_grid *g, double factor, int ks, int ke, int zchunk,
                    hipStream_t s, int *occ, long long *tiles) {
  constexpr int OX = 64 - 2 * KB, OY = W * RY - 2 * KB;
  if (OY < 1) return H3D_EARG;
  constexpr size_t lds = sizeof(double) * (size_t)pass64_lds_doubles(KB, RY, W);
  if (lds > 160 * 1024) return H3D_ELIMIT;
  const int tiles_x = (g->nx - 2 + OX - 1) / OX;
  const int tiles_y = (g->ny - 2 + OY - 1) / OY;
  auto fn = h3d_pass_f64_kernel<KB, RY, W, RHS>;
  if (lds > 64 * 1024) {
    static std::atomic<unsigned long long> big_lds_done{0};
    const int big_rc = h3d_allow_big_lds((const void *)fn, &big_lds_done);
    if (big_rc != H3D_OK) return big_rc;
  }
  if (occ) {
    // resident workgroups per CU depend on the kernel alone: asked once
    static std::atomic<int> occ_known{0};
    int o = occ_known.load(std::memory_order_relaxed);
    if (o == 0) {
      H3D_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&o, fn, 64 * W, lds));
      occ_known.store(o, std::memory_order_relaxed);
    }
    *occ = o;
    if (tiles) *tiles = (long long)tiles_x * tiles_y;
    return H3D_OK;
  }
  const long long nb =
      (long long)tiles_x * tiles_y * ((ke - ks + 1 + zchunk - 1) / zchunk);
  if (nb > 0x7fffffffLL) return H3D_ELIMIT;
  hipLaunchKernelGGL(fn, dim3((unsigned)nb), dim3(64 * W), lds, s, uo, ui, rhs, kf,
                     g->nx, g->ny, g->nz, g->sy, g->sz, factor, ks, ke, zchunk,
                     tiles_x, tiles_y);
  return h3d_launch_status();
}

template <int KB>
int launch_pass_kb(int ry, int waves, double *uo, const double *ui, const double *rhs,
                   const double *kf, const h3d_grid *g, double factor, int ks, int ke,
                   int zchunk, hipStream_t s, int *occ = nullptr,
                   long long *tiles = nullptr) {
#define H3D_P(RY, W)                                                                 \
  if (ry == RY && waves == W)                                                       \
    return rhs ? launch_pass_one<KB, RY, W, true>(uo, ui, rhs, kf, g, factor, ks, ke, \
                                                  zchunk, s, occ, tiles)            \
               : launch_pass_one<KB, RY, W, false>(uo, ui, rhs, kf, g, factor, ks, ke, \
                                                   zchunk, s, occ, tiles);
  H3D_P(2, 4) H3D_P(2, 8) H3D_P(4, 4) H3D_P(4, 8)
  // 32 rows as 16 waves of 2: four waves per SIMD where rows 4 x 8 waves
  // run two (128 KB of LDS at 3 steps).  At 4 steps the tile takes the
  // whole 160 KB and spills past the 128 VGPRs of a 16-wave workgroup; it
  // measured 1430-1600 us per pass at 512^3 against 1060 us for rows 4 x
  // 8 waves, so it is not offered.
  if constexpr (KB <= 3) {
    H3D_P(2, 16)
  }
#undef H3D_P
  return H3D_EARG;
}

bool tune_ok(const h3d_tune &t) {
  if (!(t.rows == 1 || t.rows == 2 || t.rows == 4 || t.rows == 8)) return false;
  if (t.waves < 1 || t.waves > 16) return false;
  if (t.zchunk < 1) return false;
  return true;
}

}  // namespace

int h3d_pass_f64_geometry(const h3d_grid *g, int kblock, int rows, int waves, int rhs,
                          int *wg_per_cu, long long *tiles) {
  // a non-null stand-in for rhs selects the kernel that carries it; the
  // query reads no memory
  const double *r = rhs ? reinterpret_cast<const double *>(g) : nullptr;
  switch (kblock) {
    case 2: return launch_pass_kb<2>(rows, waves, nullptr, nullptr, r, nullptr, g, 0.0, 1, g->nz - 2, 1, nullptr, wg_per_cu, tiles);
    case 3: return launch_pass_kb<3>(rows, waves, nullptr, nullptr, r, nullptr, g, 0.0, 1, g->nz - 2, 1, nullptr, wg_per_cu, tiles);
    case 4: return launch_pass_kb<4>(rows, waves, nullptr, nullptr, r, nullptr, g, 0.0, 1, g->nz - 2, 1, nullptr, wg_per_cu, tiles);
  }
  return H3D_EARG;
}

int h3d_default_tune_f64(const h3d_grid *g, int k_start, int k_end,
                         h3d_tune *t) {
  t->rows = 2;
  t->waves = 2;
  t->kblock = 1;
  // March the whole z range when that already gives 2048 wavefronts (8 per
  // CU): measured on MI355X at 512^3, the re-read edge planes of shorter
  // chunks cost more than their extra parallelism.  Small grids split z.
  const int planes = k_end - k_start + 1;
  const long long tiles = (long long)((g->nx - 2 + 63) / 64) *
                          ((g->ny - 2 + t->rows * t->waves - 1) /
                           (t->rows * t->waves));
  int zc = planes;
  while (zc > 16 && tiles * t->waves * ((planes + zc - 1) / zc) < 2048)
    zc = (zc + 1) / 2;
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

// Default shape of a fused pass: 32-row tiles of 16 waves x 2 rows at 2
// and 3 steps (one 1024-thread workgroup per CU, the tuner's pick at
// 512^3), rows 4 x waves 8 at 4 steps (16 waves x 2 rows would need 208 KB
// of LDS); the z chunk that fills whole rounds of resident workgroups
// (h3d_round_zchunks), or short enough for 1024 workgroups if the shape's
// geometry cannot be asked.
static void default_tune_pass64(const h3d_grid *g, int kb, int planes, int rhs, h3d_tune *t) {
  // KB = 2: 32 rows as 16 waves of 2; KB = 3, 4: 32 rows as 8 waves of 4
  // (measured at 512^3: 3-step passes of 739 us as 8 x 4 against 768 us as
  // 16 x 2)
  t->rows = kb == 2 ? 2 : 4;
  t->waves = kb == 2 ? 16 : 8;
  t->kblock = kb;
  int occ = 0, zc = 0, got = 0;
  long long tiles = 0;
  if (h3d_pass_f64_geometry(g, kb, t->rows, t->waves, rhs, &occ, &tiles) == H3D_OK &&
      h3d_round_zchunks(planes, kb, tiles, occ, &zc, 1, &got) == H3D_OK && got == 1) {
    t->zchunk = zc;
    return;
  }
  const int OX = 64 - 2 * kb, OY = t->rows * t->waves - 2 * kb;
  tiles = (long long)((g->nx - 2 + OX - 1) / OX) * ((g->ny - 2 + OY - 1) / OY);
  zc = planes;
  while (zc > 8 * kb && tiles * ((planes + zc - 1) / zc) < 1024) zc = (zc + 1) / 2;
  t->zchunk = zc;
}

extern "C" int h3d_pass_f64(double *u_out, const double *u_in, const double *rhs,
                            const double *kappa, const h3d_grid *g, double factor,
                            int nsteps, int k_start, int k_end, const h3d_tune *tune,
                            void *stream) {
  if (!u_out || !u_in || !kappa || u_out == u_in || !h3d_grid_ok(g)) return H3D_EARG;
  if (k_start < 1 || k_end > g->nz - 2) return H3D_EARG;
  if (nsteps < 1 || nsteps > 4) return H3D_ELIMIT;
  if (nsteps == 1)
    return h3d_step_f64(u_out, u_in, rhs, kappa, g, factor, k_start, k_end, tune, stream);
  if (k_end < k_start) return H3D_OK;  // empty range: nothing to do
  h3d_tune t{0, 0, 0, nsteps};
  if (tune) {
    t.rows = tune->rows;
    t.waves = tune->waves;
    t.zchunk = tune->zchunk;
  }
  if (!t.rows || !t.waves || !t.zchunk) {
    // the caller's non-zero fields, the rest from the default shape (a z
    // chunk from the default rule for the shape in use)
    h3d_tune d;
    default_tune_pass64(g, nsteps, k_end - k_start + 1, rhs != nullptr, &d);
    if (!t.zchunk && (t.rows || t.waves)) {
      h3d_tune e{t.rows ? t.rows : d.rows, t.waves ? t.waves : d.waves, 0, nsteps};
      int occ = 0, zc = 0, got = 0;
      long long tiles = 0;
      if (h3d_pass_f64_geometry(g, nsteps, e.rows, e.waves, rhs != nullptr, &occ, &tiles) ==
              H3D_OK &&
          h3d_round_zchunks(k_end - k_start + 1, nsteps, tiles, occ, &zc, 1, &got) == H3D_OK &&
          got == 1)
        d.zchunk = zc;
    }
    if (!t.rows) t.rows = d.rows;
    if (!t.waves) t.waves = d.waves;
    if (!t.zchunk) t.zchunk = d.zchunk;
  }
  if (t.zchunk < 1) return H3D_EARG;
  hipStream_t s = h3d_stream(stream);
  switch (nsteps) {
    case 2: return launch_pass_kb<2>(t.rows, t.waves, u_out, u_in, rhs, kappa, g, factor, k_start, k_end, t.zchunk, s);
    case 3: return launch_pass_kb<3>(t.rows, t.waves, u_out, u_in, rhs, kappa, g, factor, k_start, k_end, t.zchunk, s);
    default: return launch_pass_kb<4>(t.rows, t.waves, u_out, u_in, rhs, kappa, g, factor, k_start, k_end, t.zchunk, s);
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

// The passes of a run of `steps` at pass depth kb: `full` passes of kb
// steps, the last `deep` of them one step deeper (shape *tdeep), then one
// pass of `rem` steps when rem > 0.  A remainder of r steps is one shorter
// pass, or, without rhs at depths 2 and 3, the last r passes one step
// deeper when the autotuner has timed both depths on this grid and r
// deeper passes cost clearly less than a pass: a pass moves the same bytes
// at any depth and costs about as much, so the deeper passes win when a
// (K+1)-step pass costs less than (1 + 1/r) K-step passes.  The tuner's
// figures are the best of back-to-back samples of one shape, and a deeper
// pass runs slower than that between K-step passes: on MI355X at 512^3
// the 4-step pass on 8 waves x 4 rows took 1058-1172 us interleaved with
// 3-step passes (720-745 us), and back to back it drifted from 1080 up to
// 1300-1420 us and down to 1020-1060 us over twenty launches.  So the rule
// asks for a saving of half a pass: a 1-step remainder rides on one deeper
// pass, a 2-step remainder at depth 3 takes the short pass.
static void run_plan_f64(const h3d_grid *g, bool has_rhs, int steps, int kb, int *full,
                         int *deep, int *rem, h3d_tune *tdeep) {
  *full = steps / kb;
  *rem = steps % kb;
  *deep = 0;
  if (has_rhs || !(kb == 2 || kb == 3) || *rem == 0 || *full < *rem) return;
  float t_kb = 0.f, t_deep = 0.f;
  if (h3d_tuned_f64(g, 0, kb, nullptr, &t_kb) == H3D_OK &&
      h3d_tuned_f64(g, 0, kb + 1, tdeep, &t_deep) == H3D_OK &&
      (float)*rem * (t_deep - t_kb) < 0.5f * t_kb) {
    *deep = *rem;
    *rem = 0;
  }
}

extern "C" int h3d_run_f64_passes(const h3d_grid *g, int has_rhs, int steps,
                                  const h3d_tune *tune, int *passes, int *deep_passes) {
  if (!h3d_grid_ok(g) || steps < 0 || !passes) return H3D_EARG;
  const int kb = (tune && tune->kblock > 1) ? tune->kblock : 1;
  if (kb > 4) return H3D_ELIMIT;
  int full = 0, deep = 0, rem = 0;
  h3d_tune tdeep{};
  run_plan_f64(g, has_rhs != 0, steps, kb, &full, &deep, &rem, &tdeep);
  *passes = full + (rem ? 1 : 0);
  if (deep_passes) *deep_passes = deep;
  return H3D_OK;
}

extern "C" int h3d_run_f64(double *u_a, double *u_b, const double *rhs,
                           const double *kappa, const h3d_grid *g,
                           double factor, int steps, const h3d_tune *tune,
                           int graph, void *stream, float *ms,
                           int *final_in_b) {
  if (!u_a || !u_b || !kappa || !h3d_grid_ok(g) || steps < 0) return H3D_EARG;
  hipStream_t s = h3d_stream(stream);
  if (graph && !s) return H3D_EARG;  // capture needs a created stream
  const int kb = (tune && tune->kblock > 1) ? tune->kblock : 1;
  if (kb > 4) return H3D_ELIMIT;
  int full = 0, deep = 0, rem = 0;
  h3d_tune tdeep{};
  run_plan_f64(g, rhs != nullptr, steps, kb, &full, &deep, &rem, &tdeep);
  const int npass = full + (rem ? 1 : 0);
  // a 1-step remainder of a fused run takes the single-step march on that
  // kernel's own tuned shape (its default when untuned), not the pass's
  // tile: at 512^3 the 4-step pass's 8 waves x 4 rows ran a single step
  // in 1.46-1.57 ms against 0.73 ms on the tuned 4 waves x 2 rows
  h3d_tune t1{};
  if (kb > 1 && h3d_tuned_f64(g, rhs ? 1 : 0, 1, &t1, nullptr) != H3D_OK) t1 = h3d_tune{};
  auto pass = [&](double *out, const double *in, int ns) {
    if (ns == 1)
      return h3d_step_f64(out, in, rhs, kappa, g, factor, 1, g->nz - 2, kb > 1 ? &t1 : tune, s);
    return h3d_pass_f64(out, in, rhs, kappa, g, factor, ns, 1, g->nz - 2,
                        ns == kb + 1 ? &tdeep : tune, s);
  };
  auto pass_p = [&](int p) {
    const bool even = (p % 2) == 0;
    const int ns = p < full - deep ? kb : (p < full ? kb + 1 : rem);
    return pass(even ? u_b : u_a, even ? u_a : u_b, ns);
  };
  h3d_run_res R;
  H3D_TRY(hipEventCreate(&R.e0));
  H3D_TRY(hipEventCreate(&R.e1));
  int rc = H3D_OK;
  // graph: the run's passes captured once and launched as one graph (runs
  // of more than H3D_GRAPH_PASSES passes replay a graph of that many
  // kblock-step passes, the rest launched directly)
  int in_graph = 0, reps = 0;
  if (graph && npass >= 2) {
    if (npass <= H3D_GRAPH_PASSES) {
      in_graph = npass;
      reps = 1;
    } else {
      reps = (full - deep) / H3D_GRAPH_PASSES;
      in_graph = reps ? H3D_GRAPH_PASSES : 0;
    }
  }
  if (in_graph) {
    H3D_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int r = H3D_OK;
    for (int p = 0; r == H3D_OK && p < in_graph; ++p) r = pass_p(p);
    hipError_t ec = hipStreamEndCapture(s, &R.gr);
    if (r != H3D_OK) return r;
    H3D_TRY(ec);
    H3D_TRY(hipGraphInstantiate(&R.ge, R.gr, nullptr, nullptr, 0));
  }
  H3D_TRY(hipEventRecord(R.e0, s));
  int done = 0;
  for (int i = 0; i < reps && rc == H3D_OK; ++i, done += in_graph) {
    hipError_t el = hipGraphLaunch(R.ge, s);
    if (el != hipSuccess) rc = (int)el;
  }
  for (; rc == H3D_OK && done < npass; ++done) rc = pass_p(done);
  if (rc == H3D_OK) {
    H3D_TRY(hipEventRecord(R.e1, s));
    H3D_TRY(hipEventSynchronize(R.e1));
    float t = 0.f;
    H3D_TRY(hipEventElapsedTime(&t, R.e0, R.e1));
    if (ms) *ms = t;
    if (final_in_b) *final_in_b = npass % 2;
  }
  return rc;
}
