// FP32-storage heat update with a compensated time step, temporally blocked
// on gfx950: one HBM pass advances the field KB time steps.
//
// Arithmetic of one cell update (compiled with -ffp-contract=off, so every
// operation rounds as written and the CPU oracle can replay it bit for bit;
// the one fused multiply-add is explicit):
//   S = ((((Fx+ - Fx-) + Fy+) - Fy-) + Fz+) - Fz-,  F = (k_a + k_b)(u_a - u_b)
//   plain:       u' = fma(factor, S, u)
//   compensated: y = fma(factor, S, e); t = u + y; e' = y - (t - u); u' = t
//                (Kahan)
// The six products are the reference's terms A..F (heat3D_variable_coeff.c:
// 174-181) in FP32.  A face flux computed for one cell is reused, unchanged,
// as the opposite face of its neighbour, which is the same expression with
// the same operands and therefore the same bits.
//
// Temporal blocking.  A workgroup of W wavefronts owns an xy tile of 64 x
// (W*RY) cells (lane = x, RY rows per lane, waves stacked in y) and marches
// z.  At march iteration `it` it loads plane `it` (level 0) and computes
// level l (time step l of the pass) at plane it-l for l = 1..KB, so every
// plane is read from HBM once and written once per KB steps.  Each level
// shrinks the valid region by one cell per side (trapezoid in x and y; the
// z direction is exact), so the tile writes (64-2KB) x (W*RY-2KB) cells and
// neighbouring tiles overlap by 2KB.
//  - u of level l-1 at the plane being updated and the plane above stay in
//    registers (two slots per level, alternated by unrolling the march by
//    two, so no register moves);
//  - the z flux below a cell is the flux above it from the previous march
//    iteration (one register per level and row);
//  - x neighbours are the adjacent lanes (DPP wave shifts);
//  - y neighbours are the lane's other rows, except across wavefronts, where
//    each wave publishes the two edge rows of every level in LDS (double
//    buffered, one barrier per march iteration);
//  - kappa is kept in an LDS ring of KB+1 planes, and the face sums
//    cx = k(i+1)+k(i), cy, cz are formed where they are used (the bits a
//    stored sum would have; a ring of the three sums would take 3*KB
//    planes); the rows next to a wave's rows come from the neighbouring
//    waves' slots, written an iteration earlier;
//  - loads are unconditional, with row/column/plane indices clamped into
//    the array: a cell outside it only ever feeds cells that are outside
//    the tile's valid region, and the per-lane in-plane offsets are 32-bit
//    constants, so a plane costs RY u loads and RY kappa loads and no
//    address arithmetic.
// Residual e of each cell is carried through the KB steps in registers; in
// COMP16 mode it is also read at the start of the pass and written at the
// end as FP16 scaled by the cell's ulp, 2^(23 - E(u)).
//
// Single-step passes (KB = 1) need no level sharing, so they run a
// separate wave-independent tile (march_f32_tile): 64 x (waves * RY) cells
// per workgroup with no overlap between tiles, no LDS and no barrier, the
// x/y halos loaded directly.  It computes the same bits as the fused tile
// would at KB = 1 and streams at the HBM rate (the fused tile at KB = 1
// re-reads its 2-cell tile overlaps and waits at a barrier per plane).
#include "h3d_tiles.h"

namespace {

template <int KB, int RY, int W, int MODE, bool BF16>
__global__ void __launch_bounds__(W * 64)
h3d_pass_f32_kernel(float *__restrict__ uout, const float *__restrict__ uin,
                    const float *__restrict__ kf,
                    const unsigned short *__restrict__ kb16,
                    unsigned short *__restrict__ cout,
                    const unsigned short *__restrict__ cin, int nx, int ny,
                    int nz, int sy, long long sz, float factor, int zchunk,
                    int tiles_x, int tiles_y) {
  extern __shared__ float lds[];
  // one (tile, z chunk) per workgroup, z chunk slowest: the tiles of a
  // chunk start together and march in lockstep, so the halo columns and
  // rows they share are served from L2
  const unsigned b = h3d_xcd_tile(blockIdx.x, gridDim.x);
  const int tx = (int)(b % (unsigned)tiles_x);
  const int ty = (int)((b / (unsigned)tiles_x) % (unsigned)tiles_y);
  const int tz = (int)(b / (unsigned)tiles_x / (unsigned)tiles_y);
  const int zb = 1 + tz * zchunk;
  pass_f32_tile<KB, RY, W, MODE, BF16>(lds, uout, uin, kf, kb16, cout, cin, nx,
                                       ny, nz, sy, sz, factor, zb,
                                       min(nz - 2, zb + zchunk - 1), tx, ty);
}

template <int KB, int RY, int W, int MODE, bool BF16>
int launch_one(float *uo, const float *ui, const float *kf, const uint16_t *kb,
               uint16_t *co, const uint16_t *ci, const h3d_grid *g,
               float factor, int zchunk, hipStream_t s) {
  constexpr int OX = 64 - 2 * KB, OY = W * RY - 2 * KB;
  if (OY < 1) return H3D_EARG;
  constexpr size_t lds = sizeof(float) * (size_t)pass_lds_floats(KB, RY, W);
  if (lds > 160 * 1024) return H3D_ELIMIT;
  const int tiles_x = (g->nx - 2 + OX - 1) / OX;
  const int tiles_y = (g->ny - 2 + OY - 1) / OY;
  auto fn = h3d_pass_f32_kernel<KB, RY, W, MODE, BF16>;
  if (lds > 64 * 1024) {
    static std::atomic<unsigned long long> big_lds_done{0};
    const int big_rc = h3d_allow_big_lds((const void *)fn, &big_lds_done);
    if (big_rc != H3D_OK) return big_rc;
  }
  const long long nb =
      (long long)tiles_x * tiles_y * ((g->nz - 2 + zchunk - 1) / zchunk);
  if (nb > 0x7fffffffLL) return H3D_ELIMIT;
  hipLaunchKernelGGL(fn, dim3((unsigned)nb), dim3(64 * W), lds, s, uo, ui, kf,
                     (const unsigned short *)kb, (unsigned short *)co,
                     (const unsigned short *)ci, g->nx, g->ny, g->nz, g->sy,
                     g->sz, factor, zchunk, tiles_x, tiles_y);
  return h3d_launch_status();
}

template <int KB, int RY, int W>
int launch_kb_ry_w(float *uo, const float *ui, const float *kf,
                   const uint16_t *kb, uint16_t *co, const uint16_t *ci,
                   const h3d_grid *g, float factor, int mode, int zchunk,
                   hipStream_t s) {
  const bool bf = kb != nullptr;
  if (mode == 0)
    return bf ? launch_one<KB, RY, W, 0, true>(uo, ui, kf, kb, co, ci, g, factor, zchunk, s)
              : launch_one<KB, RY, W, 0, false>(uo, ui, kf, kb, co, ci, g, factor, zchunk, s);
  if (mode == 1)
    return bf ? launch_one<KB, RY, W, 1, true>(uo, ui, kf, kb, co, ci, g, factor, zchunk, s)
              : launch_one<KB, RY, W, 1, false>(uo, ui, kf, kb, co, ci, g, factor, zchunk, s);
  return bf ? launch_one<KB, RY, W, 2, true>(uo, ui, kf, kb, co, ci, g, factor, zchunk, s)
            : launch_one<KB, RY, W, 2, false>(uo, ui, kf, kb, co, ci, g, factor, zchunk, s);
}

template <int KB>
int launch_kb(int ry, int waves, float *uo, const float *ui, const float *kf,
              const uint16_t *kb, uint16_t *co, const uint16_t *ci,
              const h3d_grid *g, float factor, int mode, int zchunk,
              hipStream_t s) {
  if (ry == 2 && waves == 4) return launch_kb_ry_w<KB, 2, 4>(uo, ui, kf, kb, co, ci, g, factor, mode, zchunk, s);
  if (ry == 2 && waves == 8) return launch_kb_ry_w<KB, 2, 8>(uo, ui, kf, kb, co, ci, g, factor, mode, zchunk, s);
  if (ry == 4 && waves == 4) return launch_kb_ry_w<KB, 4, 4>(uo, ui, kf, kb, co, ci, g, factor, mode, zchunk, s);
  if (ry == 4 && waves == 8) return launch_kb_ry_w<KB, 4, 8>(uo, ui, kf, kb, co, ci, g, factor, mode, zchunk, s);
  return H3D_EARG;
}

// Single-step passes (KB = 1): wave-independent tiles of 64 x (waves * RY)
// cells marching z, one (tile, z chunk) per workgroup, XCD-aware order.
template <int RY, int MODE, bool BF16>
__global__ void __launch_bounds__(512)
h3d_march_f32(float *__restrict__ uout, const float *__restrict__ uin,
              const float *__restrict__ kf, const unsigned short *__restrict__ kb16,
              unsigned short *__restrict__ cout, const unsigned short *__restrict__ cin,
              int nx, int ny, int nz, int sy, long long sz, float factor,
              int zchunk, int tiles_x, int tiles_y) {
  const unsigned b = h3d_xcd_tile(blockIdx.x, gridDim.x);
  const int tx = (int)(b % (unsigned)tiles_x);
  const int ty = (int)((b / (unsigned)tiles_x) % (unsigned)tiles_y);
  const int tz = (int)(b / (unsigned)tiles_x / (unsigned)tiles_y);
  march_f32_tile<RY, MODE, BF16>(uout, uin, kf, kb16, cout, cin, nx, ny, nz, sy, sz,
                                 factor, zchunk, tx, ty, tz);
}

template <int RY, int MODE, bool BF16>
int launch_march_one(int waves, float *uo, const float *ui, const float *kf,
                     const uint16_t *kb, uint16_t *co, const uint16_t *ci,
                     const h3d_grid *g, float factor, int zchunk, hipStream_t s) {
  const int tiles_x = (g->nx - 2 + 63) / 64;
  const int tiles_y = (g->ny - 2 + RY * waves - 1) / (RY * waves);
  const long long nb =
      (long long)tiles_x * tiles_y * ((g->nz - 2 + zchunk - 1) / zchunk);
  if (nb > 0x7fffffffLL) return H3D_ELIMIT;
  hipLaunchKernelGGL((h3d_march_f32<RY, MODE, BF16>), dim3((unsigned)nb), dim3(64 * waves),
                     0, s, uo, ui, kf, (const unsigned short *)kb, (unsigned short *)co,
                     (const unsigned short *)ci, g->nx, g->ny, g->nz, g->sy, g->sz, factor,
                     zchunk, tiles_x, tiles_y);
  return h3d_launch_status();
}

template <int RY>
int launch_march_ry(int waves, float *uo, const float *ui, const float *kf,
                    const uint16_t *kb, uint16_t *co, const uint16_t *ci,
                    const h3d_grid *g, float factor, int mode, int zchunk,
                    hipStream_t s) {
  const bool bf = kb != nullptr;
#define H3D_M(M) \
  return bf ? launch_march_one<RY, M, true>(waves, uo, ui, kf, kb, co, ci, g, factor, zchunk, s) \
            : launch_march_one<RY, M, false>(waves, uo, ui, kf, kb, co, ci, g, factor, zchunk, s)
  if (mode == 0) H3D_M(0);
  if (mode == 1) H3D_M(1);
  H3D_M(2);
#undef H3D_M
}

int launch_march(int ry, int waves, float *uo, const float *ui, const float *kf,
                 const uint16_t *kb, uint16_t *co, const uint16_t *ci,
                 const h3d_grid *g, float factor, int mode, int zchunk, hipStream_t s) {
  if (waves < 1 || waves > 8) return H3D_EARG;
  if (ry == 1) return launch_march_ry<1>(waves, uo, ui, kf, kb, co, ci, g, factor, mode, zchunk, s);
  if (ry == 2) return launch_march_ry<2>(waves, uo, ui, kf, kb, co, ci, g, factor, mode, zchunk, s);
  if (ry == 4) return launch_march_ry<4>(waves, uo, ui, kf, kb, co, ci, g, factor, mode, zchunk, s);
  return H3D_EARG;
}

}  // namespace

int h3d_default_tune_f32(const h3d_grid *g, int kblock, h3d_tune *t) {
  t->kblock = kblock > 0 ? kblock : 2;
  if (t->kblock == 1) {
    // single-step tiles: the FP64 step's rule (whole z when that gives
    // 2048 wavefronts, else shorter chunks)
    t->rows = 2;
    t->waves = 4;
    const int planes = g->nz - 2;
    const long long tiles = (long long)((g->nx - 2 + 63) / 64) *
                            ((g->ny - 2 + t->rows * t->waves - 1) / (t->rows * t->waves));
    int zc = planes;
    while (zc > 16 && tiles * t->waves * ((planes + zc - 1) / zc) < 2048) zc = (zc + 1) / 2;
    t->zchunk = zc;
    return H3D_OK;
  }
  t->rows = 2;
  t->waves = 8;
  const int OX = 64 - 2 * t->kblock, OY = t->rows * t->waves - 2 * t->kblock;
  const long long tiles = (long long)((g->nx - 2 + OX - 1) / OX) *
                          ((g->ny - 2 + (OY > 0 ? OY : 1) - 1) / (OY > 0 ? OY : 1));
  const int planes = g->nz - 2;
  int zc = planes;
  // enough workgroups for 256 CUs x 2, z warm-up (2*kblock planes) small
  while (zc > 8 * t->kblock && tiles * ((planes + zc - 1) / zc) < 1024) zc = (zc + 1) / 2;
  t->zchunk = zc;
  return H3D_OK;
}

extern "C" int h3d_pass_f32(float *u_out, const float *u_in,
                            const float *kappa_f32, const uint16_t *kappa_bf16,
                            uint16_t *comp_out, const uint16_t *comp_in,
                            const h3d_grid *g, float factor, int nsteps,
                            int mode, const h3d_tune *tune, void *stream) {
  if (!u_out || !u_in || u_out == u_in || !h3d_grid_ok(g)) return H3D_EARG;
  if ((kappa_f32 == nullptr) == (kappa_bf16 == nullptr)) return H3D_EARG;
  if (mode < 0 || mode > 2) return H3D_EARG;
  if (mode == 2 && (!comp_out || !comp_in || comp_out == comp_in)) return H3D_EARG;
  if (nsteps < 1 || nsteps > 6) return H3D_ELIMIT;
  h3d_tune t;
  h3d_default_tune_f32(g, nsteps, &t);
  if (tune) {
    if (tune->rows) t.rows = tune->rows;
    if (tune->waves) t.waves = tune->waves;
    if (tune->zchunk) t.zchunk = tune->zchunk;
  }
  if (t.zchunk < 1) return H3D_EARG;
  hipStream_t s = h3d_stream(stream);
  switch (nsteps) {
    case 1:
      return launch_march(t.rows, t.waves, u_out, u_in, kappa_f32, kappa_bf16, comp_out,
                          comp_in, g, factor, mode, t.zchunk, s);
#define H3D_KB(K) \
    case K: return launch_kb<K>(t.rows, t.waves, u_out, u_in, kappa_f32, kappa_bf16, comp_out, comp_in, g, factor, mode, t.zchunk, s);
    H3D_KB(2) H3D_KB(3) H3D_KB(4) H3D_KB(5) H3D_KB(6)
#undef H3D_KB
  }
  return H3D_EARG;
}

extern "C" int h3d_run_f32(float *u_a, float *u_b, const float *kappa_f32,
                           const uint16_t *kappa_bf16, uint16_t *c_a,
                           uint16_t *c_b, const h3d_grid *g, float factor,
                           int steps, int mode, const h3d_tune *tune, int graph,
                           void *stream, float *ms, int *final_in_b) {
  if (!u_a || !u_b || !h3d_grid_ok(g) || steps < 0) return H3D_EARG;
  hipStream_t s = h3d_stream(stream);
  if (graph && !s) return H3D_EARG;
  const int kb = (tune && tune->kblock > 0) ? tune->kblock : 2;
  if (kb > 6) return H3D_ELIMIT;
  h3d_run_res R;
  H3D_TRY(hipEventCreate(&R.e0));
  H3D_TRY(hipEventCreate(&R.e1));
  const int full = steps / kb, rem = steps % kb;
  int rc = H3D_OK;
  if (graph && full >= 2) {
    H3D_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int r1 = h3d_pass_f32(u_b, u_a, kappa_f32, kappa_bf16, c_b, c_a, g, factor, kb, mode, tune, s);
    int r2 = h3d_pass_f32(u_a, u_b, kappa_f32, kappa_bf16, c_a, c_b, g, factor, kb, mode, tune, s);
    hipError_t ec = hipStreamEndCapture(s, &R.gr);
    if (r1 != H3D_OK) return r1;
    if (r2 != H3D_OK) return r2;
    H3D_TRY(ec);
    H3D_TRY(hipGraphInstantiate(&R.ge, R.gr, nullptr, nullptr, 0));
  }
  H3D_TRY(hipEventRecord(R.e0, s));
  int pass = 0;
  if (R.ge) {
    for (; pass + 2 <= full; pass += 2) {
      hipError_t el = hipGraphLaunch(R.ge, s);
      if (el != hipSuccess) { rc = (int)el; break; }
    }
  }
  const int npass = full + (rem ? 1 : 0);
  for (; rc == H3D_OK && pass < npass; ++pass) {
    const bool even = (pass % 2) == 0;
    const int ns = (pass < full) ? kb : rem;
    rc = h3d_pass_f32(even ? u_b : u_a, even ? u_a : u_b, kappa_f32, kappa_bf16,
                      even ? c_b : c_a, even ? c_a : c_b, g, factor, ns, mode,
                      tune, s);
  }
  if (rc == H3D_OK) {
    H3D_TRY(hipEventRecord(R.e1, s));
    H3D_TRY(hipEventSynchronize(R.e1));
    float tms = 0.f;
    H3D_TRY(hipEventElapsedTime(&tms, R.e0, R.e1));
    if (ms) *ms = tms;
    if (final_in_b) *final_in_b = npass % 2;
  }
  return rc;
}
