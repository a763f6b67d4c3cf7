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

// Waves per SIMD the compiler must leave room for: 4 (two 8-wave
// workgroups per CU, 128 VGPRs) for the 4-step, 32-row tile with FP32
// kappa and the residual in registers, which otherwise takes 129 VGPRs and
// one workgroup per CU; with the bound it fits without spilling and ran
// 1020 against 704 GLUP/s at 1024^3 (the other shapes either fit already
// or would spill).
__host__ __device__ constexpr int pass_waves_per_eu(int KB, int RY, int W, int MODE, bool BF16) {
  return (KB == 4 && RY == 4 && W == 8 && MODE == 1 && !BF16) ? 4 : 1;
}

template <int KB, int RY, int W, int MODE, bool BF16>
__global__ void __launch_bounds__(W * 64, pass_waves_per_eu(KB, RY, W, MODE, BF16))
h3d_pass_f32_kernel(float *__restrict__ uout, const float *__restrict__ uin,
                    const float *__restrict__ kf,
                    const unsigned short *__restrict__ kb16,
                    unsigned short *__restrict__ cout,
                    const unsigned short *__restrict__ cin, int nx, int ny,
                    int nz, int sy, long long sz, float factor, int ks, int ke,
                    int zchunk, int tiles_x, int tiles_y) {
  extern __shared__ float lds[];
  // one (tile, z chunk) per workgroup, z chunk slowest: the tiles of a
  // chunk start together and march in lockstep, so the halo columns and
  // rows they share are served from L2
  const unsigned b = h3d_xcd_tile(blockIdx.x, gridDim.x);
  const int tx = (int)(b % (unsigned)tiles_x);
  const int ty = (int)((b / (unsigned)tiles_x) % (unsigned)tiles_y);
  const int tz = (int)(b / (unsigned)tiles_x / (unsigned)tiles_y);
  const int zb = ks + tz * zchunk;
  pass_f32_tile<KB, RY, W, MODE, BF16>(lds, uout, uin, kf, kb16, cout, cin, nx,
                                       ny, nz, sy, sz, factor, zb,
                                       min(ke, zb + zchunk - 1), tx, ty);
}

// Arguments of one FP32 pass launch.  With `occ` set, nothing is launched:
// the shape's resident workgroups per CU go to *occ and its tiles per z
// chunk to *tiles (what the z-chunk choice needs to fill whole rounds).
struct F32Launch {
  float *uo;
  const float *ui;
  const float *kf;
  const uint16_t *kb;
  uint16_t *co;
  const uint16_t *ci;
  const h3d_grid *g;
  float factor;
  int ks, ke;  // output planes
  int zchunk;
  hipStream_t s;
  bool bf16;  // kappa as BF16 bits (kb) rather than FP32 (kf)
  int *occ;
  long long *tiles;
};

template <int KB, int RY, int W, int MODE, bool BF16>
int launch_one(const F32Launch &a) {
  constexpr int OX = 64 - 2 * KB, OY = W * RY - 2 * KB;
  if (OY < 1) return H3D_EARG;
  constexpr size_t lds = sizeof(float) * (size_t)pass_lds_floats(KB, RY, W);
  if (lds > 160 * 1024) return H3D_ELIMIT;
  const h3d_grid *g = a.g;
  const int tiles_x = (g->nx - 2 + OX - 1) / OX;
  const int tiles_y = (g->ny - 2 + OY - 1) / OY;
  auto fn = h3d_pass_f32_kernel<KB, RY, W, MODE, BF16>;
  if (lds > 64 * 1024) {
    static std::atomic<unsigned long long> big_lds_done{0};
    const int big_rc = h3d_allow_big_lds((const void *)fn, &big_lds_done);
    if (big_rc != H3D_OK) return big_rc;
  }
  if (a.occ) {
    // resident workgroups per CU depend on the kernel alone: asked once
    static std::atomic<int> occ_known{0};
    int occ = occ_known.load(std::memory_order_relaxed);
    if (occ == 0) {
      H3D_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, fn, 64 * W, lds));
      occ_known.store(occ, std::memory_order_relaxed);
    }
    *a.occ = occ;
    if (a.tiles) *a.tiles = (long long)tiles_x * tiles_y;
    return H3D_OK;
  }
  const long long nb =
      (long long)tiles_x * tiles_y * ((a.ke - a.ks + 1 + a.zchunk - 1) / a.zchunk);
  if (nb > 0x7fffffffLL) return H3D_ELIMIT;
  hipLaunchKernelGGL(fn, dim3((unsigned)nb), dim3(64 * W), lds, a.s, a.uo, a.ui, a.kf,
                     (const unsigned short *)a.kb, (unsigned short *)a.co,
                     (const unsigned short *)a.ci, g->nx, g->ny, g->nz, g->sy,
                     g->sz, a.factor, a.ks, a.ke, a.zchunk, tiles_x, tiles_y);
  return h3d_launch_status();
}

template <int KB, int RY, int W>
int launch_kb_ry_w(const F32Launch &a, int mode) {
  const bool bf = a.bf16;
  if (mode == 0) return bf ? launch_one<KB, RY, W, 0, true>(a) : launch_one<KB, RY, W, 0, false>(a);
  if (mode == 1) return bf ? launch_one<KB, RY, W, 1, true>(a) : launch_one<KB, RY, W, 1, false>(a);
  return bf ? launch_one<KB, RY, W, 2, true>(a) : launch_one<KB, RY, W, 2, false>(a);
}

// Shapes (rows per lane, waves) offered per depth.  Tiles of 16 waves
// (1024 threads, one workgroup per CU) keep more of their rows at depths
// 2 and 3; rows 5 and 6 fit the register file only at depth 2 (no spill
// in the register-carry mode).  Depths 7 and 8 have room only for 8-wave
// tiles of 3 or 4 rows (the levels' state takes ~40 registers per row).
template <int KB>
int launch_kb(int ry, int waves, const F32Launch &a, int mode) {
  if constexpr (KB <= 6) {
    if (ry == 2 && waves == 4) return launch_kb_ry_w<KB, 2, 4>(a, mode);
    if (ry == 2 && waves == 8) return launch_kb_ry_w<KB, 2, 8>(a, mode);
    if (ry == 4 && waves == 4) return launch_kb_ry_w<KB, 4, 4>(a, mode);
  }
  if (ry == 4 && waves == 8) return launch_kb_ry_w<KB, 4, 8>(a, mode);
  if constexpr (KB <= 3) {
    if (ry == 2 && waves == 16) return launch_kb_ry_w<KB, 2, 16>(a, mode);
    if (ry == 3 && waves == 16) return launch_kb_ry_w<KB, 3, 16>(a, mode);
    if (ry == 4 && waves == 16) return launch_kb_ry_w<KB, 4, 16>(a, mode);
  }
  if constexpr (KB == 2) {
    if (ry == 5 && waves == 16) return launch_kb_ry_w<KB, 5, 16>(a, mode);
    if (ry == 6 && waves == 16) return launch_kb_ry_w<KB, 6, 16>(a, mode);
  }
  if constexpr (KB >= 7) {
    if (ry == 3 && waves == 8) return launch_kb_ry_w<KB, 3, 8>(a, mode);
  }
  return H3D_EARG;
}

int launch_pass(int kb, int ry, int waves, const F32Launch &a, int mode) {
  switch (kb) {
#define H3D_KB(K) \
    case K: return launch_kb<K>(ry, waves, a, mode);
    H3D_KB(2) H3D_KB(3) H3D_KB(4) H3D_KB(5) H3D_KB(6) H3D_KB(7) H3D_KB(8)
#undef H3D_KB
  }
  return H3D_ELIMIT;
}

// Single-step passes (KB = 1): wave-independent tiles of 64 x (waves * RY)
// cells marching z over planes ks..ke, one (tile, z chunk) per workgroup,
// XCD-aware order.
template <int RY, int MODE, bool BF16>
__global__ void __launch_bounds__(512)
h3d_march_f32(float *__restrict__ uout, const float *__restrict__ uin,
              const float *__restrict__ kf, const unsigned short *__restrict__ kb16,
              unsigned short *__restrict__ cout, const unsigned short *__restrict__ cin,
              int nx, int ny, int nz, int sy, long long sz, float factor, int ks,
              int ke, int zchunk, int tiles_x, int tiles_y) {
  const unsigned b = h3d_xcd_tile(blockIdx.x, gridDim.x);
  const int tx = (int)(b % (unsigned)tiles_x);
  const int ty = (int)((b / (unsigned)tiles_x) % (unsigned)tiles_y);
  const int tz = (int)(b / (unsigned)tiles_x / (unsigned)tiles_y);
  march_f32_tile<RY, MODE, BF16>(uout, uin, kf, kb16, cout, cin, nx, ny, nz, sy, sz,
                                 factor, ks, ke, zchunk, tx, ty, tz);
}

template <int RY, int MODE, bool BF16>
int launch_march_one(int waves, const F32Launch &a) {
  const h3d_grid *g = a.g;
  const int tiles_x = (g->nx - 2 + 63) / 64;
  const int tiles_y = (g->ny - 2 + RY * waves - 1) / (RY * waves);
  const long long nb =
      (long long)tiles_x * tiles_y * ((a.ke - a.ks + 1 + a.zchunk - 1) / a.zchunk);
  if (nb > 0x7fffffffLL) return H3D_ELIMIT;
  hipLaunchKernelGGL((h3d_march_f32<RY, MODE, BF16>), dim3((unsigned)nb), dim3(64 * waves),
                     0, a.s, a.uo, a.ui, a.kf, (const unsigned short *)a.kb,
                     (unsigned short *)a.co, (const unsigned short *)a.ci, g->nx, g->ny,
                     g->nz, g->sy, g->sz, a.factor, a.ks, a.ke, a.zchunk, tiles_x, tiles_y);
  return h3d_launch_status();
}

template <int RY>
int launch_march_ry(int waves, const F32Launch &a, int mode) {
#define H3D_M(M) \
  return a.bf16 ? launch_march_one<RY, M, true>(waves, a) : launch_march_one<RY, M, false>(waves, a)
  if (mode == 0) H3D_M(0);
  if (mode == 1) H3D_M(1);
  H3D_M(2);
#undef H3D_M
}

int launch_march(int ry, int waves, const F32Launch &a, int mode) {
  if (waves < 1 || waves > 8) return H3D_EARG;
  if (ry == 1) return launch_march_ry<1>(waves, a, mode);
  if (ry == 2) return launch_march_ry<2>(waves, a, mode);
  if (ry == 4) return launch_march_ry<4>(waves, a, mode);
  return H3D_EARG;
}

}  // namespace

// Launch shape of an FP32 pass of `kblock` steps when the caller names
// none: single steps on the wave-independent tile; fused passes on the
// 16-wave tiles where the depth has them (rows 5 at two steps, rows 4 at
// three), else rows 4 x waves 8; the z chunk from h3d_round_zchunk.
int h3d_default_tune_f32(const h3d_grid *g, int kblock, int mode, int bf16, h3d_tune *t) {
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
  t->rows = t->kblock == 2 ? 5 : t->kblock == 3 ? 4 : 4;
  t->waves = t->kblock <= 3 ? 16 : 8;
  int zc = 0, got = 0;
  const int rc = h3d_pass_f32_zchunks(g, t->kblock, t->rows, t->waves, mode, bf16, &zc, 1, &got);
  if (rc != H3D_OK) return rc;
  if (got < 1) return H3D_EARG;
  t->zchunk = zc;
  return H3D_OK;
}

// Shape geometry: resident workgroups per CU and tiles per z chunk.
int h3d_pass_f32_geometry(const h3d_grid *g, int kblock, int rows, int waves, int mode,
                          int bf16, int *wg_per_cu, long long *tiles) {
  F32Launch a{};
  a.g = g;
  a.bf16 = bf16 != 0;
  a.occ = wg_per_cu;
  a.tiles = tiles;
  return launch_pass(kblock, rows, waves, a, mode);
}

// The n z chunks (into zc[], best first) that minimise the modelled pass
// time of a shape: workgroups = tiles x chunks run in rounds of
// CUs x (workgroups per CU), each workgroup marches its chunk plus 2*kblock
// warm-up planes, so time ~ rounds x (chunk + 2*kblock).  A round left
// part-empty costs as much as a full one, which is why the chunk count is
// chosen to fill whole rounds.  *got receives how many were written.
int h3d_pass_f32_zchunks(const h3d_grid *g, int kblock, int rows, int waves, int mode,
                         int bf16, int *zc, int n, int *got) {
  int occ = 0;
  long long tiles = 0;
  const int rc = h3d_pass_f32_geometry(g, kblock, rows, waves, mode, bf16, &occ, &tiles);
  if (rc != H3D_OK) return rc;
  return h3d_round_zchunks(g->nz - 2, kblock, tiles, occ, zc, n, got);
}

extern "C" int h3d_pass_f32_range(float *u_out, const float *u_in,
                                  const float *kappa_f32, const uint16_t *kappa_bf16,
                                  uint16_t *comp_out, const uint16_t *comp_in,
                                  const h3d_grid *g, float factor, int nsteps, int mode,
                                  int k_start, int k_end, const h3d_tune *tune,
                                  void *stream) {
  if (!u_out || !u_in || u_out == u_in || !h3d_grid_ok(g)) return H3D_EARG;
  if (k_start < 1 || k_end > g->nz - 2) return H3D_EARG;
  if ((kappa_f32 == nullptr) == (kappa_bf16 == nullptr)) return H3D_EARG;
  if (mode < 0 || mode > 2) return H3D_EARG;
  if (mode == 2 && (!comp_out || !comp_in || comp_out == comp_in)) return H3D_EARG;
  if (nsteps < 1 || nsteps > 8) return H3D_ELIMIT;
  // the caller's non-zero tune fields, the rest from this depth's default
  // shape; a z chunk left open is the modelled best for the shape in use
  const int bf = kappa_bf16 != nullptr;
  h3d_tune t{0, 0, 0, nsteps};
  if (tune) {
    t.rows = tune->rows;
    t.waves = tune->waves;
    t.zchunk = tune->zchunk;
  }
  if (!t.rows || !t.waves) {
    h3d_tune d;
    const int dr = h3d_default_tune_f32(g, nsteps, mode, bf, &d);
    if (dr != H3D_OK) return dr;
    if (!t.zchunk && !t.rows && !t.waves) t.zchunk = d.zchunk;
    if (!t.rows) t.rows = d.rows;
    if (!t.waves) t.waves = d.waves;
  }
  if (!t.zchunk) {
    if (nsteps == 1) {
      h3d_tune d;
      h3d_default_tune_f32(g, 1, mode, bf, &d);
      t.zchunk = d.zchunk;
    } else {
      int got = 0;
      const int r = h3d_pass_f32_zchunks(g, nsteps, t.rows, t.waves, mode, bf, &t.zchunk, 1, &got);
      if (r != H3D_OK) return r;
      if (got < 1) return H3D_EARG;
    }
  }
  if (t.zchunk < 1) return H3D_EARG;
  if (k_end < k_start) return H3D_OK;  // empty range: nothing to do
  F32Launch a{u_out, u_in, kappa_f32, kappa_bf16, comp_out, comp_in, g, factor, k_start,
              k_end, t.zchunk, h3d_stream(stream), bf != 0, nullptr, nullptr};
  if (nsteps == 1) return launch_march(t.rows, t.waves, a, mode);
  return launch_pass(nsteps, t.rows, t.waves, a, mode);
}

extern "C" int h3d_pass_f32(float *u_out, const float *u_in,
                            const float *kappa_f32, const uint16_t *kappa_bf16,
                            uint16_t *comp_out, const uint16_t *comp_in,
                            const h3d_grid *g, float factor, int nsteps,
                            int mode, const h3d_tune *tune, void *stream) {
  if (!h3d_grid_ok(g)) return H3D_EARG;
  return h3d_pass_f32_range(u_out, u_in, kappa_f32, kappa_bf16, comp_out, comp_in, g, factor,
                            nsteps, mode, 1, g->nz - 2, tune, stream);
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
  if (kb > 8) return H3D_ELIMIT;
  const int full = steps / kb, rem = steps % kb;
  const int npass = full + (rem ? 1 : 0);
  // pass p: ping-pong by parity; the last pass of a run whose step count
  // is not a multiple of kb is shorter, on the caller's shape where that
  // depth has it, else on that depth's own default shape
  auto pass = [&](int p) {
    const bool even = (p % 2) == 0;
    const int ns = (p < full) ? kb : rem;
    float *uo = even ? u_b : u_a;
    const float *ui = even ? u_a : u_b;
    uint16_t *co = even ? c_b : c_a;
    const uint16_t *ci = even ? c_a : c_b;
    int r = h3d_pass_f32(uo, ui, kappa_f32, kappa_bf16, co, ci, g, factor, ns, mode, tune, s);
    if (r == H3D_EARG && p >= full)
      r = h3d_pass_f32(uo, ui, kappa_f32, kappa_bf16, co, ci, g, factor, ns, mode, nullptr, s);
    return r;
  };
  h3d_run_res R;
  H3D_TRY(hipEventCreate(&R.e0));
  H3D_TRY(hipEventCreate(&R.e1));
  int rc = H3D_OK;
  // graph: the run's passes captured once and launched as one graph (runs
  // of more than H3D_GRAPH_PASSES passes replay a graph of that many
  // full passes, the rest launched directly)
  int in_graph = 0, reps = 0;
  if (graph && npass >= 2) {
    in_graph = npass <= H3D_GRAPH_PASSES ? npass : H3D_GRAPH_PASSES;
    reps = npass <= H3D_GRAPH_PASSES ? 1 : full / H3D_GRAPH_PASSES;
    H3D_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int r = H3D_OK;
    for (int p = 0; r == H3D_OK && p < in_graph; ++p) r = pass(p);
    hipError_t ec = hipStreamEndCapture(s, &R.gr);
    if (r != H3D_OK) return r;
    H3D_TRY(ec);
    H3D_TRY(hipGraphInstantiate(&R.ge, R.gr, nullptr, nullptr, 0));
  }
  H3D_TRY(hipEventRecord(R.e0, s));
  int p = 0;
  for (int i = 0; i < reps && rc == H3D_OK; ++i, p += in_graph) {
    hipError_t el = hipGraphLaunch(R.ge, s);
    if (el != hipSuccess) rc = (int)el;
  }
  for (; rc == H3D_OK && p < npass; ++p) rc = pass(p);
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
