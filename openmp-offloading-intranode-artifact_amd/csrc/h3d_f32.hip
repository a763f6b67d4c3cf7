// FP32-storage heat update with a compensated time step, temporally blocked
// on gfx950: one HBM pass advances the field KB time steps.
//
// Arithmetic of one cell update (compiled with -ffp-contract=off, so every
// operation rounds as written and the CPU oracle can replay it bit for bit):
//   S = ((((Fx+ - Fx-) + Fy+) - Fy-) + Fz+) - Fz-,  F = (k_a + k_b)(u_a - u_b)
//   d = factor * S
//   plain:       u' = u + d
//   compensated: y = d + e; t = u + y; e' = TwoSum error of (u, y); u' = t
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
//  - face coefficients cx = k(i+1)+k(i), cy, cz are computed once per plane
//    when kappa is loaded and kept in an LDS ring of KB+1 planes.
// Residual e of each cell is carried through the KB steps in registers; in
// COMP16 mode it is also read at the start of the pass and written at the
// end as FP16 scaled by the cell's ulp, 2^(23 - E(u)).
#include "h3d_internal.h"

namespace {

// ulp exponent of a stored value: max(ilogb(u), -126), -126 for 0.
__device__ __forceinline__ int ulp_exp(float u) {
  const unsigned b = __float_as_uint(u) & 0x7fffffffu;
  const int e = (int)(b >> 23) - 127;
  if (b == 0u || e == 128) return -126;  // zero or non-finite
  if (e == -127) {                      // subnormal: exponent of leading bit
    return -126;                        // clamped to -126 like ilogb >= -126
  }
  return e;
}

__device__ __forceinline__ unsigned short comp_encode(float e, float u) {
  const float s = __builtin_ldexpf(e, 23 - ulp_exp(u));
  const _Float16 h = (_Float16)s;  // round to nearest even
  return __builtin_bit_cast(unsigned short, h);
}

__device__ __forceinline__ float comp_decode(unsigned short hb, float u) {
  const _Float16 h = __builtin_bit_cast(_Float16, hb);
  return __builtin_ldexpf((float)h, ulp_exp(u) - 23);
}

__device__ __forceinline__ float bf16_to_f32(unsigned short b) {
  return __uint_as_float((unsigned)b << 16);
}

template <int KB, int RY, int MODE, bool BF16>
struct Pass {
  // ---- per-lane state ----
  float A[KB][2][RY];    // u of level l: slot P = plane it-l-1, 1-P = it-l
  float E[KB][2][RY];    // residual of level l, same slots
  float FZ[KB + 1][RY];  // z flux above the last plane of level l
  float kprev[2][RY];    // kappa rows of the previous loaded plane
  float pu[2][RY];       // prefetched u rows (slot per phase)
  float pk[2][RY + 1];   // prefetched kappa rows + the row above
  unsigned short pc[2][RY];  // prefetched residual bits (COMP16)

  // ---- geometry ----
  const float *uin;
  const float *kf;
  const unsigned short *kb16;
  const unsigned short *cin;
  float *uout;
  unsigned short *cout;
  float factor;
  int nx, ny, nz, sy;
  long long sz;
  int lane, w, W, TY;
  int i, j0;          // global x of this lane, global y of row 0
  int zb, ze;         // output planes of this workgroup
  bool lane_ld, lane_int, lane_out;
  bool row_ld[RY + 1], row_int[RY], row_out[RY];
  float *cx, *cy, *cz;  // LDS coefficient ring [KB+1][TY][64]
  float *ue;            // LDS edge rows [2][KB][W][2][64]

  __device__ __forceinline__ int slot(int p) const {
    int s = p % (KB + 1);
    return s < 0 ? s + KB + 1 : s;
  }
  __device__ __forceinline__ int cidx(int p, int trow) const {
    return (slot(p) * TY + trow) * 64 + lane;
  }
  __device__ __forceinline__ int eidx(int ph, int l, int wave, int hi) const {
    return (((ph * KB + l) * W + wave) * 2 + hi) * 64 + lane;
  }

  __device__ __forceinline__ void load_plane(int p, int ph) {
    const bool pin = p >= 0 && p <= nz - 1;
    const long long base = (long long)p * sz;
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const bool ok = pin && lane_ld && row_ld[r];
      const long long o = base + (long long)(j0 + r) * sy + i;
      pu[ph][r] = ok ? uin[o] : 0.0f;
      if (BF16) pk[ph][r] = ok ? bf16_to_f32(kb16[o]) : 0.0f;
      else pk[ph][r] = ok ? kf[o] : 0.0f;
      if (MODE == 2) pc[ph][r] = ok ? cin[o] : (unsigned short)0;
    }
    const bool ok = pin && lane_ld && row_ld[RY];
    const long long o = base + (long long)(j0 + RY) * sy + i;
    if (BF16) pk[ph][RY] = ok ? bf16_to_f32(kb16[o]) : 0.0f;
    else pk[ph][RY] = ok ? kf[o] : 0.0f;
  }

  // One march iteration at phase P (P alternates; slot roles swap).
  template <int P>
  __device__ __forceinline__ void iter(int it, int it_last) {
#pragma clang fp contract(off)
    // 1. Face coefficients of plane it (cx, cy) and cz of plane it-1.
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const float k = pk[P][r];
      const float kr = h3d_from_right(k, 0.0f);
      const float ku = pk[P][r + 1];
      const int tr = w * RY + r;
      cx[cidx(it, tr)] = kr + k;
      cy[cidx(it, tr)] = ku + k;
      cz[cidx(it - 1, tr)] = kprev[1 - P][r] + k;
      kprev[P][r] = k;
    }
    // 2. Level 0: the loaded plane becomes the newest slot of level 0.
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      A[0][1 - P][r] = pu[P][r];
      if (MODE == 2) E[0][1 - P][r] = comp_decode(pc[P][r], pu[P][r]);
      else E[0][1 - P][r] = 0.0f;
    }
    // 3. Prefetch the next plane into the other phase's buffers.
    if (it + 1 <= it_last) load_plane(it + 1, 1 - P);

    // 4. Levels 1..KB at planes it-1 .. it-KB.
    const int wb = w > 0 ? w - 1 : 0;
    const int wa = w < W - 1 ? w + 1 : W - 1;
    const int trb = w > 0 ? w * RY - 1 : 0;  // tile row below the strip
#pragma unroll
    for (int l = 1; l <= KB; ++l) {
      const int q = it - l;
      const bool q_int = q >= 1 && q <= nz - 2;
      const bool q_out = (l == KB) && q >= zb && q <= ze;
      const float ub = ue[eidx(1 - P, l - 1, wb, 1)];  // row below (mid plane)
      const float ua = ue[eidx(1 - P, l - 1, wa, 0)];  // row above
      const float cyb = cy[cidx(q, trb)];
      float fy_prev = 0.0f;
#pragma unroll
      for (int r = 0; r < RY; ++r) {
        const float u = A[l - 1][P][r];
        const float un = A[l - 1][1 - P][r];
        const int tr = w * RY + r;
        const float ccx = cx[cidx(q, tr)];
        const float ccy = cy[cidx(q, tr)];
        const float ccz = cz[cidx(q, tr)];
        const float uip = h3d_from_right(u, 0.0f);
        const float fxp = ccx * (uip - u);
        const float fxm = h3d_from_left(fxp, 0.0f);
        const float ujp = (r < RY - 1) ? A[l - 1][P][r + 1] : ua;
        const float fyp = ccy * (ujp - u);
        const float fym = (r > 0) ? fy_prev : cyb * (u - ub);
        fy_prev = fyp;
        const float fzp = ccz * (un - u);
        const float fzm = FZ[l][r];
        FZ[l][r] = fzp;
        const float s = ((((fxp - fxm) + fyp) - fym) + fzp) - fzm;
        const float d = factor * s;
        const bool inside = lane_int && row_int[r] && q_int;
        float res, e2 = 0.0f;
        if (MODE == 0) {
          res = u + d;
        } else {
          const float y = d + E[l - 1][P][r];
          const float t = u + y;
          const float bp = t - u;
          const float ap = t - bp;
          e2 = (u - ap) + (y - bp);
          res = t;
        }
        res = inside ? res : u;
        e2 = inside ? e2 : 0.0f;
        if (l < KB) {
          A[l][1 - P][r] = res;
          E[l][1 - P][r] = e2;
        } else if (q_out && lane_out && row_out[r]) {
          const long long o = (long long)q * sz + (long long)(j0 + r) * sy + i;
          uout[o] = res;
          if (MODE == 2) cout[o] = comp_encode(e2, res);
        }
      }
    }
    // 5. Publish this iteration's newest edge rows of levels 0..KB-1.
#pragma unroll
    for (int l = 0; l < KB; ++l) {
      ue[eidx(P, l, w, 0)] = A[l][1 - P][0];
      ue[eidx(P, l, w, 1)] = A[l][1 - P][RY - 1];
    }
    __syncthreads();
  }
};

template <int KB, int RY, int MODE, bool BF16>
__global__ void __launch_bounds__(512)
h3d_pass_f32_kernel(float *__restrict__ uout, const float *__restrict__ uin,
                    const float *__restrict__ kf,
                    const unsigned short *__restrict__ kb16,
                    unsigned short *__restrict__ cout,
                    const unsigned short *__restrict__ cin, int nx, int ny,
                    int nz, int sy, long long sz, float factor, int zchunk,
                    int tiles_x, int tiles_y) {
  extern __shared__ float lds[];
  Pass<KB, RY, MODE, BF16> st;
  st.uin = uin; st.kf = kf; st.kb16 = kb16; st.cin = cin;
  st.uout = uout; st.cout = cout; st.factor = factor;
  st.nx = nx; st.ny = ny; st.nz = nz; st.sy = sy; st.sz = sz;
  st.lane = threadIdx.x & 63;
  st.w = threadIdx.x >> 6;
  st.W = blockDim.x >> 6;
  st.TY = st.W * RY;

  unsigned t = h3d_xcd_tile(blockIdx.x, gridDim.x);
  const int tx = (int)(t % (unsigned)tiles_x);
  t /= (unsigned)tiles_x;
  const int ty = (int)(t % (unsigned)tiles_y);
  const int tz = (int)(t / (unsigned)tiles_y);
  const int OX = 64 - 2 * KB, OY = st.TY - 2 * KB;

  st.i = 1 - KB + tx * OX + st.lane;
  const int y0 = 1 - KB + ty * OY;  // global y of tile row 0
  st.j0 = y0 + st.w * RY;
  st.zb = 1 + tz * zchunk;
  st.ze = min(nz - 2, st.zb + zchunk - 1);

  st.lane_ld = st.i >= 0 && st.i <= nx - 1;
  st.lane_int = st.i >= 1 && st.i <= nx - 2;
  st.lane_out = st.lane >= KB && st.lane <= 63 - KB && st.i <= nx - 2;
#pragma unroll
  for (int r = 0; r <= RY; ++r) {
    const int j = st.j0 + r;
    st.row_ld[r] = j >= 0 && j <= ny - 1;
  }
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int j = st.j0 + r, tr = st.w * RY + r;
    st.row_int[r] = j >= 1 && j <= ny - 2;
    st.row_out[r] = tr >= KB && tr <= st.TY - 1 - KB && j <= ny - 2;
  }
  const int ring = (KB + 1) * st.TY * 64;
  st.cx = lds;
  st.cy = lds + ring;
  st.cz = lds + 2 * ring;
  st.ue = lds + 3 * ring;

#pragma unroll
  for (int l = 0; l <= KB; ++l)
#pragma unroll
    for (int r = 0; r < RY; ++r) st.FZ[l][r] = 0.0f;
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    st.kprev[0][r] = st.kprev[1][r] = 0.0f;
#pragma unroll
    for (int l = 0; l < KB; ++l)
      st.A[l][0][r] = st.A[l][1][r] = st.E[l][0][r] = st.E[l][1][r] = 0.0f;
  }
  // Edge-row buffers start defined (read before first written at warm-up).
  for (int x = threadIdx.x; x < 2 * KB * st.W * 2 * 64; x += blockDim.x) st.ue[x] = 0.0f;

  const int it0 = st.zb - KB, it1 = st.ze + KB;
  st.load_plane(it0, 0);
  __syncthreads();
  for (int it = it0; it <= it1; it += 2) {
    st.template iter<0>(it, it1);
    if (it + 1 <= it1) st.template iter<1>(it + 1, it1);
  }
}

struct F32Launch {
  int kb, ry, waves, zchunk;
};

template <int KB, int RY>
int launch_kb_ry(float *uo, const float *ui, const float *kf,
                 const uint16_t *kb, uint16_t *co, const uint16_t *ci,
                 const h3d_grid *g, float factor, int mode, int waves,
                 int zchunk, hipStream_t s) {
  const int TY = waves * RY;
  const int OX = 64 - 2 * KB, OY = TY - 2 * KB;
  if (OY < 1) return H3D_EARG;
  const int tiles_x = (g->nx - 2 + OX - 1) / OX;
  const int tiles_y = (g->ny - 2 + OY - 1) / OY;
  const int chunks = (g->nz - 2 + zchunk - 1) / zchunk;
  const long long nb = (long long)tiles_x * tiles_y * chunks;
  if (nb > 0x7fffffffLL) return H3D_ELIMIT;
  const size_t lds = sizeof(float) * ((size_t)3 * (KB + 1) * TY * 64 +
                                      (size_t)2 * KB * waves * 2 * 64);
  if (lds > 160 * 1024) return H3D_ELIMIT;
  dim3 grid((unsigned)nb), block(64 * waves);
  const bool bf = kb != nullptr;
  // Dynamic LDS above 64 KB must be allowed per kernel before launch.
#define H3D_LAUNCH(M, B)                                                      \
  do {                                                                        \
    auto fn = h3d_pass_f32_kernel<KB, RY, M, B>;                              \
    static bool big_lds_ok = false;                                           \
    if (lds > 64 * 1024 && !big_lds_ok) {                                     \
      H3D_TRY(hipFuncSetAttribute((const void *)fn,                           \
                                  hipFuncAttributeMaxDynamicSharedMemorySize, \
                                  160 * 1024));                               \
      big_lds_ok = true;                                                      \
    }                                                                         \
    hipLaunchKernelGGL(fn, grid, block, lds, s, uo, ui, kf, kb, co, ci,       \
                       g->nx, g->ny, g->nz, g->sy, g->sz, factor, zchunk,     \
                       tiles_x, tiles_y);                                     \
  } while (0)
  if (mode == 0) { if (bf) H3D_LAUNCH(0, true); else H3D_LAUNCH(0, false); }
  else if (mode == 1) { if (bf) H3D_LAUNCH(1, true); else H3D_LAUNCH(1, false); }
  else { if (bf) H3D_LAUNCH(2, true); else H3D_LAUNCH(2, false); }
#undef H3D_LAUNCH
  return h3d_launch_status();
}

template <int KB>
int launch_kb(int ry, float *uo, const float *ui, const float *kf,
              const uint16_t *kb, uint16_t *co, const uint16_t *ci,
              const h3d_grid *g, float factor, int mode, int waves, int zchunk,
              hipStream_t s) {
  switch (ry) {
    case 2: return launch_kb_ry<KB, 2>(uo, ui, kf, kb, co, ci, g, factor, mode, waves, zchunk, s);
    case 4: return launch_kb_ry<KB, 4>(uo, ui, kf, kb, co, ci, g, factor, mode, waves, zchunk, s);
    default: return H3D_EARG;
  }
}

}  // namespace

int h3d_default_tune_f32(const h3d_grid *g, int kblock, h3d_tune *t) {
  t->kblock = kblock > 0 ? kblock : 2;
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
  if (nsteps < 1 || nsteps > 8) return H3D_ELIMIT;
  h3d_tune t;
  h3d_default_tune_f32(g, nsteps, &t);
  if (tune) {
    if (tune->rows) t.rows = tune->rows;
    if (tune->waves) t.waves = tune->waves;
    if (tune->zchunk) t.zchunk = tune->zchunk;
  }
  if (t.waves < 1 || t.waves > 8 || t.zchunk < 1) return H3D_EARG;
  hipStream_t s = h3d_stream(stream);
  switch (nsteps) {
#define H3D_KB(K) \
    case K: return launch_kb<K>(t.rows, u_out, u_in, kappa_f32, kappa_bf16, comp_out, comp_in, g, factor, mode, t.waves, t.zchunk, s);
    H3D_KB(1) H3D_KB(2) H3D_KB(3) H3D_KB(4) H3D_KB(5) H3D_KB(6) H3D_KB(7) H3D_KB(8)
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
  if (kb > 8) return H3D_ELIMIT;
  hipEvent_t e0, e1;
  H3D_TRY(hipEventCreate(&e0));
  H3D_TRY(hipEventCreate(&e1));
  const int full = steps / kb, rem = steps % kb;
  int rc = H3D_OK;
  hipGraph_t gr = nullptr;
  hipGraphExec_t ge = nullptr;
  if (graph && full >= 2) {
    H3D_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int r1 = h3d_pass_f32(u_b, u_a, kappa_f32, kappa_bf16, c_b, c_a, g, factor, kb, mode, tune, s);
    int r2 = h3d_pass_f32(u_a, u_b, kappa_f32, kappa_bf16, c_a, c_b, g, factor, kb, mode, tune, s);
    hipError_t ec = hipStreamEndCapture(s, &gr);
    if (r1 != H3D_OK) return r1;
    if (r2 != H3D_OK) return r2;
    H3D_TRY(ec);
    H3D_TRY(hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0));
  }
  H3D_TRY(hipEventRecord(e0, s));
  int pass = 0;
  if (ge) {
    for (; pass + 2 <= full; pass += 2) {
      hipError_t el = hipGraphLaunch(ge, s);
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
    H3D_TRY(hipEventRecord(e1, s));
    H3D_TRY(hipEventSynchronize(e1));
    float tms = 0.f;
    H3D_TRY(hipEventElapsedTime(&tms, e0, e1));
    if (ms) *ms = tms;
    if (final_in_b) *final_in_b = npass % 2;
  }
  if (ge) (void)hipGraphExecDestroy(ge);
  if (gr) (void)hipGraphDestroy(gr);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return rc;
}
