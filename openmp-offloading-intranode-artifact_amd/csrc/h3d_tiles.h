// Tile-level device code shared by the standalone kernels (h3d_f64.hip,
// h3d_f32.hip) and the ensemble kernel (h3d_ensemble.hip).  Including
// translation units must be compiled with -ffp-contract=off: the FP64 tile
// is bit-identical to the reference's update only without contraction, and
// the FP32 tile to the CPU oracle's step-by-step FP32 arithmetic.
#pragma once

#include "h3d_internal.h"

namespace {

// Raw buffer access to one plane: the resource (base = the plane's first
// element, range = its bytes) is built in scalar registers and each lane
// supplies a 32-bit byte offset, so a load or store costs no 64-bit vector
// address arithmetic.  An offset at or past the range reads 0 and drops a
// write, which is how the tiles below predicate edge lanes and rows.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t plane_rsrc(const void *base,
                                                            unsigned bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(base), 0, (int)bytes,
                                           0x00020000);
}
__device__ __forceinline__ double ld_f64(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 0));
}
// Non-temporal load (cache policy bit 1), for rhs in the single step: each
// cell's rhs is read by one lane only, so it streams through without
// displacing the u and kappa lines that neighbouring waves share in L2
// (151.3 -> 153.8 GLUP/s on the reference's n = 512 case).  The fused pass
// keeps cached rhs loads: its tiles overlap, and a non-temporal rhs there
// cost 9%.
__device__ __forceinline__ double ld_f64_nt(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(r, (int)off, 0, 2));
}
__device__ __forceinline__ void st_f64(__amdgpu_buffer_rsrc_t r, unsigned off, double v) {
  using v2u = decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0));
  __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, v), r, (int)off, 0, 0);
}

// One tile (tx, ty, tz) of the FP64 step: waves x 64 lanes, RY rows per
// lane, planes k_start + tz*zchunk ... (see h3d_f64.hip).  Lane l of tile
// tx holds x = x0 + 64*tx + l; x0 <= 1 is chosen by the caller so that
// tiles start on 128-byte lines of u_new where the layout allows (x0 = 1:
// the classic origin; x0 = 0 or negative: lanes left of the interior idle).
// Wavefronts are independent (no barrier), so a wave with nothing to do
// returns at once.  z-neighbours stay in registers, x-neighbours come from
// the adjacent lanes (DPP) with the tile's first and last interior lanes
// loading the x halo, and each wave loads its y halo rows.  Accesses are
// buffer instructions; lanes and rows outside the array (loads) or off the
// interior (stores) carry an out-of-range offset instead of a branch.
template <int RY, bool RHS>
__device__ __forceinline__ void march_f64_tile(
    double *__restrict__ un, const double *__restrict__ uo,
    const double *__restrict__ rhs, const double *__restrict__ kap, int nx,
    int ny, int sy, long long sz, double factor, int k_start, int k_end,
    int zchunk, int x0, int tx, int ty, int tz) {
#pragma clang fp contract(off)
  // past every plane: planes hold at most 2^28 elements (2^31 bytes)
  constexpr unsigned OOB = 0xfffffff8u;
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int waves = blockDim.x >> 6;

  const int i = x0 + tx * 64 + lane;
  const int j0 = 1 + (ty * waves + wave) * RY;
  if (j0 > ny - 2) return;  // wave-uniform; no barriers in this function
  const int kb = k_start + tz * zchunk;
  if (kb > k_end) return;
  const int ke = min(k_end, kb + zchunk - 1);
  const int nz_last = k_end + 1;  // last plane any load of this range needs
  const unsigned b8 = (unsigned)((unsigned long long)sz * 8ull);

  const bool lane_ld = i >= 0 && i <= nx - 1;   // this lane's column exists
  const bool lane_cmp = i >= 1 && i <= nx - 2;  // and is interior
  // x halo: lane 0 holds the point left of the segment, lane 63 the point
  // right of it (each needed only when that lane itself is interior; a
  // lane left of it is then the boundary column, which lane 0 is not).
  const bool xh_ld = (lane == 0 || lane == 63) && lane_cmp;
  const int xh_di = (lane == 0) ? -1 : 1;

  unsigned ro[RY], xo[RY], so[RY];
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int j = j0 + r;
    const unsigned e = (unsigned)(j * sy + i);
    ro[r] = (lane_ld && j <= ny - 1) ? e * 8u : OOB;
    xo[r] = (xh_ld && j <= ny - 2) ? (e + xh_di) * 8u : OOB;
    so[r] = (lane_cmp && j <= ny - 2) ? e * 8u : OOB;
  }
  const unsigned lo = lane_ld ? (unsigned)((j0 - 1) * sy + i) * 8u : OOB;
  const unsigned hi =
      (lane_ld && j0 + RY <= ny - 1) ? (unsigned)((j0 + RY) * sy + i) * 8u : OOB;
  auto rs = [&](const double *p, int plane) { return plane_rsrc(p + (long long)plane * sz, b8); };

  double um[RY], uc[RY], up[RY], km[RY], kc[RY], kq[RY];
  double uxh[RY], kxh[RY], ulo, uhi, klo, khi;
  double rc[RY];
  {
    const auto u0 = rs(uo, kb - 1), k0 = rs(kap, kb - 1);
    const auto u1 = rs(uo, kb), k1 = rs(kap, kb);
    const auto u2 = rs(uo, kb + 1), k2 = rs(kap, kb + 1);
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      um[r] = ld_f64(u0, ro[r]);
      km[r] = ld_f64(k0, ro[r]);
      uc[r] = ld_f64(u1, ro[r]);
      kc[r] = ld_f64(k1, ro[r]);
      up[r] = ld_f64(u2, ro[r]);
      kq[r] = ld_f64(k2, ro[r]);
      uxh[r] = ld_f64(u1, xo[r]);
      kxh[r] = ld_f64(k1, xo[r]);
      if (RHS) rc[r] = ld_f64_nt(rs(rhs, kb), ro[r]);
    }
    ulo = ld_f64(u1, lo);
    klo = ld_f64(k1, lo);
    uhi = ld_f64(u1, hi);
    khi = ld_f64(k1, hi);
  }

  for (int k = kb; k <= ke; ++k) {
    // Prefetch: halos of plane k+1 and rows of plane k+2 (plane indices
    // clamped to the range's last plane; values fetched past ke are unused).
    const int p1 = min(k + 1, nz_last), p2 = min(k + 2, nz_last);
    double nu[RY], nk[RY], nuxh[RY], nkxh[RY], nulo, nuhi, nklo, nkhi, nr[RY];
    {
      const auto u1 = rs(uo, p1), k1 = rs(kap, p1);
      const auto u2 = rs(uo, p2), k2 = rs(kap, p2);
#pragma unroll
      for (int r = 0; r < RY; ++r) {
        nu[r] = ld_f64(u2, ro[r]);
        nk[r] = ld_f64(k2, ro[r]);
        nuxh[r] = ld_f64(u1, xo[r]);
        nkxh[r] = ld_f64(k1, xo[r]);
        if (RHS) nr[r] = ld_f64_nt(rs(rhs, p1), ro[r]);
      }
      nulo = ld_f64(u1, lo);
      nklo = ld_f64(k1, lo);
      nuhi = ld_f64(u1, hi);
      nkhi = ld_f64(k1, hi);
    }

    const auto out = rs(un, k);
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
      st_f64(out, so[r], res);
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


// ---- FP32 pass: see h3d_f32.hip for the temporal-blocking design ----

// ulp exponent of a stored value: max(ilogb(u), -126); -126 for zero,
// subnormals and non-finite values.
__device__ __forceinline__ int ulp_exp(float u) {
  const unsigned b = __float_as_uint(u) & 0x7fffffffu;
  const int e = (int)(b >> 23) - 127;
  return (e <= -127 || e == 128) ? -126 : e;
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

// Lane shifts for the FP32 pass.  Lanes without a source read 0: the two
// edge lanes of a tile are never part of its valid region.
__device__ __forceinline__ float lane_right(float v) {  // lane l <- l+1
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xf, 0xf, true));
}
__device__ __forceinline__ float lane_left(float v) {  // lane l <- l-1
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xf, 0xf, true));
}
// The same for the FP64 pass, one 32-bit half at a time.
__device__ __forceinline__ double lane_right(double v) {  // lane l <- l+1
  const int2 s = __builtin_bit_cast(int2, v);
  return __builtin_bit_cast(double, make_int2(__builtin_amdgcn_mov_dpp(s.x, 0x130, 0xf, 0xf, true),
                                              __builtin_amdgcn_mov_dpp(s.y, 0x130, 0xf, 0xf, true)));
}
__device__ __forceinline__ double lane_left(double v) {  // lane l <- l-1
  const int2 s = __builtin_bit_cast(int2, v);
  return __builtin_bit_cast(double, make_int2(__builtin_amdgcn_mov_dpp(s.x, 0x138, 0xf, 0xf, true),
                                              __builtin_amdgcn_mov_dpp(s.y, 0x138, 0xf, 0xf, true)));
}

__device__ __forceinline__ float ld_f32(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __uint_as_float(__builtin_amdgcn_raw_buffer_load_b32(r, (int)off, 0, 0));
}
__device__ __forceinline__ unsigned short ld_u16(__amdgpu_buffer_rsrc_t r, unsigned off) {
  return __builtin_amdgcn_raw_buffer_load_b16(r, (int)off, 0, 0);
}
__device__ __forceinline__ void st_f32(__amdgpu_buffer_rsrc_t r, unsigned off, float v) {
  __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), r, (int)off, 0, 0);
}
__device__ __forceinline__ void st_u16(__amdgpu_buffer_rsrc_t r, unsigned off,
                                       unsigned short v) {
  __builtin_amdgcn_raw_buffer_store_b16(v, r, (int)off, 0, 0);
}

// LDS of one workgroup, in floats, for KB fused steps and W waves of RY rows:
//   kr : [KB+1 slots][TY rows][64]  kappa of planes it-KB .. it (plane p in
//        slot p mod KB+1)
//   ue : [2][KB][W][2][64]  edge rows of every level, double buffered
// Every face sum k_a + k_b is formed where it is used, from the same two
// operands, so it has the bits a stored sum would have; holding kappa
// instead of three sums per cell takes KB+1 planes of LDS instead of 3*KB.
__host__ __device__ constexpr int pass_lds_floats(int KB, int RY, int W) {
  return (KB + 1) * (W * RY) * 64 + 2 * KB * W * 2 * 64;
}

// How far ahead the FP32 pass loads planes, and how many prefetch slots it
// cycles through (the march loop is unrolled by that many iterations, so
// every slot index is static).  16-wave tiles (one workgroup per CU, a
// barrier per plane) hold two planes in flight where the registers allow
// it without spilling (up to 3 steps and 4 rows per lane, residual in
// registers): 3-step passes on 64 x 64 tiles ran 81-84% of the stream rate
// against 77% with one plane in flight.  Unrolling by 4 alone helps the
// 16-wave tiles too; with the FP16 residual field those tiles spill, and
// the longer loop made that worse, so they keep the 2-slot loop, as do
// the 8- and 4-wave tiles (two or more workgroups per CU hide the latency).
__host__ __device__ constexpr int pass_prefetch(int KB, int RY, int W, int MODE) {
  return (W == 16 && MODE < 2 && KB <= 3 && RY <= 4) ? 2 : 1;
}
__host__ __device__ constexpr int pass_slots(int KB, int RY, int W, int MODE) {
  return (W == 16 && MODE < 2) ? 4 : 2;
}

template <int KB, int RY, int W, int MODE, bool BF16>
struct Pass {
  // past the range of every plane resource (planes hold at most 2^28
  // elements, 1 GiB of floats)
  static constexpr unsigned OUT_OF_RANGE = 0x80000000u;
  static constexpr int TY = W * RY;
  static constexpr int SLOT = TY * 64;  // floats per kappa slot
  // ---- per-lane state ----
  float A[KB][2][RY];        // u of level l: slot P = plane it-l-1, 1-P = it-l
  float E[KB][2][RY];        // residual of level l, same slots
  float FZ[KB + 1][RY];      // z flux above the last plane of level l
  static constexpr int PF = pass_prefetch(KB, RY, W, MODE);  // planes ahead
  static constexpr int NS = pass_slots(KB, RY, W, MODE);     // prefetch slots
  float pu[NS][RY];           // prefetched u rows (slot = iteration mod NS)
  float pk[NS][RY];           // prefetched kappa rows
  unsigned short pc[NS][RY];  // prefetched residual bits (COMP16)

  // ---- geometry ----
  const float *uin;
  const float *kf;
  const unsigned short *kb16;
  const unsigned short *cin;
  float *uout;
  unsigned short *cout;
  float factor;
  int nz, w;
  long long sz;
  unsigned o4[RY];  // byte offsets of rows j0..j0+RY-1 in a float plane
  unsigned o2[RY];  // the same in a 16-bit plane
  unsigned s4[RY];  // store offsets: o4, or OUT_OF_RANGE off the output region
  unsigned s2[RY];  // the same for the 16-bit residual plane
  bool lane_int;
  bool row_int[RY];
  int zb, ze;
  float *kr;        // this lane, slot 0, this wave's row 0
  const float *kb;  // this lane, slot 0, the row below this wave's rows
  const float *ka;  // this lane, slot 0, the row above this wave's rows
  float *ue;        // this lane, (phase 0, level 0, wave 0, lo)

  __device__ __forceinline__ int eslot(int p) const {  // p mod KB+1
    const int s = p % (KB + 1);
    return s < 0 ? s + KB + 1 : s;
  }
  __device__ __forceinline__ static constexpr int ue_at(int ph, int l, int wave, int hi) {
    return (((ph * KB + l) * W + wave) * 2 + hi) * 64;
  }

  __device__ __forceinline__ void load_plane(int p, int ph) {
    const int pcl = p < 0 ? 0 : (p > nz - 1 ? nz - 1 : p);
    const long long plane = (long long)pcl * sz;
    const unsigned b4 = (unsigned)sz * 4u, b2 = (unsigned)sz * 2u;
    const auto ru = plane_rsrc(uin + plane, b4);
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      pu[ph][r] = ld_f32(ru, o4[r]);
      if (MODE == 2) pc[ph][r] = ld_u16(plane_rsrc(cin + plane, b2), o2[r]);
      if (BF16) pk[ph][r] = bf16_to_f32(ld_u16(plane_rsrc(kb16 + plane, b2), o2[r]));
      else pk[ph][r] = ld_f32(plane_rsrc(kf + plane, b4), o4[r]);
    }
  }

  // One march iteration at phase PH (iteration mod NS): the level slots
  // alternate with P = PH mod 2, the prefetch slots cycle with PH.
  template <int PH>
  __device__ __forceinline__ void iter(int it) {
#pragma clang fp contract(off)
    constexpr int P = PH & 1;
    // 1. kappa of plane it into its ring slot (the slot of it-KB-1, which no
    //    level reads any more).  Other waves read these rows from the next
    //    iteration on, after the barrier; this lane's own rows of plane it
    //    are taken from registers below.
    {
      float *ks = kr + eslot(it) * SLOT;
#pragma unroll
      for (int r = 0; r < RY; ++r) ks[r * 64] = pk[PH][r];
    }
    // 2. Level 0: the loaded plane becomes the newest slot of level 0.
    float kq1[RY];  // kappa of plane q+1 (this lane's rows)
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      A[0][1 - P][r] = pu[PH][r];
      if (MODE == 2) E[0][1 - P][r] = comp_decode(pc[PH][r], pu[PH][r]);
      else E[0][1 - P][r] = 0.0f;
      kq1[r] = pk[PH][r];
    }
    // 3. Prefetch the plane PF ahead (clamped, so always a valid address).
    load_plane(it + PF, (PH + PF) % NS);

    // 4. Levels 1..KB at planes it-1 .. it-KB.
    const int wb = w > 0 ? w - 1 : 0;
    const int wa = w < W - 1 ? w + 1 : W - 1;
#pragma unroll
    for (int l = 1; l <= KB; ++l) {
      const int q = it - l;
      const bool q_int = q >= 1 && q <= nz - 2;
      const bool q_out = (l == KB) && q >= zb && q <= ze;
      const float ub = ue[ue_at(1 - P, l - 1, wb, 1)];  // row below (mid plane)
      const float ua = ue[ue_at(1 - P, l - 1, wa, 0)];  // row above
      const int so = eslot(q) * SLOT;
      float kc[RY];
#pragma unroll
      for (int r = 0; r < RY; ++r) kc[r] = kr[so + r * 64];
      const float k_below = kb[so];
      const float k_above = ka[so];
      const long long qplane = (long long)(q_out ? q : zb) * sz;
      const auto ro = plane_rsrc(uout + qplane, (unsigned)sz * 4u);
      const auto rc = plane_rsrc(MODE == 2 ? cout + qplane : nullptr, (unsigned)sz * 2u);
      float fy_prev = 0.0f;
#pragma unroll
      for (int r = 0; r < RY; ++r) {
        const float u = A[l - 1][P][r];
        const float un = A[l - 1][1 - P][r];
        const float k = kc[r];
        const float ccx = lane_right(k) + k;
        const float ccy = ((r < RY - 1) ? kc[r + 1] : k_above) + k;
        const float ccz = k + kq1[r];
        const float fxp = ccx * (lane_right(u) - u);
        const float fxm = lane_left(fxp);
        const float ujp = (r < RY - 1) ? A[l - 1][P][r + 1] : ua;
        const float fyp = ccy * (ujp - u);
        const float fym = (r > 0) ? fy_prev : (k + k_below) * (u - ub);
        fy_prev = fyp;
        const float fzp = ccz * (un - u);
        const float fzm = FZ[l][r];
        FZ[l][r] = fzp;
        const float s = ((((fxp - fxm) + fyp) - fym) + fzp) - fzm;
        const bool inside = lane_int && row_int[r] && q_int;
        float res, e2 = 0.0f;
        if (MODE == 0) {
          res = __builtin_fmaf(factor, s, u);
        } else {
          // Kahan: carry the part of the increment the FP32 sum dropped;
          // the increment factor*s joins the carried residual in one fused
          // multiply-add (one rounding)
          const float y = __builtin_fmaf(factor, s, E[l - 1][P][r]);
          const float t = u + y;
          e2 = y - (t - u);
          res = t;
        }
        if (l < KB) {
          // Dirichlet cells keep their value; the residual of a cell is used
          // only by that cell, so it needs no mask.
          A[l][1 - P][r] = inside ? res : u;
          E[l][1 - P][r] = e2;
        } else {
          // Stored cells are interior on every axis, so the last level needs
          // no Dirichlet select.  Every row issues its store: rows and lanes
          // off the output region, and every lane on a warm-up plane of the
          // chunk (q outside zb..ze), carry an out-of-range offset and the
          // buffer range check drops the write.  No branch, so the whole
          // iteration is one basic block for the scheduler.
          st_f32(ro, q_out ? s4[r] : OUT_OF_RANGE, res);
          if (MODE == 2) st_u16(rc, q_out ? s2[r] : OUT_OF_RANGE, comp_encode(e2, res));
        }
      }
#pragma unroll
      for (int r = 0; r < RY; ++r) kq1[r] = kc[r];
    }
    // 5. This iteration's newest edge rows of levels 0..KB-1.
#pragma unroll
    for (int l = 0; l < KB; ++l) {
      ue[ue_at(P, l, w, 0)] = A[l][1 - P][0];
      ue[ue_at(P, l, w, 1)] = A[l][1 - P][RY - 1];
    }
    __syncthreads();
  }
};

// One workgroup's tile (tx, ty), output planes zb..ze (1 <= zb <= ze <=
// nz-2), of an FP32 pass with W waves; `lds` holds pass_lds_floats(KB, RY,
// W) floats.  Every thread of the workgroup must call it (it synchronises
// the workgroup).
template <int KB, int RY, int W, int MODE, bool BF16>
__device__ __forceinline__ void pass_f32_tile(
    float *lds, float *__restrict__ uout, const float *__restrict__ uin,
    const float *__restrict__ kf, const unsigned short *__restrict__ kb16,
    unsigned short *__restrict__ cout, const unsigned short *__restrict__ cin,
    int nx, int ny, int nz, int sy, long long sz, float factor, int zb,
    int ze, int tx, int ty) {
  using S = Pass<KB, RY, W, MODE, BF16>;
  S st;
  st.uin = uin; st.kf = kf; st.kb16 = kb16; st.cin = cin;
  st.uout = uout; st.cout = cout; st.factor = factor;
  st.nz = nz; st.sz = sz;
  const int lane = threadIdx.x & 63;
  st.w = threadIdx.x >> 6;
  constexpr int OX = 64 - 2 * KB, OY = S::TY - 2 * KB;

  const int i = 1 - KB + tx * OX + lane;
  const int y0 = 1 - KB + ty * OY;  // global y of tile row 0
  const int j0 = y0 + st.w * RY;
  st.zb = zb;
  st.ze = ze;

  // Loads are clamped into the array: cells outside it, and the Dirichlet
  // ring, are carried through every level unchanged, and no interior cell
  // reads anything but interior or ring cells, so clamped values never
  // reach a written result.
  const int ic = i < 0 ? 0 : (i > nx - 1 ? nx - 1 : i);
  st.lane_int = i >= 1 && i <= nx - 2;
  const bool lane_out = lane >= KB && lane <= 63 - KB && i <= nx - 2;
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int j = j0 + r;
    const int jc = j < 0 ? 0 : (j > ny - 1 ? ny - 1 : j);
    const unsigned e = (unsigned)(jc * sy + ic);
    st.o4[r] = e * 4u;
    st.o2[r] = e * 2u;
  }
  constexpr unsigned OUT_OF_RANGE = S::OUT_OF_RANGE;
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int j = j0 + r, tr = st.w * RY + r;
    st.row_int[r] = j >= 1 && j <= ny - 2;
    const bool out = lane_out && tr >= KB && tr <= S::TY - 1 - KB && j <= ny - 2;
    st.s4[r] = out ? st.o4[r] : OUT_OF_RANGE;
    st.s2[r] = out ? st.o2[r] : OUT_OF_RANGE;
  }
  // kappa rows next to this wave's rows: the other waves' edge rows.  The
  // tile's own edge rows (row 0 of wave 0, the last row of wave W-1) take
  // their own row instead; their cells are off the output region at every
  // level above 0, so the value never reaches a stored result.
  st.kr = lds + st.w * RY * 64 + lane;
  st.kb = lds + (st.w > 0 ? st.w * RY - 1 : 0) * 64 + lane;
  st.ka = lds + (st.w < W - 1 ? (st.w + 1) * RY : S::TY - 1) * 64 + lane;
  st.ue = lds + (KB + 1) * S::SLOT + lane;

#pragma unroll
  for (int l = 0; l <= KB; ++l)
#pragma unroll
    for (int r = 0; r < RY; ++r) st.FZ[l][r] = 0.0f;
#pragma unroll
  for (int r = 0; r < RY; ++r) {
#pragma unroll
    for (int l = 0; l < KB; ++l)
      st.A[l][0][r] = st.A[l][1][r] = st.E[l][0][r] = st.E[l][1][r] = 0.0f;
  }
  // All LDS starts defined (warm-up iterations read slots not yet written).
  for (int x = threadIdx.x; x < pass_lds_floats(KB, RY, W); x += W * 64) lds[x] = 0.0f;
  __syncthreads();

  // Iterations it0 .. cover zb-KB .. ze+KB in groups of NS; trailing
  // iterations of the last group write nothing (their planes are > ze).
  const int it0 = st.zb - KB;
  constexpr int NS = S::NS;
  const int groups = (st.ze + KB - it0 + NS) / NS;
#pragma unroll
  for (int f = 0; f < S::PF; ++f) st.load_plane(it0 + f, f);
  for (int p = 0; p < groups; ++p) {
    st.template iter<0>(it0 + NS * p);
    st.template iter<1>(it0 + NS * p + 1);
    if constexpr (NS == 4) {
      st.template iter<2>(it0 + NS * p + 2);
      st.template iter<3>(it0 + NS * p + 3);
    }
  }
}

// ---- FP64 fused pass: KB steps of the reference's update per HBM pass ----
//
// The FP32 pass's tile and march (above) with the FP64 update of
// march_f64_tile: every face term is the reference's expression with the
// same operands (a flux computed for one cell is the opposite face's term
// of its neighbour, bit for bit), u + factor*s or (u + rhs) + factor*s
// without contraction, so KB fused steps give the bits of KB single steps.
//
// The kappa face sums of a plane (the x, y and z faces above each cell
// and the y face below a wave's first row) are formed once, when the
// plane enters level 1, from the operands the single step adds, and kept
// in a register ring of KB planes that levels 2..KB read: a sum formed
// once has the bits of the same sum formed at every level.  kappa passes
// through LDS once: plane it is written at iteration it and read at
// iteration it+1 (the edge rows of the waves next to it, and at depth 2 a
// wave's own rows), so two LDS slots suffice.  With rhs, each lane keeps
// the rhs of its cells for the last KB planes in a second register ring
// (every level adds it; a cell reads only its own).  The march is unrolled by
// lcm(2, KB) iterations, so every ring slot, level slot and LDS slot is a
// compile-time index.

// LDS of one workgroup, in doubles (rhs needs none):
//   kl : [2 slots][TY rows][64]           kappa of planes it-1 and it
//   ue : [2][KB][W][2][64]  edge rows of every level, double buffered
__host__ __device__ constexpr int pass64_lds_doubles(int KB, int RY, int W) {
  return 2 * (W * RY) * 64 + 2 * KB * W * 2 * 64;
}

// Iterations per unrolled group of the FP64 march: lcm(2, KB).
__host__ __device__ constexpr int pass64_unroll(int KB) { return KB % 2 == 0 ? KB : 2 * KB; }

template <int KB, int RY, int W, bool RHS>
struct Pass64 {
  // at or past the range of every plane resource (planes hold at most
  // 2^28 elements, 2 GiB of doubles)
  static constexpr unsigned OUT_OF_RANGE = 0x80000000u;
  static constexpr int TY = W * RY;
  static constexpr int SLOT = TY * 64;  // doubles per kappa slot
  static constexpr int U = pass64_unroll(KB);
  // A wave's own kappa rows of plane it-1 are read from its prefetch
  // registers instead of back from LDS (only the edge rows then pass
  // through LDS) at depths 3 and 4: on MI355X at 512^3, 4-step passes on
  // 8 waves x 4 rows 996 -> 975 us, 3-step passes on 8 x 4 748 -> 741 us,
  // 3-step passes on 16 x 2 unchanged.  The 2-step pass on 16 x 2 ran
  // 768 -> 793 us that way, so depth 2 keeps the LDS copy.
  static constexpr bool KQ_REG = KB >= 3;
  // A Dirichlet cell keeps its value through the levels (`inside ? res :
  // u`).  hipcc turns that select into a branch on the lane mask around the
  // update's dependent add chain, one per row and level, and the rows then
  // run one after the other.  An empty asm statement on the result keeps
  // the update outside the select, so the rows' chains interleave; a
  // scheduling barrier after each level keeps the register use near the
  // branchy form's.  Measured at 512^3: 4-step passes on 8 waves x 4 rows
  // 1046-1056 -> 996-1001 us per launch, 3-step passes 0.8-1% faster; the
  // 2-step passes ran 2% slower, so depth 2 keeps the branches.  With rhs
  // (a ring of rhs values in registers) the form spills: 3-step passes on
  // 16 waves x 2 rows 1178 -> 2221 us, 4-step on 8 x 4 3650 -> 6129 us
  // (3-step on 8 x 4: 1015 -> 1007), so passes with rhs keep the branches.
  static constexpr bool SELECT_OPAQUE = KB >= 3 && !RHS;
  // ---- per-lane state ----
  double A[KB][2][RY];    // u of level l: slot P = plane it-l-1, 1-P = it-l
  double FZ[KB + 1][RY];  // z flux above the last plane of level l
  double pu[2][RY];       // prefetched u rows (slot = iteration parity)
  double pk[2][RY];       // prefetched kappa rows
  double pr[2][RY];       // prefetched rhs rows (RHS)
  // rings over the last KB planes: plane q in slot (q - it0) mod KB
  double CX[KB][RY];      // k(i+1) + k(i): x face above each cell
  double CY[KB][RY];      // k(j+1) + k(j): y face above
  double CZ[KB][RY];      // k(q) + k(q+1): z face above
  double CB[KB];          // k(j) + k(j-1): y face below the wave's first row
  double RH[KB][RY];      // rhs (RHS)

  // ---- geometry ----
  const double *uin;
  const double *kf;
  const double *rh;
  double *uout;
  double factor;
  int nz, w;
  long long sz;
  unsigned o8[RY];      // byte offsets of rows j0..j0+RY-1 in a plane
  unsigned s8[RY];      // store offsets: o8, or OUT_OF_RANGE off the output region
  bool lane_int;
  bool row_int[RY];
  int zb, ze;
  double *kl;        // this lane, slot 0, this wave's row 0
  const double *kb;  // this lane, slot 0, the row below this wave's rows
  const double *ka;  // this lane, slot 0, the row above this wave's rows
  double *ue;        // this lane, (phase 0, level 0, wave 0, lo)

  __device__ __forceinline__ static constexpr int ue_at(int ph, int l, int wave, int hi) {
    return (((ph * KB + l) * W + wave) * 2 + hi) * 64;
  }
  __device__ __forceinline__ static constexpr int ring(int ph) {  // ph mod KB, ph > -KB
    return (ph + KB) % KB;
  }

  __device__ __forceinline__ void load_plane(int p, int ph) {
    const int pcl = p < 0 ? 0 : (p > nz - 1 ? nz - 1 : p);
    const long long plane = (long long)pcl * sz;
    const unsigned b8 = (unsigned)((unsigned long long)sz * 8ull);
    const auto ru = plane_rsrc(uin + plane, b8);
    const auto rk = plane_rsrc(kf + plane, b8);
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      pu[ph][r] = ld_f64(ru, o8[r]);
      if (RHS) pr[ph][r] = ld_f64(plane_rsrc(rh + plane, b8), o8[r]);
      pk[ph][r] = ld_f64(rk, o8[r]);
    }
  }

  // One march iteration, the PH-th of an unrolled group (it - it0 = PH
  // mod U): level slots and LDS slots alternate with P = PH mod 2, the
  // prefetch slots likewise, the rings cycle with PH mod KB.
  template <int PH>
  __device__ __forceinline__ void iter(int it) {
#pragma clang fp contract(off)
    constexpr int P = PH & 1;
    // 1. kappa of plane it into LDS slot P: the edge rows the neighbouring
    //    waves read, and at KB = 2 every row (see KQ_REG).  The slot held
    //    plane it-2, last read at iteration it-1, before its closing
    //    barrier.
#pragma unroll
    for (int r = 0; r < RY; ++r)
      if (!KQ_REG || r == 0 || r == RY - 1) kl[P * SLOT + r * 64] = pk[P][r];
    // 2. Face sums of plane it-1 (level 1's plane) into ring slot PH-1:
    //    its kappa rows (prefetch slot 1-P, loaded at iteration it-2 and
    //    refilled by this iteration's prefetch below, or LDS slot 1-P),
    //    the edge rows of the waves next to this one from LDS slot 1-P
    //    (written at iteration it-1), kappa of plane it from registers.
    {
      constexpr int c1 = ring(PH - 1);
      double kq[RY];
#pragma unroll
      for (int r = 0; r < RY; ++r) kq[r] = KQ_REG ? pk[1 - P][r] : kl[(1 - P) * SLOT + r * 64];
      const double k_below = kb[(1 - P) * SLOT];
      const double k_above = ka[(1 - P) * SLOT];
#pragma unroll
      for (int r = 0; r < RY; ++r) {
        CX[c1][r] = lane_right(kq[r]) + kq[r];
        CY[c1][r] = ((r < RY - 1) ? kq[r + 1] : k_above) + kq[r];
        CZ[c1][r] = kq[r] + pk[P][r];
      }
      CB[c1] = kq[0] + k_below;
    }
    double rn[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) rn[r] = RHS ? pr[P][r] : 0.0;  // rhs of plane it
    // 3. Level 0: the loaded plane becomes the newest slot of level 0.
#pragma unroll
    for (int r = 0; r < RY; ++r) A[0][1 - P][r] = pu[P][r];
    // 4. Prefetch plane it+1 (clamped, so always a valid address).
    load_plane(it + 1, 1 - P);

    // 5. Levels 1..KB at planes it-1 .. it-KB.
    const int wb = w > 0 ? w - 1 : 0;
    const int wa = w < W - 1 ? w + 1 : W - 1;
    const unsigned b8 = (unsigned)((unsigned long long)sz * 8ull);
#pragma unroll
    for (int l = 1; l <= KB; ++l) {
      const int c = ring(PH - l);  // ring slot of plane q
      const int q = it - l;
      const bool q_int = q >= 1 && q <= nz - 2;
      const bool q_out = (l == KB) && q >= zb && q <= ze;
      const double ub = ue[ue_at(1 - P, l - 1, wb, 1)];  // row below (mid plane)
      const double ua = ue[ue_at(1 - P, l - 1, wa, 0)];  // row above
      const long long qplane = (long long)(q_out ? q : zb) * sz;
      const auto ro = plane_rsrc(uout + qplane, b8);
      double fy_prev = 0.0;
#pragma unroll
      for (int r = 0; r < RY; ++r) {
        const double u = A[l - 1][P][r];
        const double un = A[l - 1][1 - P][r];
        const double fxp = CX[c][r] * (lane_right(u) - u);
        const double fxm = lane_left(fxp);
        const double ujp = (r < RY - 1) ? A[l - 1][P][r + 1] : ua;
        const double fyp = CY[c][r] * (ujp - u);
        const double fym = (r > 0) ? fy_prev : CB[c] * (u - ub);
        fy_prev = fyp;
        const double fzp = CZ[c][r] * (un - u);
        const double fzm = FZ[l][r];
        FZ[l][r] = fzp;
        const double s = ((((fxp - fxm) + fyp) - fym) + fzp) - fzm;
        const double res = RHS ? (u + RH[c][r]) + factor * s : u + factor * s;
        if (l < KB) {
          // Dirichlet cells keep their value
          const bool inside = lane_int && row_int[r] && q_int;
          double rv = res;
          if constexpr (SELECT_OPAQUE) __asm__("" : "+v"(rv));
          A[l][1 - P][r] = inside ? rv : u;
        } else {
          // stored cells are interior on every axis; rows and lanes off the
          // output region, and every lane on a warm-up plane of the chunk,
          // carry an out-of-range offset (dropped write), so no branch
          st_f64(ro, q_out ? s8[r] : OUT_OF_RANGE, res);
        }
      }
      if constexpr (SELECT_OPAQUE) __builtin_amdgcn_sched_barrier(0);
    }
    // 6. rhs ring: plane it takes the slot of plane it-KB, which level KB
    //    has just read
    if (RHS) {
#pragma unroll
      for (int r = 0; r < RY; ++r) RH[ring(PH)][r] = rn[r];
    }
    // 7. This iteration's newest edge rows of levels 0..KB-1.
#pragma unroll
    for (int l = 0; l < KB; ++l) {
      ue[ue_at(P, l, w, 0)] = A[l][1 - P][0];
      ue[ue_at(P, l, w, 1)] = A[l][1 - P][RY - 1];
    }
    __syncthreads();
  }

  // Iterations PH .. U-1 of a group starting at iteration it.
  template <int PH>
  __device__ __forceinline__ void group(int it) {
    iter<PH>(it + PH);
    if constexpr (PH + 1 < U) group<PH + 1>(it);
  }
};

// One workgroup's tile (tx, ty), output planes zb..ze (1 <= zb <= ze <=
// nz-2), of an FP64 pass of KB steps with W waves; `lds` holds
// pass64_lds_doubles(KB, RY, W) doubles.  Every thread of the
// workgroup must call it (it synchronises the workgroup).
template <int KB, int RY, int W, bool RHS>
__device__ __forceinline__ void pass_f64_tile(
    double *lds, double *__restrict__ uout, const double *__restrict__ uin,
    const double *__restrict__ rhs, const double *__restrict__ kf, int nx,
    int ny, int nz, int sy, long long sz, double factor, int zb, int ze,
    int tx, int ty) {
  using S = Pass64<KB, RY, W, RHS>;
  S st;
  st.uin = uin; st.kf = kf; st.rh = rhs; st.uout = uout; st.factor = factor;
  st.nz = nz; st.sz = sz;
  const int lane = threadIdx.x & 63;
  st.w = threadIdx.x >> 6;
  constexpr int OX = 64 - 2 * KB, OY = S::TY - 2 * KB;

  const int i = 1 - KB + tx * OX + lane;
  const int y0 = 1 - KB + ty * OY;  // global y of tile row 0
  const int j0 = y0 + st.w * RY;
  st.zb = zb;
  st.ze = ze;

  // Loads are clamped into the array: cells outside it, and the Dirichlet
  // ring, are carried through every level unchanged, and no interior cell
  // reads anything but interior or ring cells, so clamped values never
  // reach a written result.
  const int ic = i < 0 ? 0 : (i > nx - 1 ? nx - 1 : i);
  st.lane_int = i >= 1 && i <= nx - 2;
  const bool lane_out = lane >= KB && lane <= 63 - KB && i <= nx - 2;
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int j = j0 + r;
    const int jc = j < 0 ? 0 : (j > ny - 1 ? ny - 1 : j);
    st.o8[r] = (unsigned)(jc * sy + ic) * 8u;
  }
  constexpr unsigned OUT_OF_RANGE = S::OUT_OF_RANGE;
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int j = j0 + r, tr = st.w * RY + r;
    st.row_int[r] = j >= 1 && j <= ny - 2;
    const bool out = lane_out && tr >= KB && tr <= S::TY - 1 - KB && j <= ny - 2;
    st.s8[r] = out ? st.o8[r] : OUT_OF_RANGE;
  }
  // kappa rows next to this wave's rows: the other waves' edge rows.  The
  // tile's own edge rows (row 0 of wave 0, the last row of wave W-1) take
  // their own row instead; their cells are off the output region at every
  // level above 0, so the value never reaches a stored result.
  st.kl = lds + st.w * RY * 64 + lane;
  st.kb = lds + (st.w > 0 ? st.w * RY - 1 : 0) * 64 + lane;
  st.ka = lds + (st.w < W - 1 ? (st.w + 1) * RY : S::TY - 1) * 64 + lane;
  st.ue = lds + 2 * S::SLOT + lane;

#pragma unroll
  for (int l = 0; l <= KB; ++l)
#pragma unroll
    for (int r = 0; r < RY; ++r) st.FZ[l][r] = 0.0;
#pragma unroll
  for (int l = 0; l < KB; ++l) {
    st.CB[l] = 0.0;
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      st.A[l][0][r] = st.A[l][1][r] = st.RH[l][r] = 0.0;
      st.CX[l][r] = st.CY[l][r] = st.CZ[l][r] = 0.0;
    }
  }
  // the first iteration forms plane it0-1's face sums from prefetch slot 1
  // or LDS slot 1 (a warm-up plane: zero in both)
#pragma unroll
  for (int r = 0; r < RY; ++r) st.pk[1][r] = 0.0;
  // All LDS starts defined (warm-up iterations read slots not yet written).
  for (int x = threadIdx.x; x < pass64_lds_doubles(KB, RY, W); x += W * 64) lds[x] = 0.0;
  __syncthreads();

  // Iterations it0 .. cover zb-KB .. ze+KB in groups of U; trailing
  // iterations of the last group write nothing (their planes are > ze).
  const int it0 = st.zb - KB;
  constexpr int U = S::U;
  const int groups = (st.ze + KB - it0 + U) / U;
  st.load_plane(it0, 0);
  for (int p = 0; p < groups; ++p) st.template group<0>(it0 + U * p);
}

// One tile (tx, ty, tz) of a single-step FP32 pass (KB = 1) over planes
// k_start..k_end (z chunk tz), built like the
// FP64 tile: waves x 64 lanes, RY rows per lane, independent wavefronts (no
// LDS, no barrier), z-neighbours in registers, x-neighbours through DPP with
// lanes 0 and 63 loading the x halo, y halo rows loaded per wave.  All
// accesses are raw buffer instructions with per-plane resources; an access
// outside the array or off the tile's interior cells carries an
// out-of-range offset, so it reads 0 or writes nothing without a branch.
// The update is the same FP32 expression as the temporally blocked pass at
// KB = 1 (each face term (k_a + k_b)(u_a - u_b) is evaluated directly, with
// the operands the pass's shared flux has, so the results are the same
// bits); MODE 2 reads and writes the FP16 residual field.
template <int RY, int MODE, bool BF16>
__device__ __forceinline__ void march_f32_tile(
    float *__restrict__ uout, const float *__restrict__ uin,
    const float *__restrict__ kf, const unsigned short *__restrict__ kb16,
    unsigned short *__restrict__ cout, const unsigned short *__restrict__ cin,
    int nx, int ny, int nz, int sy, long long sz, float factor, int k_start,
    int k_end, int zchunk, int tx, int ty, int tz) {
#pragma clang fp contract(off)
  constexpr unsigned OOB = 0x80000000u;  // past every plane (<= 2^28 elements)
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int waves = blockDim.x >> 6;
  const int i = 1 + tx * 64 + lane;
  const int j0 = 1 + (ty * waves + wave) * RY;
  if (j0 > ny - 2) return;  // wave-uniform; no barriers in this function
  const int kb = k_start + tz * zchunk;
  if (kb > k_end) return;
  const int ke = min(k_end, kb + zchunk - 1);
  const unsigned b4 = (unsigned)sz * 4u, b2 = (unsigned)sz * 2u;

  const bool lane_ld = i <= nx - 1;
  const bool lane_cmp = i <= nx - 2;
  const bool xh_ld = (lane == 0) || (lane == 63 && lane_cmp);
  const int xh_di = (lane == 0) ? -1 : 1;
  // element offsets -> byte offsets, OOB where nothing is to be read/written
  unsigned ro[RY], xo[RY], so[RY];
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    const int j = j0 + r;
    const unsigned e = (unsigned)(j * sy + i);
    ro[r] = (lane_ld && j <= ny - 1) ? e : OOB;
    xo[r] = (xh_ld && j <= ny - 2) ? e + xh_di : OOB;
    so[r] = (lane_cmp && j <= ny - 2) ? e : OOB;
  }
  const unsigned lo = lane_ld ? (unsigned)((j0 - 1) * sy + i) : OOB;
  const unsigned hi = (lane_ld && j0 + RY <= ny - 1) ? (unsigned)((j0 + RY) * sy + i) : OOB;
  auto b4o = [](unsigned e) { return e == OOB ? OOB : e * 4u; };
  auto b2o = [](unsigned e) { return e == OOB ? OOB : e * 2u; };
  auto ldu = [&](int p, unsigned e) {
    return ld_f32(plane_rsrc(uin + (long long)p * sz, b4), b4o(e));
  };
  auto ldk = [&](int p, unsigned e) {
    if (BF16) return bf16_to_f32(ld_u16(plane_rsrc(kb16 + (long long)p * sz, b2), b2o(e)));
    return ld_f32(plane_rsrc(kf + (long long)p * sz, b4), b4o(e));
  };
  auto ldc = [&](int p, unsigned e) {
    return ld_u16(plane_rsrc(cin + (long long)p * sz, b2), b2o(e));
  };

  float um[RY], uc[RY], up[RY], km[RY], kc[RY], kq[RY], uxh[RY], kxh[RY];
  unsigned short cc[RY];
  float ulo, uhi, klo, khi;
#pragma unroll
  for (int r = 0; r < RY; ++r) {
    um[r] = ldu(kb - 1, ro[r]);
    km[r] = ldk(kb - 1, ro[r]);
    uc[r] = ldu(kb, ro[r]);
    kc[r] = ldk(kb, ro[r]);
    up[r] = ldu(kb + 1, ro[r]);
    kq[r] = ldk(kb + 1, ro[r]);
    uxh[r] = ldu(kb, xo[r]);
    kxh[r] = ldk(kb, xo[r]);
    if (MODE == 2) cc[r] = ldc(kb, ro[r]);
  }
  ulo = ldu(kb, lo);
  klo = ldk(kb, lo);
  uhi = ldu(kb, hi);
  khi = ldk(kb, hi);

  for (int k = kb; k <= ke; ++k) {
    // Prefetch halos of plane k+1 and rows of plane k+2 (planes clamped
    // into the array; the values fetched past ke are never used).
    const int p1 = min(k + 1, nz - 1), p2 = min(k + 2, nz - 1);
    float nu[RY], nk[RY], nuxh[RY], nkxh[RY], nulo, nuhi, nklo, nkhi;
    unsigned short nc[RY];
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      nu[r] = ldu(p2, ro[r]);
      nk[r] = ldk(p2, ro[r]);
      nuxh[r] = ldu(p1, xo[r]);
      nkxh[r] = ldk(p1, xo[r]);
      if (MODE == 2) nc[r] = ldc(p1, ro[r]);
    }
    nulo = ldu(p1, lo);
    nklo = ldk(p1, lo);
    nuhi = ldu(p1, hi);
    nkhi = ldk(p1, hi);

    const auto ro4 = plane_rsrc(uout + (long long)k * sz, b4);
    const auto rc2 = plane_rsrc(MODE == 2 ? cout + (long long)k * sz : nullptr, b2);
#pragma unroll
    for (int r = 0; r < RY; ++r) {
      const float u = uc[r], kk = kc[r];
      const float uim = __int_as_float(__builtin_amdgcn_update_dpp(
          __float_as_int(uxh[r]), __float_as_int(u), 0x138, 0xf, 0xf, false));
      const float uip = __int_as_float(__builtin_amdgcn_update_dpp(
          __float_as_int(uxh[r]), __float_as_int(u), 0x130, 0xf, 0xf, false));
      const float kim = __int_as_float(__builtin_amdgcn_update_dpp(
          __float_as_int(kxh[r]), __float_as_int(kk), 0x138, 0xf, 0xf, false));
      const float kip = __int_as_float(__builtin_amdgcn_update_dpp(
          __float_as_int(kxh[r]), __float_as_int(kk), 0x130, 0xf, 0xf, false));
      const float ujm = (r == 0) ? ulo : uc[r - 1];
      const float ujp = (r == RY - 1) ? uhi : uc[r + 1];
      const float kjm = (r == 0) ? klo : kc[r - 1];
      const float kjp = (r == RY - 1) ? khi : kc[r + 1];
      const float s = (kip + kk) * (uip - u) - (kk + kim) * (u - uim) +
                      (kjp + kk) * (ujp - u) - (kk + kjm) * (u - ujm) +
                      (kq[r] + kk) * (up[r] - u) - (kk + km[r]) * (u - um[r]);
      if (MODE == 0) {
        st_f32(ro4, b4o(so[r]), __builtin_fmaf(factor, s, u));
      } else {
        const float e = (MODE == 2) ? comp_decode(cc[r], u) : 0.0f;
        const float y = __builtin_fmaf(factor, s, e);
        const float t = u + y;
        st_f32(ro4, b4o(so[r]), t);
        if (MODE == 2) st_u16(rc2, b2o(so[r]), comp_encode(y - (t - u), t));
      }
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
      if (MODE == 2) cc[r] = nc[r];
    }
    ulo = nulo;
    uhi = nuhi;
    klo = nklo;
    khi = nkhi;
  }
}

}  // namespace
