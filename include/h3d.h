/*
 * h3d.h -- C ABI of the MI355X-native 3-D variable-coefficient heat update.
 *
 * The reference (ezhilmathik/OpenMP-Offloading-Intranode-Artifact) has no
 * plugin layer; its boundary is the kernel signature shared by its OpenMP,
 * HIP and CUDA programs plus the `./bin n T` process contract.  The first
 * group of entry points below reproduces that kernel signature exactly;
 * the rest are the grid/stream forms the h3d host program and the tests use.
 *
 * Conventions (all entry points):
 *  - Memory is owned by the caller.  Device pointers come from
 *    h3d_target_alloc (or any hipMalloc'd memory on the same device).
 *  - u_new != u_old; the kernels write interior points only, so the caller
 *    seeds u_new's Dirichlet ring (1-hip.cpp:222-225 does the same).
 *  - Layout [k][j][i], i fastest; point (k,j,i) at k*sz + j*sy + i.  The
 *    dense reference layout is sy = nx, sz = nx*ny (1-openmp.c:28-29).
 *  - Calls are ordered on the given HIP stream (void* so that this header
 *    needs no HIP include; NULL = the device's default stream).  They are
 *    re-entrant on distinct streams and never exit or abort.
 *  - Return value: 0 on success, a negative H3D_E* code for a rejected
 *    argument, or a positive hipError_t code from the runtime.
 */
#ifndef H3D_H
#define H3D_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define H3D_OK 0
#define H3D_EARG (-1)   /* invalid argument: null pointer, size, range  */
#define H3D_ELIMIT (-2) /* supported limit exceeded (e.g. time block)   */
#define H3D_ENODEV (-3) /* no such device                                */
#define H3D_EHOST (-4)  /* host-side failure (malloc, file I/O)          */

typedef struct h3d_grid {
  int nx, ny, nz; /* points per axis, Dirichlet ring included          */
  int sy;         /* row stride in elements, sy >= nx                  */
  long long sz;   /* plane stride in elements, sz >= (long long)sy*ny  */
} h3d_grid;

/* Launch shape of the stencil kernels.  Zero fields take the library's
 * defaults (or the autotuner's cached choice).  Supported shapes (others
 * return H3D_EARG):
 *   FP64 step and FP32 single-step pass: rows 1, 2, 4 (FP64 also 8),
 *     waves 1..16 for FP64, 1..8 for FP32;
 *   FP32 fused passes (kblock 2..6) and FP64 fused passes (kblock 2..4):
 *     rows 2 or 4, waves 4 or 8; FP64 passes of 2 or 3 steps also rows 2
 *     with waves 16. */
typedef struct h3d_tune {
  int rows;   /* y rows per lane held in registers                   */
  int waves;  /* wavefronts per workgroup, stacked in y               */
  int zchunk; /* planes marched by one workgroup                      */
  int kblock; /* time steps fused per HBM pass: FP32 1..8, FP64 1..4  */
} h3d_tune;

/* ---------------------------------------------------------------------- */
/* Reference kernel signature, dense layout, device_id's default stream.  */

/* Replaces heat_update (OpenMP_Single_Source/1-openmp.c:55-58;
 * 3D_Heat_Benchmark/AMD-MI250X/Source-Results/HIP/1-hip.cpp:117-120):
 * every interior point k = 1..nz-2.  rhs may be NULL (treated as zero). */
int h3d_heat_update(double *u_new, const double *u_old, const double *rhs,
                    const double *kappa, int nz, int ny, int nx,
                    double factor, int device_id);

/* Replaces heat_update_interior (OpenMP_Single_Source/2-openmp.c:214-219;
 * 4-openmp-stream-omp-p2p.c:308-313): planes k_start..k_end inclusive. */
int h3d_heat_update_interior(double *u_new, const double *u_old,
                             const double *rhs, const double *kappa, int nz,
                             int ny, int nx, double factor, int k_start,
                             int k_end, int device_id);

/* Replace heat_update_boundary_bottom / _top (2-openmp.c:145-148,
 * 4-openmp-stream-omp-p2p.c:230-234 and :267-271): plane k = 1 / nz-2. */
int h3d_heat_update_boundary_bottom(double *u_new, const double *u_old,
                                    const double *rhs, const double *kappa,
                                    int nz, int ny, int nx, double factor,
                                    int device_id);
int h3d_heat_update_boundary_top(double *u_new, const double *u_old,
                                 const double *rhs, const double *kappa,
                                 int nz, int ny, int nx, double factor,
                                 int device_id);

/* ---------------------------------------------------------------------- */
/* Grid/stream forms.                                                      */

/* One FP64 update of planes k_start..k_end (1 <= k_start, k_end <= nz-2)
 * in the reference's association order, without FMA contraction: bit
 * identical to heat3D_variable_coeff.c:174-181. tune may be NULL. */
int h3d_step_f64(double *u_new, const double *u_old, const double *rhs,
                 const double *kappa, const h3d_grid *g, double factor,
                 int k_start, int k_end, const h3d_tune *tune, void *stream);

/* nsteps (1..4) FP64 updates in one pass over HBM: planes k_start..k_end
 * (1 <= k_start, k_end <= nz-2) of u_out receive those planes of nsteps
 * steps of the reference's update applied to all of u_in, bit identical
 * to nsteps calls of h3d_step_f64 over planes 1..nz-2 (the same expression
 * and association order, no contraction).  An output plane depends on the
 * input planes within nsteps of it; other planes of u_out are not written.
 * Planes are read once and written once per pass, 24/nsteps B per update
 * (32/nsteps with rhs).  u_out's Dirichlet ring must hold u_in's.  No
 * reference counterpart: the reference runs one kernel launch per step
 * (1-openmp.c:193-205).  Shapes: rows 2 or 4, waves 4 or 8, and rows 2
 * with waves 16 at nsteps 2 and 3; any zchunk (H3D_ELIMIT when the
 * shape's LDS exceeds 160 KB).  nsteps == 1 is h3d_step_f64 over the same
 * planes. */
int h3d_pass_f64(double *u_out, const double *u_in, const double *rhs,
                 const double *kappa, const h3d_grid *g, double factor,
                 int nsteps, int k_start, int k_end, const h3d_tune *tune,
                 void *stream);

/* `steps` FP64 updates with the u_a/u_b ping-pong of the reference's time
 * loop (1-openmp.c:193-205).  With tune->kblock = K > 1, every K steps are
 * one h3d_pass_f64 and the buffers swap once per pass.  A remainder of r
 * steps is one shorter pass; without rhs and for K = 2 or 3, when the run
 * has at least r passes and h3d_autotune has timed depths K and K+1 on
 * this grid (it does so when it picks K), the last r passes are K+1 steps
 * deep instead (the tuner's K+1 shape) if by those times that saves more
 * than half a K-step pass, one pass fewer.  A 1-step remainder runs as a
 * single step on the single-step shape h3d_autotune found on this grid
 * (with fixed_kblock 1), else on that kernel's default shape.  The bits
 * are those of `steps` single steps either way.  graph != 0 captures the
 * run's passes (up to 1024 of them; longer runs replay a graph of 1024
 * K-step passes) in a hipGraph and launches it.  *ms receives
 * the device time of the loop measured with HIP events on `stream`;
 * *final_in_b is 1 if the result is in u_b (an odd number of passes). */
int h3d_run_f64(double *u_a, double *u_b, const double *rhs,
                const double *kappa, const h3d_grid *g, double factor,
                int steps, const h3d_tune *tune, int graph, void *stream,
                float *ms, int *final_in_b);

/* The passes h3d_run_f64 launches for `steps` with this tune on grid g
 * (has_rhs: whether rhs is passed) -> *passes, of which *deep_passes
 * (may be NULL) are one step deeper than tune->kblock.  No reference
 * counterpart. */
int h3d_run_f64_passes(const h3d_grid *g, int has_rhs, int steps,
                       const h3d_tune *tune, int *passes, int *deep_passes);

/* Face pack/unpack: the whole XY plane k including its x/y ring, written
 * to/read from a dense [j][i] buffer of nx*ny values -- the order of the
 * reference's plane transfer (2-openmp.c:321,
 * 4-openmp-stream-omp-p2p.c:609-616).  Bit-exact copies. */
int h3d_pack_face_f64(double *buf, const double *u, const h3d_grid *g, int k,
                      void *stream);
int h3d_unpack_face_f64(double *u, const double *buf, const h3d_grid *g,
                        int k, void *stream);
int h3d_pack_face_f32(float *buf, const float *u, const h3d_grid *g, int k,
                      void *stream);
int h3d_unpack_face_f32(float *u, const float *buf, const h3d_grid *g, int k,
                        void *stream);
/* Band pack/unpack: planes k .. k+count-1 to/from a dense [k][j][i] buffer
 * of count*nx*ny values, the exchange unit of slabs with `count` ghost
 * planes per side (h3d_slab_pass_f64).  No reference counterpart: the
 * reference moves one plane per step (4-openmp-stream-omp-p2p.c:609-616). */
int h3d_pack_band_f64(double *buf, const double *u, const h3d_grid *g, int k,
                      int count, void *stream);
int h3d_unpack_band_f64(double *u, const double *buf, const h3d_grid *g, int k,
                        int count, void *stream);
/* The same for FP32 fields and for the FP16 residual field of
 * H3D_F32_COMP16 (16-bit words). */
int h3d_pack_band_f32(float *buf, const float *u, const h3d_grid *g, int k,
                      int count, void *stream);
int h3d_unpack_band_f32(float *u, const float *buf, const h3d_grid *g, int k,
                        int count, void *stream);
int h3d_pack_band_u16(uint16_t *buf, const uint16_t *u, const h3d_grid *g,
                      int k, int count, void *stream);
int h3d_unpack_band_u16(uint16_t *u, const uint16_t *buf, const h3d_grid *g,
                        int k, int count, void *stream);

/* One step of one z-slab (nz >= 4 planes: 2 ghosts + owned planes) in the
 * shape of the reference's middle devices (4-openmp-stream-omp-p2p.c:
 * 655-727): boundary planes 1 and nz-2, interior 2..nz-3, then the faces
 * pushed into the ghost planes of u_new through face_lo / face_hi (device
 * buffers of nx*ny values).  With one GPU the slab is its own neighbour:
 * plane nz-2 lands in ghost 0 and plane 1 in ghost nz-1 (z-periodic
 * stand-in for the peer exchange, same kernels and bytes). */
int h3d_slab_step_f64(double *u_new, const double *u_old, const double *rhs,
                      const double *kappa, const h3d_grid *g, double factor,
                      double *face_lo, double *face_hi, const h3d_tune *tune,
                      void *stream);
/* `steps` slab steps with u_a/u_b ping-pong; the interior kernel runs on
 * a second stream overlapped with the boundary planes and face copies (as
 * the reference overlaps its interior kernel with the exchange); graph != 0
 * replays a captured pair of steps.  stream must be a created stream. */
int h3d_run_slab_f64(double *u_a, double *u_b, const double *rhs,
                     const double *kappa, const h3d_grid *g, double factor,
                     double *face_lo, double *face_hi, int steps,
                     const h3d_tune *tune, int graph, void *stream, float *ms,
                     int *final_in_b);

/* A slab with `ghost` (1..4) ghost planes per side: owned planes
 * ghost .. nz-1-ghost, at least `ghost` of them (nz >= 3*ghost, nz >= 4).
 * One pass advances the owned planes by nsteps (1..ghost) steps with one
 * fused pass (h3d_pass_f64), the bands of `ghost` planes next to each side
 * first, then pushes the bands into the ghosts of u_new through face_lo /
 * face_hi (device buffers of ghost*nx*ny values): owned planes
 * nz-2*ghost .. nz-1-ghost into ghosts 0 .. ghost-1 and owned planes
 * ghost .. 2*ghost-1 into ghosts nz-ghost .. nz-1 (the z-periodic
 * stand-in of h3d_slab_step_f64, `ghost` planes deep).  Between passes
 * the ghosts must hold the neighbours' planes, as the exchange leaves
 * them; kappa's and rhs's ghost planes hold the neighbours' values.  The
 * same bytes cross per step as with one ghost plane, in 1/ghost of the
 * messages.  ghost = 1, nsteps = 1 is h3d_slab_step_f64.  No reference
 * counterpart: the reference exchanges one plane per step
 * (4-openmp-stream-omp-p2p.c:609-616). */
int h3d_slab_pass_f64(double *u_new, const double *u_old, const double *rhs,
                      const double *kappa, const h3d_grid *g, double factor,
                      int ghost, int nsteps, double *face_lo, double *face_hi,
                      const h3d_tune *tune, void *stream);
/* `steps` steps of such a slab as passes of `ghost` steps (the remainder
 * one shorter pass) with u_a/u_b ping-pong, swapping once per pass; the
 * owned planes between the bands run on a second stream; graph != 0
 * replays a captured pair of passes.  tune shapes those interior passes
 * (a fused-pass shape when ghost > 1).  stream must be a created stream.
 * ghost = 1 is h3d_run_slab_f64. */
int h3d_run_slab_pass_f64(double *u_a, double *u_b, const double *rhs,
                          const double *kappa, const h3d_grid *g,
                          double factor, int ghost, double *face_lo,
                          double *face_hi, int steps, const h3d_tune *tune,
                          int graph, void *stream, float *ms, int *final_in_b);

/* FP32 slabs with `ghost` (1..8) ghost planes per side: h3d_slab_pass_f64
 * with FP32 fields and the FP32 pass (h3d_pass_f32_range, mode as there).
 * In mode H3D_F32_COMP16 the residual field's bands move too, through
 * cface_lo / cface_hi (ghost*nx*ny 16-bit words each; NULL in the other
 * modes): a ghost plane's residual carries into the next pass as an owned
 * plane's does.  Owned planes after a pass equal those planes of the same
 * pass over the whole field, bit for bit.  No reference counterpart (the
 * reference is FP64). */
int h3d_slab_pass_f32(float *u_new, const float *u_old, const float *kappa_f32,
                      const uint16_t *kappa_bf16, uint16_t *comp_new,
                      const uint16_t *comp_old, const h3d_grid *g, float factor,
                      int mode, int ghost, int nsteps, float *face_lo,
                      float *face_hi, uint16_t *cface_lo, uint16_t *cface_hi,
                      const h3d_tune *tune, void *stream);
/* `steps` steps of such a slab as passes of `ghost` steps (the remainder
 * one shorter pass) with u_a/u_b (and c_a/c_b) ping-pong, as
 * h3d_run_slab_pass_f64; tune shapes the interior's full passes. */
int h3d_run_slab_pass_f32(float *u_a, float *u_b, const float *kappa_f32,
                          const uint16_t *kappa_bf16, uint16_t *c_a, uint16_t *c_b,
                          const h3d_grid *g, float factor, int mode, int ghost,
                          float *face_lo, float *face_hi, uint16_t *cface_lo,
                          uint16_t *cface_hi, int steps, const h3d_tune *tune,
                          int graph, void *stream, float *ms, int *final_in_b);

/* The `tune` of a slab's interior launch for the slab functions above,
 * timed on the slab pass itself (no reference counterpart).  kind 0: FP64
 * slab (kappa FP64; mode 0 without rhs, 1 with rhs; ghost 1..4, 1 for
 * h3d_run_slab_f64); kind 1: FP32 slab (kappa FP32, or BF16 bits when
 * kappa_bf16; mode as h3d_pass_f32; ghost 1..8; comp_a, the FP16 residual
 * field, and cface_lo / cface_hi in mode H3D_F32_COMP16).  The interior
 * launch is first tuned alone (h3d_autotune on a grid of exactly its
 * planes); then whole passes of `ghost` steps (bands, exchange and the
 * interior on a second stream) are timed for that shape and the
 * precision's 16-wave fused tiles, each with its candidate z chunks:
 * alone, the interior may prefer another shape than beside the bands.  Reads u_a, kappa and comp_a only (passes run on
 * scratch copies); face_lo / face_hi (and cface_*) are used as workspace.
 * stream must be a created stream.  Results are cached per (kind, grid,
 * mode, kappa format, ghost); any tune gives the same bits. */
int h3d_autotune_slab(int kind, const void *u_a, const void *kappa, int kappa_bf16,
                      const void *comp_a, const h3d_grid *g, int mode, int ghost,
                      void *face_lo, void *face_hi, void *cface_lo, void *cface_hi,
                      void *stream, h3d_tune *out);

/* FP64 diagnostics over the whole array (ring included), deterministic:
 * out[0] = sum u (checksum), out[1] = sum u^2, out[2] = max |u|,
 * out[3] = number of non-finite values.  Host `out`; synchronises stream. */
int h3d_stats_f64(const double *u, const h3d_grid *g, double *out,
                  void *stream);
int h3d_stats_f32(const float *u, const h3d_grid *g, double *out,
                  void *stream);

/* h3d_stats_f64 for callers that cannot allocate device memory (code under
 * HIP graph capture, serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  The partial records live in a fixed
 * device array of the library (1024 records of 48 bytes, one copy per GPU,
 * held until the process ends) instead of a stream-ordered allocation, and
 * the four results (same meaning and bits as h3d_stats_f64's out[]) are
 * written to the caller's DEVICE memory out_dev, ordered on `stream`,
 * without synchronising.  Calls that overlap on one GPU share the array
 * and race: order them on one stream.  H3D_ELIMIT if the field would need
 * more partial records than the array holds (h3d_stats_f64 uses at most
 * 1024 for any grid, so every valid grid fits). */
int h3d_stats_f64_noalloc(const double *u, const h3d_grid *g, double *out_dev,
                          void *stream);

/* Error metrics of compare.py:47-76 and the PASS rule of verify.sh:239-286,
 * computed on the device in FP64 over the whole array:
 * out[0] max_abs, [1] mean_abs, [2] rms, [3] rel_l2, [4] max_rel over
 * |ref| > 1e-12, [5] argmax (flat k,j,i index of the first max |d|),
 * [6] NaN count, [7] Inf count, [8] cells failing |d| <= atol + rtol|ref|
 * or non-finite.  cand is FP64 (_f64) or FP32 (_f32). */
int h3d_compare_f64(const double *ref, const double *cand, const h3d_grid *g,
                    double atol, double rtol, double *out, void *stream);
int h3d_compare_f32(const double *ref, const float *cand, const h3d_grid *g,
                    double atol, double rtol, double *out, void *stream);

/* max over interior cells of sum_{6 faces}(kappa_c + kappa_nb), in FP64;
 * the explicit scheme keeps the discrete maximum principle iff
 * dt <= 2 h^2 / result.  No reference counterpart (the reference fixes
 * dt = h^2/(12*0.95), heat3D_variable_coeff.c:80-82). */
int h3d_face_sum_max_f64(const double *kappa, const h3d_grid *g, double *out,
                         void *stream);

/* ---------------------------------------------------------------------- */
/* FP32 storage paths (no reference counterpart; the reference is FP64).   */

#define H3D_F32_PLAIN 0 /* u' = fma(factor, S, u)                          */
#define H3D_F32_CARRY 1 /* compensated; residual carried in registers
                           across the fused steps of one pass             */
#define H3D_F32_COMP16 2 /* compensated; residual also kept between passes
                           in an FP16 field scaled by the cell's ulp      */

/* One HBM pass of `nsteps` (1..8) fused time steps over the interior of
 * a grid stored in FP32.  kappa is exactly one of kappa_f32 / kappa_bf16
 * (BF16 bits of kappa/scale; pass factor*scale as `factor`).  comp_in /
 * comp_out (FP16 bits) are used in mode H3D_F32_COMP16 only.  Cell update:
 * S in the FP64 kernel's order in FP32, then Kahan's compensated sum with
 * the increment joining the residual in one rounding: y = fma(factor, S,
 * e), u' = u + y, e' = y - (u' - u).  u_out's ring must hold the Dirichlet
 * values. */
int h3d_pass_f32(float *u_out, const float *u_in, const float *kappa_f32,
                 const uint16_t *kappa_bf16, uint16_t *comp_out,
                 const uint16_t *comp_in, const h3d_grid *g, float factor,
                 int nsteps, int mode, const h3d_tune *tune, void *stream);
/* The same pass writing planes k_start..k_end only (1 <= k_start, k_end <=
 * nz-2): those planes of u_out (and comp_out) receive those planes of
 * nsteps steps applied to all of u_in, the bits of h3d_pass_f32; other
 * planes are not written (as h3d_pass_f64 with a plane range).  The
 * exchange unit of FP32 slabs with deep ghosts (h3d_slab_pass_f32). */
int h3d_pass_f32_range(float *u_out, const float *u_in, const float *kappa_f32,
                       const uint16_t *kappa_bf16, uint16_t *comp_out,
                       const uint16_t *comp_in, const h3d_grid *g, float factor,
                       int nsteps, int mode, int k_start, int k_end,
                       const h3d_tune *tune, void *stream);

/* `steps` FP32 updates as passes of tune->kblock steps (remainder last),
 * ping-ponging u_a/u_b (and c_a/c_b in COMP16 mode).  graph != 0 captures
 * the run's passes (up to 1024; longer runs replay a graph of 1024 passes)
 * in a hipGraph and launches it.  *ms: loop time by HIP events on
 * `stream`; *final_in_b: 1 if the result (and residual field) ended in
 * u_b (c_b). */
int h3d_run_f32(float *u_a, float *u_b, const float *kappa_f32,
                const uint16_t *kappa_bf16, uint16_t *c_a, uint16_t *c_b,
                const h3d_grid *g, float factor, int steps, int mode,
                const h3d_tune *tune, int graph, void *stream, float *ms,
                int *final_in_b);

/* Choose tune fields for a problem by timing candidate shapes on the
 * device (a few passes each, reading u_a and kappa and ping-ponging
 * through two scratch fields the tuner allocates and seeds from u_a; u_b
 * is not used); results are cached per
 * (kernel, grid, mode, kappa format, fixed_kblock).  kind: 0 = FP64 step
 * or pass (mode 0: no rhs, mode 1: with rhs, which the shapes' LDS must
 * hold; kblock chosen too, 1..4, unless fixed_kblock 1..4 is given), 1 =
 * FP32 pass (kblock chosen too, 1..4, unless fixed_kblock 1..8 is given).
 * The figure of merit is time per step.  The caller's buffers are not
 * modified. */
int h3d_autotune(int kind, void *u_a, void *u_b, const void *kappa,
                 int kappa_bf16, const h3d_grid *g, int mode,
                 int fixed_kblock, void *stream, h3d_tune *out);

/* ---------------------------------------------------------------------- */
/* Ensembles: many independent problems of uneven sizes, one persistent
 * launch per time step over a device-side tile work list (no reference
 * counterpart).                                                           */

typedef struct h3d_member {
  void *u_a, *u_b;   /* device: ping-pong fields, FP64 or FP32, ring seeded */
  const void *kappa; /* device: FP64 (precision 0) or FP32 (precision 1, 2) */
  h3d_grid g;        /* member grid                                         */
  double factor;     /* dt/h/h/2 of this member (FP32 members use it
                        rounded to float)                                   */
  int precision;     /* 0 = FP64 step (as h3d_step_f64); 1 = FP32 storage,
                        compensated, with an FP16 residual field kept by the
                        ensemble across passes (as h3d_pass_f32 with mode
                        H3D_F32_COMP16 and the ensemble's kblock); 2 = FP32
                        storage, compensated, the residual carried in
                        registers through each pass's fused steps (as
                        h3d_pass_f32 with mode H3D_F32_CARRY)               */
} h3d_member;

typedef struct h3d_ensemble h3d_ensemble;

/* Builds the tile work lists on `device_id` and allocates the FP32
 * members' residual fields (zeroed; precision 1 only).  Tiles are fixed in x and y (1024-
 * thread workgroups of 32 rows for FP64 members and 48 rows for FP32
 * members; one persistent kernel serves both).  tune may be NULL;
 * tune->zchunk, if non-zero, is the z-march
 * depth (planes per work item, default 64) and tune->kblock the
 * temporal-blocking depth (steps fused per pass of every member, 1..3,
 * default 1; H3D_ELIMIT above). */
int h3d_ensemble_create(h3d_ensemble **out, const h3d_member *members,
                        int count, const h3d_tune *tune, int device_id);
/* `steps` time steps of every member.  With kblock K > 1, each group of K
 * steps is one persistent launch of K-step passes for every member
 * (FP64 members bit-identical to h3d_pass_f64); steps past the last full
 * group are single-step launches.  graph != 0 captures the run's launches into one
 * hipGraph (kept and reused while the step count and starting parity
 * repeat) and launches that.  *ms is the device time of the steps.
 * Members ping-pong u_a -> u_b once per launch:
 * h3d_ensemble_where says which buffer holds a member's current field. */
int h3d_ensemble_run(h3d_ensemble *e, int steps, int graph, void *stream,
                     float *ms);
/* Times the ensemble's step for z-march depths 32..256 planes and pass
 * depths 1..3, and keeps the fastest per step
 * (written to *out if out != NULL).  The timed launches recompute the next
 * step into the idle buffers, so the members' state and the results of
 * later runs are unchanged. */
int h3d_ensemble_autotune(h3d_ensemble *e, void *stream, h3d_tune *out);
/* *in_b = 1 if member `member`'s current field is in its u_b, else 0. */
int h3d_ensemble_where(const h3d_ensemble *e, int member, int *in_b);
/* Work-list size, grid size and steps taken so far. */
int h3d_ensemble_info(const h3d_ensemble *e, int *nwork, int *nblocks,
                      long long *steps_done);
int h3d_ensemble_destroy(h3d_ensemble *e);

/* ---------------------------------------------------------------------- */
/* Runtime helpers (stand-ins for omp_target_alloc / omp_target_memcpy of
 * 1-openmp.c:166-179,214).                                                */

int h3d_device_count(int *count);
int h3d_target_alloc(void **ptr, size_t bytes, int device_id);
int h3d_target_free(void *ptr, int device_id);
/* Synchronous copy; dst_device/src_device = -1 for host memory. */
int h3d_target_memcpy(void *dst, const void *src, size_t bytes,
                      int dst_device, int src_device);
int h3d_device_synchronize(int device_id);
/* A non-blocking HIP stream on device_id (default priority), returned as
 * void*; h3d_stream_synchronize waits for all work queued on it. */
int h3d_stream_create(void **stream, int device_id);
int h3d_stream_destroy(void *stream);
int h3d_stream_synchronize(void *stream);
const char *h3d_error_string(int code);
/* Name of the GPU architecture of device_id (e.g. "gfx950") into buf. */
int h3d_device_arch(int device_id, char *buf, int buflen);

/* HBM bandwidth probe: device-to-device copy and triad (a = b + s*c) over
 * `bytes` per array, best of `reps` timed launches, in GB/s (1e9 B/s). */
int h3d_stream_probe(size_t bytes, int reps, int device_id, double *copy_gbs,
                     double *triad_gbs);

/* Host <-> device copy rates, as the reference's transfer benchmarks
 * measure them (Bandwidth_Measurements/AMD-MI250X/Synchronous-test-H-to-D-
 * Measurements/hip-1.hip:56-92, and the D-to-H / asynchronous twins): host
 * clock around blocking copies, best of `reps`.  out[0] H2D GB/s, out[1]
 * D2H GB/s for `bytes` (>= 4096); out[2] H2D, out[3] D2H microseconds of a
 * 4 KiB copy.  pinned != 0: hipHostMalloc'ed host buffer, else malloc'ed. */
int h3d_transfer_probe(size_t bytes, int reps, int pinned, int device_id,
                       double *out);

/* One host<->device transfer of `bytes` (a multiple of 8, >= 8) timed as
 * the reference's transfer benchmarks time it, with the data it carried
 * checked afterwards (Bandwidth_Measurements/AMD-MI250X/Synchronous-test-
 * H-to-D-Measurements/hip-1.hip:67-120 and its D-to-H twin;
 * Asynchronous-test-H-to-D-Measurements/hip-1.hip:79-98 for async = 1):
 * two warm-up copies, then one copy timed by the host clock around a
 * blocking hipMemcpy (async = 0) or by HIP events around hipMemcpyAsync on
 * a stream (async = 1).  direction 0 = host to device, 1 = device to host;
 * pinned != 0: hipHostMalloc'ed host buffer, else malloc'ed (the
 * reference's).  Every 64-bit word carries its own index-derived value;
 * all words are checked up to 1 GiB, above that the first and last 4096
 * and 4096 spread through the buffer.  out[0] GB/s, out[1] seconds,
 * out[2] words that arrived wrong (0 = verified), out[3] words checked. */
int h3d_transfer_bench(size_t bytes, int pinned, int async, int direction,
                       int device_id, double *out);

/* ---------------------------------------------------------------------- */
/* Host-side problem set-up (plain C; no device involved).                 */

/* heat3D_variable_coeff.c:79-84: h = 1/(n-1), dt = h^2/(12*0.95),
 * factor = dt/h/h/2. */
int h3d_params_reference(int n, double *h, double *dt, double *factor);
/* heat3D_variable_coeff.c:161-193: number of passes of
 * `while (t < T) { ...; t += dt; }`. */
int h3d_count_steps(double T, double dt);
/* heat3D_variable_coeff.c:110-150 / 1-openmp.c:153-163 on a grid with
 * nx = ny = nz = n (strides honoured).  rhs may be NULL. */
int h3d_init_reference(double *u, double *rhs, double *kappa,
                       const h3d_grid *g);
/* 4-openmp-stream-omp-p2p.c:208-223 (init_partition), dense slab. */
int h3d_init_partition(double *u, double *rhs, double *kappa, int nz_local,
                       int gk_start, int ng, int n, int ny, int nx, double h,
                       double dt);
/* 4-openmp-stream-omp-p2p.c:391-415 (parts = 4) / 2-openmp.c:294-316
 * (parts = 2): nz_d[p] planes incl. 2 ghosts, first owned global plane
 * gk_start[p].  H3D_EARG if a slab has fewer than 4 planes. */
int h3d_slab_decompose(int n, int parts, int *nz_d, int *gk_start);
/* 4-openmp-stream-omp-p2p.c:776-820: copies the planes slab `part` of a
 * `parts`-way split owns (global k = gk_start[part] .. gk_start[part+1]-1,
 * plus k = 0 for the first and k = n-1 for the last slab) from its dense
 * slab array (nz_d[part] planes of n*n) into the global n^3 array. */
int h3d_slab_reconstruct(double *global, const double *slab, int n, int parts,
                         int part);

/* Seeded inputs (no reference counterpart: the reference has no RNG).
 * Interior m^3 of a global grid of (m+2)^3 points, h = 1/(m+1), x = i*h.
 *   kappa_kind 0: kappa = (1 + 0.5 sin2pix sin2piy sin2piz) * U[0.9,1.1]
 *   kappa_kind 1: kappa = exp(0.5 Z), Z ~ N(0,1) (Box-Muller)
 *   u0 = exp(-|x-0.5|^2 / (2*0.1^2)) + U[0,1e-3] inside, 0 on the ring.
 * Random numbers are a counter-based hash of (seed, member, stream,
 * global linear index), so slabs and tiles of one problem generate the
 * same values.  This fills a block of nz planes starting at global plane
 * gk0, of a global problem whose axes are gnx, gny, gnz points (the block
 * covers all of x and y: g->nx == gnx, g->ny == gny).  Either output may
 * be NULL. */
int h3d_gen_seeded(double *u0, double *kappa, const h3d_grid *g, int gk0,
                   int gnx, int gny, int gnz, int kappa_kind, uint64_t seed,
                   int member);
/* The generator's uniform variate in [0,1) (exposed for testing). */
double h3d_rng_uniform(uint64_t seed, int member, int stream, uint64_t idx);

/* Host FP64 max over interior of sum_{6 faces}(kappa_c + kappa_nb). */
double h3d_face_sum_max_host(const double *kappa, const h3d_grid *g);

/* BF16 storage of kappa with a per-field power-of-two scale s: writes
 * bf16(kappa/s) (round to nearest even) and s; kappa is then represented
 * exactly by bf16_to_f32(bits)*s (use it, not the original, as the FP64
 * reference input).  n values. */
int h3d_quantize_bf16(const double *kappa, uint16_t *bits, size_t n,
                      double *scale);
/* The two halves of h3d_quantize_bf16, for fields split across slabs:
 * every slab must store its kappa with the whole field's scale, or its bits
 * differ from the whole field's.  h3d_bf16_scale gives the power-of-two
 * scale of n values (2^e for max|kappa| = m * 2^e, m in [0.5, 1); 1 for all
 * zeros); h3d_quantize_bf16_scaled writes bf16(kappa/scale) with a given
 * power-of-two scale (H3D_EARG for another scale, non-finite values or
 * values BF16 cannot hold). */
int h3d_bf16_scale(const double *kappa, size_t n, double *scale);
int h3d_quantize_bf16_scaled(const double *kappa, uint16_t *bits, size_t n, double scale);

/* `%.17g` one value per line, k,j,i order, all nz*ny*nx points
 * (1-openmp.c:41-51). */
int h3d_write_dump(const char *fname, const double *u, const h3d_grid *g);

#ifdef __cplusplus
}
#endif
#endif /* H3D_H */
