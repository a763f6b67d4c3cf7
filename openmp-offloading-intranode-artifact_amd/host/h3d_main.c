/*
 * h3d -- host program of the MI355X-native 3-D variable-coefficient heat
 * update.  Drop-in for the reference's `./bin n T` binaries
 * (OpenMP_Single_Source/1-openmp.c:87-241, HIP/1-hip.cpp:137-293):
 *
 *   h3d [n [T]] [options]
 *
 * n counts points per axis including the Dirichlet ring (default 51) and T
 * is the final time (default 1.0); the step count is the number of passes
 * of `while (t < T) t += dt` (heat3D_variable_coeff.c:161-193).  stdout
 * carries the lines the reference's pipeline parses ("Estimated
 * timesteps: ~N", "N timesteps complete", "Time = S seconds",
 * aggregate.sh:31-40) and WRITE_TXT selects the %.17g solution dump
 * (default on, as in 1-openmp.c:31-39), here u_h3d_1gpu.txt.
 *
 * Options (all new; the reference has none):
 *   --steps S          fixed step count instead of T
 *   --mode reference|seeded   reference fields (default) or seeded fields
 *   --kappa sine|lognormal    seeded kappa generator (default sine)
 *   --seed N           seed of the seeded generators (default 20260101)
 *   --precision f64|f32|f32c|f32c16
 *                      FP64 plain (default); FP32 storage plain; FP32
 *                      compensated with register carry; with FP16 residual
 *   --kappa-bf16       store kappa in BF16 with a per-field scale (FP32)
 *   --kblock K         time steps fused per HBM pass (FP32 1..8, FP64 1..4;
 *                      the results are the same bits as single steps)
 *   --tile R,W,Z       rows per lane, waves per workgroup, planes per chunk
 *   --autotune         time candidate launch shapes first and use the best
 *                      (the default when there are at least 128^3 interior
 *                      points and no --tile is given)
 *   --no-autotune      use the library's default launch shapes
 *   --graph            replay the step loop from a hipGraph
 *   --slab P:p         run slab p of a P-way z split (P = 2 or 4) with its
 *                      two ghost planes refreshed every step by in-HBM face
 *                      pack/unpack from the slab's own edge planes; with
 *                      --kblock K, K ghost planes per side refreshed by a
 *                      band of K planes after every K-step fused pass.
 *                      FP32 precisions (seeded mode) run passes of K steps
 *                      (1 without --kblock) with bands of K FP32 planes,
 *                      and of the FP16 residual too in f32c16
 *   --check            FP32 runs: also run the FP64 kernel (an FP64 slab
 *                      with the same exchange, for --slab) on the same
 *                      inputs and report rel-L2 / max errors against it
 *   --device D         HIP device (default 0)
 * Exit status: 0 on success, 1 on any error (message on stderr).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <strings.h>
#include <time.h>
#include <unistd.h>

#include "h3d.h"


#define CHECK(call)                                                        \
  do {                                                                     \
    int rc_ = (call);                                                      \
    if (rc_ != H3D_OK) {                                                   \
      fprintf(stderr, "ERROR: %s failed at %s:%d: %s (%d)\n", #call,       \
              __FILE__, __LINE__, h3d_error_string(rc_), rc_);             \
      exit(1);                                                             \
    }                                                                      \
  } while (0)

static int should_write_txt(void)
{
  const char *e = getenv("WRITE_TXT");
  if (!e) return 1;
  if (!strcmp(e, "0")) return 0;
  if (!strcasecmp(e, "false")) return 0;
  if (!strcasecmp(e, "no")) return 0;
  return 1;
}

static double wtime(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static void *xmalloc(size_t bytes)
{
  void *p = malloc(bytes ? bytes : 1);
  if (!p) {
    fprintf(stderr, "ERROR: malloc failed (host arrays, %zu bytes)\n", bytes);
    exit(1);
  }
  return p;
}

static void usage(void)
{
  fprintf(stderr,
          "usage: h3d [n [T]] [--steps S] [--mode reference|seeded] "
          "[--kappa sine|lognormal] [--seed N] [--precision f64|f32|f32c|f32c16] "
          "[--kappa-bf16] [--kblock K] [--tile R,W,Z] [--autotune|--no-autotune] [--graph] "
          "[--slab P:p] [--check] [--device D]\n");
  exit(1);
}

typedef struct opts {
  int n;
  double T;
  int steps; /* -1: from T */
  int seeded, kappa_kind;
  uint64_t seed;
  int prec; /* 0 f64, 1 f32 plain, 2 f32c carry, 3 f32c fp16 */
  int bf16, autotune, graph, check, device;
  int slab_parts, slab_index;
  h3d_tune tune;
} opts;

static opts parse(int argc, char **argv)
{
  opts o;
  memset(&o, 0, sizeof o);
  o.n = 51;
  o.T = 1.0;
  o.steps = -1;
  o.seed = 20260101ULL;
  o.autotune = -1; /* decided from the problem size */
  int pos = 0;
  for (int a = 1; a < argc; a++) {
    const char *s = argv[a];
    if (!strncmp(s, "--", 2)) {
      const char *v = (a + 1 < argc) ? argv[a + 1] : NULL;
      if (!strcmp(s, "--steps") && v) { o.steps = atoi(v); a++; }
      else if (!strcmp(s, "--mode") && v) {
        if (!strcmp(v, "reference")) o.seeded = 0;
        else if (!strcmp(v, "seeded")) o.seeded = 1;
        else usage();
        a++;
      } else if (!strcmp(s, "--kappa") && v) {
        if (!strcmp(v, "sine")) o.kappa_kind = 0;
        else if (!strcmp(v, "lognormal")) o.kappa_kind = 1;
        else usage();
        a++;
      } else if (!strcmp(s, "--seed") && v) { o.seed = strtoull(v, NULL, 10); a++; }
      else if (!strcmp(s, "--precision") && v) {
        if (!strcmp(v, "f64")) o.prec = 0;
        else if (!strcmp(v, "f32")) o.prec = 1;
        else if (!strcmp(v, "f32c")) o.prec = 2;
        else if (!strcmp(v, "f32c16")) o.prec = 3;
        else usage();
        a++;
      } else if (!strcmp(s, "--kappa-bf16")) o.bf16 = 1;
      else if (!strcmp(s, "--kblock") && v) { o.tune.kblock = atoi(v); a++; }
      else if (!strcmp(s, "--tile") && v) {
        if (sscanf(v, "%d,%d,%d", &o.tune.rows, &o.tune.waves, &o.tune.zchunk) < 2) usage();
        a++;
      } else if (!strcmp(s, "--autotune")) o.autotune = 1;
      else if (!strcmp(s, "--no-autotune")) o.autotune = 0;
      else if (!strcmp(s, "--graph")) o.graph = 1;
      else if (!strcmp(s, "--check")) o.check = 1;
      else if (!strcmp(s, "--slab") && v) {
        if (sscanf(v, "%d:%d", &o.slab_parts, &o.slab_index) != 2) usage();
        a++;
      } else if (!strcmp(s, "--device") && v) { o.device = atoi(v); a++; }
      else usage();
    } else if (pos == 0) { o.n = atoi(s); pos++; }
    else if (pos == 1) { o.T = atof(s); pos++; }
    else usage();
  }
  if (o.n < 3) {
    fprintf(stderr, "ERROR: n=%d: need at least 3 points per axis\n", o.n);
    exit(1);
  }
  if (o.bf16 && o.prec == 0) {
    fprintf(stderr, "ERROR: --kappa-bf16 needs an FP32 precision\n");
    exit(1);
  }
  if (o.slab_parts && o.tune.kblock > 4) {
    fprintf(stderr, "ERROR: --slab --kblock K: K ghost planes per side, 1..4\n");
    exit(1);
  }
  return o;
}

/* FP32 copy of an FP64 field. */
static float *to_f32(const double *src, size_t n)
{
  float *d = (float *)xmalloc(n * sizeof(float));
  for (size_t c = 0; c < n; c++) d[c] = (float)src[c];
  return d;
}

/* Device layout of every field: rows padded to a multiple of 32 elements
 * and the buffer offset so that interior point i = 1 of every row starts a
 * 128-byte line.  Host fields stay in the reference's dense layout (what
 * the initialisers write and the dump reads); the copies below convert.
 * The kernels read and write whole lines this way: at n = 1024 the FP64
 * step runs 27% faster than on the dense layout, whose rows start 8 bytes
 * into a line. */
typedef struct {
  void *base; /* allocation (freed) */
  void *p;    /* aligned field       */
} dfield;

static int pitch_of(int n) { return (n + 31) / 32 * 32; }

static dfield dev_alloc(const h3d_grid *gd, size_t esz, int dev)
{
  dfield f;
  CHECK(h3d_target_alloc(&f.base, (size_t)gd->sz * gd->nz * esz + 256, dev));
  const uintptr_t b = (uintptr_t)f.base;
  f.p = (void *)(b + ((128 - (b + esz) % 128) % 128));
  return f;
}

/* dense host field (nx x ny x nz, elements of esz bytes) -> device layout */
static void dev_upload(dfield f, const void *host, const h3d_grid *gd, size_t esz, int dev)
{
  const size_t bytes = (size_t)gd->sz * gd->nz * esz;
  char *stage = (char *)calloc(bytes ? bytes : 1, 1);
  if (!stage) {
    fprintf(stderr, "ERROR: malloc failed (staging, %zu bytes)\n", bytes);
    exit(1);
  }
  const size_t row = (size_t)gd->nx * esz;
  for (size_t r = 0; r < (size_t)gd->ny * gd->nz; r++)
    memcpy(stage + r * (size_t)gd->sy * esz, (const char *)host + r * row, row);
  CHECK(h3d_target_memcpy(f.p, stage, bytes, dev, -1));
  free(stage);
}

/* device layout -> dense host field */
static void dev_download(void *host, const void *dptr, const h3d_grid *gd, size_t esz, int dev)
{
  const size_t bytes = (size_t)gd->sz * gd->nz * esz;
  char *stage = (char *)xmalloc(bytes);
  CHECK(h3d_target_memcpy(stage, dptr, bytes, -1, dev));
  const size_t row = (size_t)gd->nx * esz;
  for (size_t r = 0; r < (size_t)gd->ny * gd->nz; r++)
    memcpy((char *)host + r * row, stage + r * (size_t)gd->sy * esz, row);
  free(stage);
}

int main(int argc, char **argv)
{
  setvbuf(stdout, NULL, _IOLBF, 0);
  opts o = parse(argc, argv);
  const int n = o.n;
  printf("\n**** ONE GPU HIP (h3d, MI355X) STARTS ****\n");

  /* ---- slab geometry -------------------------------------------------- */
  int nz = n, gk0 = 0; /* planes held and first held global plane */
  /* a slab holds `ghost` ghost planes per side and advances its owned
   * planes `ghost` steps per fused pass between band exchanges */
  const int ghost = o.slab_parts ? (o.tune.kblock > 1 ? o.tune.kblock : 1) : 0;
  if (o.slab_parts) {
    int nz_d[4], gk_start[4];
    if (o.slab_index < 0 || o.slab_index >= o.slab_parts ||
        h3d_slab_decompose(n, o.slab_parts, nz_d, gk_start) != H3D_OK) {
      fprintf(stderr, "ERROR: n=%d cannot be split %d ways (each slab needs >= 4 planes)\n",
              n, o.slab_parts);
      return 1;
    }
    printf("Block sizes:");
    for (int p = 0; p < o.slab_parts; p++) printf(" nz%d=%d", p, nz_d[p]);
    printf("\n");
    nz = nz_d[o.slab_index] - 2 + 2 * ghost;
    gk0 = gk_start[o.slab_index] - ghost; /* first ghost plane */
    if (nz < 3 * ghost) {
      fprintf(stderr, "ERROR: a slab of %d owned planes cannot hold %d ghost planes per side\n",
              nz_d[o.slab_index] - 2, ghost);
      return 1;
    }
    if (ghost > 1) printf("Ghost planes per side: %d (band exchange every %d steps)\n", ghost, ghost);
  }
  const h3d_grid g = {n, n, nz, n, (long long)n * n}; /* host fields */
  const int sy = pitch_of(n);
  const h3d_grid gd = {n, n, nz, sy, (long long)sy * n}; /* device fields */
  const size_t cells = (size_t)nz * n * n;
  const size_t total_points = (size_t)n * n * n;

  printf("\n=== MEMORY ESTIMATE (GLOBAL) ===\n");
  printf("Grid size: %d x %d x %d = %zu points\n", n, n, n, total_points);
  printf("Memory per global array: %.3f GB\n", total_points * 8 / 1e9);
  printf("================================\n\n");

  /* ---- fields and parameters ----------------------------------------- */
  double *u0 = (double *)xmalloc(cells * sizeof(double));
  double *kap = (double *)xmalloc(cells * sizeof(double));
  double *rhs = NULL;
  double h, dt, factor, kappa_max;
  double kscale_field = 0.0; /* BF16 kappa scale of the whole field (slabs) */
  if (!o.seeded) {
    CHECK(h3d_params_reference(n, &h, &dt, &factor));
    kappa_max = 0.95;
    rhs = (double *)xmalloc(cells * sizeof(double));
    if (o.slab_parts)
      CHECK(h3d_init_partition(u0, rhs, kap, nz, gk0 + ghost, ghost, n, n, n, h, dt));
    else
      CHECK(h3d_init_reference(u0, rhs, kap, &g));
  } else {
    /* planes of a deep slab beyond the global domain (first and last slab)
     * repeat the nearest domain plane; they are ghosts, which the first
     * band exchange overwrites */
    const int lo = gk0 < 0 ? 0 : gk0, hi = gk0 + nz > n ? n : gk0 + nz;
    const size_t plane = (size_t)n * n;
    const h3d_grid gin = {n, n, hi - lo, n, (long long)plane};
    CHECK(h3d_gen_seeded(u0 + (size_t)(lo - gk0) * plane, kap + (size_t)(lo - gk0) * plane,
                         &gin, lo, n, n, n, o.kappa_kind, o.seed, 0));
    for (int k = 0; k < nz; k++) {
      const int src = k < lo - gk0 ? lo - gk0 : (k >= hi - gk0 ? hi - gk0 - 1 : k);
      if (src == k) continue;
      memcpy(u0 + (size_t)k * plane, u0 + (size_t)src * plane, plane * sizeof(double));
      memcpy(kap + (size_t)k * plane, kap + (size_t)src * plane, plane * sizeof(double));
    }
    h = 1.0 / (n - 1);
    double fmax;
    if (o.slab_parts) {
      /* the bound must be global: every slab of the split uses one dt */
      double *kg = (double *)xmalloc(total_points * sizeof(double));
      const h3d_grid gg = {n, n, n, n, (long long)n * n};
      CHECK(h3d_gen_seeded(NULL, kg, &gg, 0, n, n, n, o.kappa_kind, o.seed, 0));
      fmax = h3d_face_sum_max_host(kg, &gg);
      /* and so must BF16 kappa's scale, or slabs' bits differ from the
       * whole field's */
      if (o.bf16) CHECK(h3d_bf16_scale(kg, total_points, &kscale_field));
      free(kg);
    } else {
      fmax = h3d_face_sum_max_host(kap, &g);
    }
    dt = 0.9 * 2.0 * h * h / fmax;
    factor = dt / h / h / 2.0;
    kappa_max = fmax / 12.0;
  }

  int steps = o.steps >= 0 ? o.steps : h3d_count_steps(o.T, dt);
  double T = o.steps >= 0 ? steps * dt : o.T;
  printf("=== SIMULATION PARAMETERS ===\n");
  printf("Grid size (n):           %d\n", n);
  printf("Final time (T):          %.6f\n", T);
  printf("Grid spacing (h):        %.6e\n", h);
  printf("Time step (dt):          %.6e\n", dt);
  printf("Kappa_max:               %.6f\n", kappa_max);
  printf("Diffusion factor:        %.6e\n", factor);
  printf("Estimated timesteps:     ~%d\n", o.steps >= 0 ? steps : (int)(T / dt));
  printf("CFL condition (dt/h^2):  %.6f (should be < 0.5)\n", dt / (h * h));
  printf("Mode:                    %s%s, precision %s%s, seed %llu\n",
         o.seeded ? "seeded" : "reference",
         o.seeded ? (o.kappa_kind ? " (lognormal kappa)" : " (sine kappa)") : "",
         (const char *[]){"f64", "f32", "f32c", "f32c16"}[o.prec],
         o.bf16 ? " + bf16 kappa" : "", (unsigned long long)o.seed);
  printf("==============================\n\n");

  int ndev = 0;
  CHECK(h3d_device_count(&ndev));
  if (ndev < 1 || o.device >= ndev) {
    fprintf(stderr, "ERROR: this version needs 1 HIP device, %d present\n", ndev);
    return 1;
  }
  char arch[64] = "";
  CHECK(h3d_device_arch(o.device, arch, sizeof arch));
  printf("Using HIP device %d of %d (%s)\n", o.device, ndev, arch);
  void *stream = NULL;
  CHECK(h3d_stream_create(&stream, o.device));

  /* ---- device buffers ------------------------------------------------- */
  const int f32 = o.prec != 0;
  const size_t esz = f32 ? sizeof(float) : sizeof(double);
  dfield fa = dev_alloc(&gd, esz, o.device), fb = dev_alloc(&gd, esz, o.device);
  dfield fk = {NULL, NULL}, fr = {NULL, NULL}, fca = {NULL, NULL}, fcb = {NULL, NULL};
  void *d_ua = fa.p, *d_ub = fb.p, *d_kap = NULL, *d_rhs = NULL, *d_ca = NULL, *d_cb = NULL;
  double kscale = 1.0;
  float *h_u32 = NULL;
  if (!f32) {
    fk = dev_alloc(&gd, 8, o.device);
    d_kap = fk.p;
    dev_upload(fa, u0, &gd, 8, o.device);
    dev_upload(fb, u0, &gd, 8, o.device);
    dev_upload(fk, kap, &gd, 8, o.device);
    if (rhs) {
      fr = dev_alloc(&gd, 8, o.device);
      d_rhs = fr.p;
      dev_upload(fr, rhs, &gd, 8, o.device);
    }
  } else {
    if (rhs) {
      fprintf(stderr, "ERROR: FP32 precisions run seeded mode only (--mode seeded)\n");
      return 1;
    }
    h_u32 = to_f32(u0, cells);
    dev_upload(fa, h_u32, &gd, 4, o.device);
    dev_upload(fb, h_u32, &gd, 4, o.device);
    if (o.bf16) {
      uint16_t *kb = (uint16_t *)xmalloc(cells * 2);
      if (kscale_field > 0.0) {
        kscale = kscale_field;
        CHECK(h3d_quantize_bf16_scaled(kap, kb, cells, kscale));
      } else {
        CHECK(h3d_quantize_bf16(kap, kb, cells, &kscale));
      }
      /* the FP64 check runs on exactly the kappa the FP32 run sees */
      for (size_t c = 0; c < cells; c++) {
        uint32_t w = (uint32_t)kb[c] << 16;
        float f;
        memcpy(&f, &w, 4);
        kap[c] = (double)f * kscale;
      }
      fk = dev_alloc(&gd, 2, o.device);
      d_kap = fk.p;
      dev_upload(fk, kb, &gd, 2, o.device);
      free(kb);
    } else {
      float *k32 = to_f32(kap, cells);
      for (size_t c = 0; c < cells; c++) kap[c] = (double)k32[c];
      fk = dev_alloc(&gd, 4, o.device);
      d_kap = fk.p;
      dev_upload(fk, k32, &gd, 4, o.device);
      free(k32);
    }
    if (o.prec == 3) {
      fca = dev_alloc(&gd, 2, o.device);
      fcb = dev_alloc(&gd, 2, o.device);
      d_ca = fca.p;
      d_cb = fcb.p;
      uint16_t *z = (uint16_t *)calloc(cells, 2);
      if (!z) {
        fprintf(stderr, "ERROR: malloc failed (residual fields)\n");
        return 1;
      }
      dev_upload(fca, z, &gd, 2, o.device);
      dev_upload(fcb, z, &gd, 2, o.device);
      free(z);
    }
  }
  const int mode = o.prec == 1 ? H3D_F32_PLAIN : o.prec == 2 ? H3D_F32_CARRY : H3D_F32_COMP16;
  const float f32factor = (float)(factor * kscale);

  h3d_tune tune = o.tune;
  if (f32 && tune.kblock == 0) tune.kblock = 2;
  if (o.autotune < 0) {
    /* tuning costs tens of launches: worth it for long runs on large grids,
     * and the shapes give bit-identical results, so only speed changes */
    const double interior = (double)(gd.nx - 2) * (gd.ny - 2) * (gd.nz - 2);
    o.autotune = interior >= 128.0 * 128.0 * 128.0 && o.tune.rows == 0 && o.tune.waves == 0;
  }

  /* slab: face (band) buffers for the in-HBM exchange, sized for FP64;
   * the FP16 residual's bands in f32c16 */
  double *d_face_lo = NULL, *d_face_hi = NULL;
  uint16_t *d_cface_lo = NULL, *d_cface_hi = NULL;
  if (o.slab_parts) {
    CHECK(h3d_target_alloc((void **)&d_face_lo, (size_t)ghost * n * n * 8, o.device));
    CHECK(h3d_target_alloc((void **)&d_face_hi, (size_t)ghost * n * n * 8, o.device));
    if (o.prec == 3) {
      CHECK(h3d_target_alloc((void **)&d_cface_lo, (size_t)ghost * n * n * 2, o.device));
      CHECK(h3d_target_alloc((void **)&d_cface_hi, (size_t)ghost * n * n * 2, o.device));
    }
  }
  if (o.autotune) {
    h3d_tune t;
    /* FP64: mode 1 tunes shapes that hold rhs.  A slab runs passes of as
     * many steps as it has ghost planes per side, its interior beside the
     * bands: tuned on the slab pass itself */
    if (o.slab_parts)
      CHECK(h3d_autotune_slab(f32 ? 1 : 0, d_ua, d_kap, o.bf16, d_ca, &gd,
                              f32 ? mode : (rhs ? 1 : 0), ghost, d_face_lo, d_face_hi,
                              d_cface_lo, d_cface_hi, stream, &t));
    else
      CHECK(h3d_autotune(f32 ? 1 : 0, d_ua, d_ub, d_kap, o.bf16, &gd,
                         f32 ? mode : (rhs ? 1 : 0), o.tune.kblock, stream, &t));
    tune = t;
    printf("Autotuned: rows=%d waves=%d zchunk=%d kblock=%d\n", tune.rows,
           tune.waves, tune.zchunk, tune.kblock);
  }


  printf("\n=== SIMULATION STARTS ===\n");
  CHECK(h3d_device_synchronize(o.device));
  double timer = -wtime();
  int in_b = 0;
  float dev_ms = 0.0f;
  if (!f32 && o.slab_parts) {
    /* middle-slab shape (2-openmp.c:25-55): boundary planes and the face
     * exchange overlapped with the interior on a second stream */
    if (ghost == 1)
      CHECK(h3d_run_slab_f64(d_ua, d_ub, d_rhs, d_kap, &gd, factor, d_face_lo,
                             d_face_hi, steps, &tune, o.graph, stream, &dev_ms,
                             &in_b));
    else /* K ghost planes: K-step passes, one band exchange per pass */
      CHECK(h3d_run_slab_pass_f64(d_ua, d_ub, d_rhs, d_kap, &gd, factor, ghost,
                                  d_face_lo, d_face_hi, steps, &tune, o.graph,
                                  stream, &dev_ms, &in_b));
  } else if (!f32 && (o.graph || tune.kblock > 1)) {
    /* replayed from a hipGraph, or passes of tune.kblock fused steps */
    CHECK(h3d_run_f64(d_ua, d_ub, d_rhs, d_kap, &gd, factor, steps, &tune, o.graph,
                      stream, &dev_ms, &in_b));
  } else if (!f32) {
    /* the reference's loop (1-openmp.c:193-205): one update, then swap */
    double *uo = d_ua, *un = d_ub;
    for (int s = 0; s < steps; s++) {
      CHECK(h3d_step_f64(un, uo, d_rhs, d_kap, &gd, factor, 1, nz - 2, &tune, stream));
      double *t = uo; uo = un; un = t;
    }
    in_b = (uo == d_ub);
  } else if (o.slab_parts) {
    /* FP32 slab: passes of `ghost` steps, one band exchange per pass */
    CHECK(h3d_run_slab_pass_f32(d_ua, d_ub, o.bf16 ? NULL : d_kap, o.bf16 ? d_kap : NULL,
                                d_ca, d_cb, &gd, f32factor, mode, ghost, (float *)d_face_lo,
                                (float *)d_face_hi, d_cface_lo, d_cface_hi, steps, &tune,
                                o.graph, stream, &dev_ms, &in_b));
  } else {
    CHECK(h3d_run_f32(d_ua, d_ub, o.bf16 ? NULL : d_kap, o.bf16 ? d_kap : NULL,
                      d_ca, d_cb, &gd, f32factor, steps, mode, &tune, o.graph,
                      stream, &dev_ms, &in_b));
  }
  CHECK(h3d_stream_synchronize(stream));
  timer += wtime();

  printf("\n%d timesteps complete\n", steps);
  printf("Time = %f seconds\n", timer);
  printf("=== SIMULATION ENDS ===\n");

  const double lups = (double)(n - 2) * (n - 2) * (nz - 2 * (ghost ? ghost : 1)) * steps;
  double bpl = f32 ? (o.bf16 ? 10.0 : 12.0) : (rhs ? 32.0 : 24.0);
  if (o.prec == 3) bpl += 4.0;
  bpl /= o.slab_parts ? ghost : (tune.kblock > 0 ? tune.kblock : 1);
  if (timer > 0)
    printf("Performance: %.3f GLUP/s, %.1f GB/s algorithmic (%.2f B/LUP)\n",
           lups / timer / 1e9, lups * bpl / timer / 1e9, bpl);

  /* ---- diagnostics (FP64 reductions on the device) -------------------- */
  void *d_res = in_b ? d_ub : d_ua;
  double st[4];
  if (f32) CHECK(h3d_stats_f32((const float *)d_res, &gd, st, stream));
  else CHECK(h3d_stats_f64((const double *)d_res, &gd, st, stream));
  const double h3 = h * h * h;
  printf("checksum = %.17g  L2 = %.17g  max|u| = %.17g  energy = %.17g  nonfinite = %.0f\n",
         st[0], sqrt(st[1] * h3), st[2], 0.5 * st[1] * h3, st[3]);

  if (f32 && o.check) {
    /* FP64 kernel on the same (rounded) inputs as the reference answer */
    dfield r0 = dev_alloc(&gd, 8, o.device), r1 = dev_alloc(&gd, 8, o.device),
           rk = dev_alloc(&gd, 8, o.device);
    for (size_t c = 0; c < cells; c++) u0[c] = (double)h_u32[c];
    dev_upload(r0, u0, &gd, 8, o.device);
    dev_upload(r1, u0, &gd, 8, o.device);
    dev_upload(rk, kap, &gd, 8, o.device);
    int rb = 0;
    float rms = 0;
    if (o.slab_parts) /* the same slab, stand-in exchange and pass length */
      CHECK(h3d_run_slab_pass_f64(r0.p, r1.p, NULL, rk.p, &gd, factor, ghost, d_face_lo,
                                  d_face_hi, steps, NULL, 0, stream, &rms, &rb));
    else
      CHECK(h3d_run_f64(r0.p, r1.p, NULL, rk.p, &gd, factor, steps, NULL, 0, stream, &rms, &rb));
    double m[9];
    CHECK(h3d_compare_f32(rb ? r1.p : r0.p, (const float *)d_res, &gd, 1e-6, 1e-5, m, stream));
    printf("vs FP64: rel_l2 = %.3e  max_abs = %.3e  max_rel = %.3e  rms = %.3e  %s\n",
           m[3], m[0], m[4], m[2], m[8] == 0 ? "PASS" : "FAIL");
    h3d_target_free(r0.base, o.device);
    h3d_target_free(r1.base, o.device);
    h3d_target_free(rk.base, o.device);
  }

  if (should_write_txt()) {
    double *out = (double *)xmalloc(cells * sizeof(double));
    if (f32) {
      float *o32 = (float *)xmalloc(cells * sizeof(float));
      dev_download(o32, d_res, &gd, 4, o.device);
      for (size_t c = 0; c < cells; c++) out[c] = o32[c];
      free(o32);
    } else {
      dev_download(out, d_res, &gd, 8, o.device);
    }
    char name[64];
    if (o.slab_parts)
      snprintf(name, sizeof name, "u_h3d_slab%dof%d.txt", o.slab_index, o.slab_parts);
    else
      snprintf(name, sizeof name, "u_h3d_1gpu.txt");
    CHECK(h3d_write_dump(name, out, &g));
    free(out);
  }

  h3d_target_free(fa.base, o.device);
  h3d_target_free(fb.base, o.device);
  h3d_target_free(fk.base, o.device);
  if (fr.base) h3d_target_free(fr.base, o.device);
  if (fca.base) h3d_target_free(fca.base, o.device);
  if (fcb.base) h3d_target_free(fcb.base, o.device);
  if (d_face_lo) h3d_target_free(d_face_lo, o.device);
  if (d_face_hi) h3d_target_free(d_face_hi, o.device);
  if (d_cface_lo) h3d_target_free(d_cface_lo, o.device);
  if (d_cface_hi) h3d_target_free(d_cface_hi, o.device);
  h3d_stream_destroy(stream);
  free(u0);
  free(kap);
  free(rhs);
  free(h_u32);
  printf("\n**** ONE GPU HIP (h3d, MI355X) FINISHED ****\n");
  return 0;
}
