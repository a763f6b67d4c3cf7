/*
 * Host-side problem set-up of the h3d library: reference parameters and
 * fields, the step-count rule, z-slab decomposition, seeded generators,
 * the FP64 time-step bound, BF16 storage of kappa and the solution dump.
 *
 * Compiled with -ffp-contract=off: the reference-mode fields must round
 * exactly as the reference's initialisation does, because the GPU path is
 * checked bit for bit against it.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "h3d.h"

int h3d_params_reference(int n, double *h, double *dt, double *factor)
{
  if (n < 3 || !h || !dt || !factor) return H3D_EARG;
  /* heat3D_variable_coeff.c:79-84 (identical in 1-openmp.c:103-106) */
  double hh = 1.0 / (n - 1);
  double kappa_max = 0.95;
  double d = hh * hh / (12.0 * kappa_max);
  *h = hh;
  *dt = d;
  *factor = d / hh / hh / 2.0; /* left to right, as the reference writes it */
  return H3D_OK;
}

int h3d_count_steps(double T, double dt)
{
  /* heat3D_variable_coeff.c:161-193: the loop test is on the accumulated
   * FP64 time, so T = STEPS*dt printed with %.7g gives STEPS or STEPS+1. */
  if (!(dt > 0.0)) return 0;
  double t = 0.0;
  int steps = 0;
  while (t < T) {
    steps++;
    t += dt;
  }
  return steps;
}

int h3d_init_reference(double *u, double *rhs, double *kappa,
                       const h3d_grid *g)
{
  if (!u || !kappa || !g) return H3D_EARG;
  int n = g->nx;
  if (g->ny != n || g->nz != n || n < 3 || g->sy < n ||
      g->sz < (long long)g->sy * n)
    return H3D_EARG;
  double h, dt, factor;
  h3d_params_reference(n, &h, &dt, &factor);
  /* heat3D_variable_coeff.c:110-150 / 1-openmp.c:153-163 */
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        size_t idx = (size_t)k * (size_t)g->sz + (size_t)j * (size_t)g->sy + i;
        double sx = sin(M_PI * i * h), sy = sin(M_PI * j * h),
               sk = sin(M_PI * k * h);
        if (rhs)
          rhs[idx] = dt * (1.0 + cos(M_PI * i * h) * cos(M_PI * j * h) *
                                     cos(M_PI * k * h));
        kappa[idx] = 0.5 + 0.45 * sx * sy * sk;
        u[idx] = (i == 0 || i == n - 1 || j == 0 || j == n - 1 || k == 0 ||
                  k == n - 1)
                     ? 0.0
                     : sx * sy * sk;
      }
  return H3D_OK;
}

int h3d_init_partition(double *u, double *rhs, double *kappa, int nz_local,
                       int gk_start, int ng, int n, int ny, int nx, double h,
                       double dt)
{
  if (!u || !kappa || nz_local < 1 || nx < 1 || ny < 1) return H3D_EARG;
  /* 4-openmp-stream-omp-p2p.c:208-223: global plane gk = k - ng + gk_start */
  for (int k = 0; k < nz_local; k++) {
    int gk = (k - ng) + gk_start;
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        size_t idx = ((size_t)k * ny + j) * nx + i;
        double sx = sin(M_PI * i * h), sy = sin(M_PI * j * h),
               sk = sin(M_PI * gk * h);
        u[idx] = (gk <= 0 || gk >= n - 1 || j <= 0 || j >= n - 1 || i <= 0 ||
                  i >= n - 1)
                     ? 0.0
                     : sx * sy * sk;
        if (rhs)
          rhs[idx] = dt * (1.0 + cos(M_PI * i * h) * cos(M_PI * j * h) *
                                     cos(M_PI * gk * h));
        kappa[idx] = 0.5 + 0.45 * sx * sy * sk;
      }
  }
  return H3D_OK;
}

int h3d_slab_decompose(int n, int parts, int *nz_d, int *gk_start)
{
  if (!nz_d || !gk_start || n < 3) return H3D_EARG;
  const int ng = 1, nz_int = n - 2;
  if (parts == 1) {
    nz_d[0] = n;
  } else if (parts == 2) {
    /* 2-openmp.c:294-303: remainder to the higher device */
    int mid = nz_int / 2;
    nz_d[0] = mid + 2 * ng;
    nz_d[1] = (nz_int - mid) + 2 * ng;
  } else if (parts == 4) {
    /* 4-openmp-stream-omp-p2p.c:391-405: remainder to the lower devices */
    int q = nz_int / 4, r = nz_int % 4;
    for (int p = 0; p < 4; p++) nz_d[p] = q + (r > p ? 1 : 0) + 2 * ng;
  } else {
    return H3D_EARG;
  }
  gk_start[0] = 1;
  for (int p = 1; p < parts; p++)
    gk_start[p] = gk_start[p - 1] + (nz_d[p - 1] - 2 * ng);
  if (parts > 1)
    for (int p = 0; p < parts; p++)
      if (nz_d[p] < 4) return H3D_EARG;
  return H3D_OK;
}

/* 4-openmp-stream-omp-p2p.c:776-820: slab `part` contributes its owned
 * planes to the global [k][j][i] array; interface ghosts are left to their
 * owners, and the first and last slab also contribute the physical
 * boundary planes k = 0 and k = n-1. */
int h3d_slab_reconstruct(double *global, const double *slab, int n, int parts,
                         int part)
{
  int nz_d[4], gk_start[4];
  if (!global || !slab || parts < 1 || parts > 4 || part < 0 || part >= parts)
    return H3D_EARG;
  int rc = h3d_slab_decompose(n, parts, nz_d, gk_start);
  if (rc != H3D_OK) return rc;
  const int lo = part == 0 ? 0 : gk_start[part];
  const int hi = part == parts - 1 ? n - 1 : gk_start[part + 1] - 1;
  const size_t plane = (size_t)n * (size_t)n;
  for (int k = 0; k < nz_d[part]; k++) {
    const int gk = k - 1 + gk_start[part];
    if (gk >= lo && gk <= hi)
      memcpy(global + (size_t)gk * plane, slab + (size_t)k * plane, plane * sizeof(double));
  }
  return H3D_OK;
}

/* ---- seeded inputs ------------------------------------------------------ */

static inline uint64_t mix64(uint64_t z)
{
  /* SplitMix64 finaliser (Steele, Lea, Flood 2014) */
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

double h3d_rng_uniform(uint64_t seed, int member, int stream, uint64_t idx)
{
  uint64_t key =
      mix64(seed ^ mix64(((uint64_t)(uint32_t)member << 8) | (uint64_t)(stream & 0xff)));
  uint64_t z = mix64(key + 0x9E3779B97F4A7C15ULL * (idx + 1));
  return (double)(z >> 11) * 0x1.0p-53;
}

enum { RNG_KAPPA = 1, RNG_U0 = 2, RNG_BM1 = 3, RNG_BM2 = 4 };

int h3d_gen_seeded(double *u0, double *kappa, const h3d_grid *g, int gk0,
                   int gnx, int gny, int gnz, int kappa_kind, uint64_t seed,
                   int member)
{
  if (!g || g->nx != gnx || g->ny != gny || gnz < 3 || gnx < 3 || gny < 3)
    return H3D_EARG;
  if (g->sy < g->nx || g->sz < (long long)g->sy * g->ny) return H3D_EARG;
  if (gk0 < 0 || gk0 + g->nz > gnz) return H3D_EARG;
  if (kappa_kind != 0 && kappa_kind != 1) return H3D_EARG;
  const double h = 1.0 / (gnx - 1);
  const double cx = 0.5 * (gnx - 1) * h, cy = 0.5 * (gny - 1) * h,
               cz = 0.5 * (gnz - 1) * h;
  const double sig = 0.1, two_pi = 2.0 * M_PI;
#pragma omp parallel for schedule(static)
  for (int kl = 0; kl < g->nz; kl++) {
    const int k = gk0 + kl;
    const double z = k * h;
    for (int j = 0; j < gny; j++) {
      const double y = j * h;
      for (int i = 0; i < gnx; i++) {
        const double x = i * h;
        const size_t idx = (size_t)kl * (size_t)g->sz + (size_t)j * (size_t)g->sy + i;
        const uint64_t gid = ((uint64_t)k * (uint64_t)gny + (uint64_t)j) * (uint64_t)gnx + (uint64_t)i;
        if (kappa) {
          if (kappa_kind == 0) {
            double base = 1.0 + 0.5 * sin(two_pi * x) * sin(two_pi * y) * sin(two_pi * z);
            kappa[idx] = base * (0.9 + 0.2 * h3d_rng_uniform(seed, member, RNG_KAPPA, gid));
          } else {
            double a = 1.0 - h3d_rng_uniform(seed, member, RNG_BM1, gid); /* (0,1] */
            double b = h3d_rng_uniform(seed, member, RNG_BM2, gid);
            double zn = sqrt(-2.0 * log(a)) * cos(two_pi * b);
            kappa[idx] = exp(0.5 * zn);
          }
        }
        if (u0) {
          const int ring = (i == 0 || i == gnx - 1 || j == 0 || j == gny - 1 ||
                            k == 0 || k == gnz - 1);
          if (ring) {
            u0[idx] = 0.0;
          } else {
            double r2 = (x - cx) * (x - cx) + (y - cy) * (y - cy) + (z - cz) * (z - cz);
            u0[idx] = exp(-r2 / (2.0 * sig * sig)) +
                      1e-3 * h3d_rng_uniform(seed, member, RNG_U0, gid);
          }
        }
      }
    }
  }
  return H3D_OK;
}

double h3d_face_sum_max_host(const double *kappa, const h3d_grid *g)
{
  if (!kappa || !g) return -1.0;
  const size_t sy = (size_t)g->sy, sz = (size_t)g->sz;
  double best = 0.0;
#pragma omp parallel for schedule(static) reduction(max : best)
  for (int k = 1; k < g->nz - 1; k++)
    for (int j = 1; j < g->ny - 1; j++)
      for (int i = 1; i < g->nx - 1; i++) {
        size_t c = (size_t)k * sz + (size_t)j * sy + i;
        double kc = kappa[c];
        double s = (kappa[c + 1] + kc) + (kc + kappa[c - 1]) +
                   (kappa[c + sy] + kc) + (kc + kappa[c - sy]) +
                   (kappa[c + sz] + kc) + (kc + kappa[c - sz]);
        if (s > best) best = s;
      }
  return best;
}

static uint16_t f32_to_bf16_rne(float x)
{
  uint32_t w;
  memcpy(&w, &x, 4);
  w += 0x7fffu + ((w >> 16) & 1u);
  return (uint16_t)(w >> 16);
}

int h3d_bf16_scale(const double *kappa, size_t n, double *scale)
{
  if (!kappa || !scale || n == 0) return H3D_EARG;
  double mx = 0.0;
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(max : mx) reduction(| : bad)
  for (size_t c = 0; c < n; c++) {
    if (!isfinite(kappa[c])) bad = 1;
    else if (fabs(kappa[c]) > mx) mx = fabs(kappa[c]);
  }
  if (bad) return H3D_EARG;
  int e = 0;
  if (mx > 0.0) frexp(mx, &e); /* mx = m * 2^e, m in [0.5, 1) */
  *scale = ldexp(1.0, e);
  return H3D_OK;
}

int h3d_quantize_bf16_scaled(const double *kappa, uint16_t *bits, size_t n, double scale)
{
  if (!kappa || !bits || n == 0 || !(scale > 0.0) || !isfinite(scale)) return H3D_EARG;
  int e = 0;
  if (frexp(scale, &e) != 0.5) return H3D_EARG; /* a power of two: division is exact */
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (size_t c = 0; c < n; c++) {
    const double q = kappa[c] / scale;
    if (!isfinite(q) || fabs(q) >= 0x1.fep127) bad = 1; /* bf16 holds it finitely */
    else bits[c] = f32_to_bf16_rne((float)q);
  }
  return bad ? H3D_EARG : H3D_OK;
}

int h3d_quantize_bf16(const double *kappa, uint16_t *bits, size_t n,
                      double *scale)
{
  if (!kappa || !bits || !scale || n == 0) return H3D_EARG;
  double s = 0.0;
  const int rc = h3d_bf16_scale(kappa, n, &s);
  if (rc != H3D_OK) return rc;
  const int rq = h3d_quantize_bf16_scaled(kappa, bits, n, s);
  if (rq == H3D_OK) *scale = s;
  return rq;
}

int h3d_write_dump(const char *fname, const double *u, const h3d_grid *g)
{
  if (!fname || !u || !g) return H3D_EARG;
  FILE *f = fopen(fname, "w");
  if (!f) return H3D_EHOST;
  /* 1-openmp.c:41-51: one value per line, k, j, i order, %.17g */
  for (int k = 0; k < g->nz; k++)
    for (int j = 0; j < g->ny; j++)
      for (int i = 0; i < g->nx; i++)
        fprintf(f, "%.17g\n", u[(size_t)k * (size_t)g->sz + (size_t)j * (size_t)g->sy + i]);
  if (fclose(f) != 0) return H3D_EHOST;
  printf("Solution written to %s\n", fname);
  return H3D_OK;
}
