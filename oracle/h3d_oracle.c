/*
 * h3d_oracle.c -- CPU oracle for the 3-D variable-coefficient heat update.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for tests/, smoke() and bench.py's
 * cpu_baseline leg.  Nothing in the product links it.
 *
 * Built with -ffp-contract=off (oracle/Makefile) so that every expression
 * below rounds exactly as written, like the reference, whose builds all use
 * -ffp-contract=off (3D_Heat_Benchmark/README.md:112-113).
 *
 * Pinning (how each function is checked against something other than
 * itself; see tests/test_oracle.py):
 *   orc_params_reference, orc_count_steps  -- banner values and step counts
 *       recorded in AMD-MI250X/Source-Results/HIP/verify-20370244.out and
 *       the summary.csv steps column;
 *   orc_init_reference, orc_step_f64, orc_run_f64, orc_write_dump
 *       -- byte-compared with u_serial.txt written by the reference's own
 *       heat3D_variable_coeff.c compiled from source into oracle/_ref
 *       (committed as golden fixtures), plus the closed-form eigenmode decay
 *       g^s for constant kappa;
 *   orc_slab_decompose -- integer closed forms of the reference's
 *       decomposition (n=1026 -> 258 planes) and slab-vs-global equality;
 *   orc_metrics -- hand-computed vectors;
 *   orc_stats, orc_face_sum_max -- math.fsum / brute force in Python;
 *   orc_run_f32 -- FP32 paths have no reference counterpart.  All modes
 *       are pinned against a numpy float32 restatement (whose fused
 *       multiply-add is an exact FP64 emulation, itself pinned against
 *       rational arithmetic), and modes 1-2 against the FP64 oracle
 *       (error bounds).
 */
#include "h3d_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_params_reference(int n, double *h, double *dt, double *factor)
{
  /* heat3D_variable_coeff.c:79-84: factor is dt/h/h/2.0, evaluated left
   * to right -- not dt/(2h^2), which rounds differently. */
  double hh = 1.0 / (n - 1);
  double kappa_max = 0.95;
  double d = hh * hh / (12.0 * kappa_max);
  *h = hh;
  *dt = d;
  *factor = d / hh / hh / 2.0;
}

int orc_count_steps(double T, double dt)
{
  double t = 0.0;
  int steps = 0;
  while (t < T) {
    steps++;
    t += dt;
  }
  return steps;
}

void orc_init_reference(double *u_old, double *u_new, double *rhs,
                        double *kappa, int n)
{
  double h, dt, factor;
  orc_params_reference(n, &h, &dt, &factor);
  for (int k = 0; k < n; k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        size_t idx = ((size_t)k * n + j) * n + i;
        rhs[idx] = dt * (1.0 + cos(M_PI * i * h) * cos(M_PI * j * h) *
                                   cos(M_PI * k * h));
        kappa[idx] = 0.5 + 0.45 * sin(M_PI * i * h) * sin(M_PI * j * h) *
                               sin(M_PI * k * h);
        int ring = (i == 0 || i == n - 1 || j == 0 || j == n - 1 || k == 0 ||
                    k == n - 1);
        u_old[idx] = ring ? 0.0
                          : sin(M_PI * i * h) * sin(M_PI * j * h) *
                                sin(M_PI * k * h);
        if (u_new) u_new[idx] = 0.0;
      }
}

void orc_init_partition(double *u, double *rhs, double *kappa, int nz_local,
                        int gk_start, int ng, int n, int ny, int nx, double h,
                        double dt)
{
  for (int k = 0; k < nz_local; k++) {
    int gk = (k - ng) + gk_start;
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        size_t idx = ((size_t)k * ny + j) * nx + i;
        u[idx] = (gk <= 0 || gk >= n - 1 || j <= 0 || j >= n - 1 || i <= 0 ||
                  i >= n - 1)
                     ? 0.0
                     : sin(M_PI * i * h) * sin(M_PI * j * h) *
                           sin(M_PI * gk * h);
        rhs[idx] = dt * (1.0 + cos(M_PI * i * h) * cos(M_PI * j * h) *
                                   cos(M_PI * gk * h));
        kappa[idx] = 0.5 + 0.45 * sin(M_PI * i * h) * sin(M_PI * j * h) *
                               sin(M_PI * gk * h);
      }
  }
}

static inline double cell_update(const double *uo, const double *kp,
                                 const double *rhs, size_t idx, size_t sy,
                                 size_t sz, double factor)
{
  /* heat3D_variable_coeff.c:174-181, term order and signs as written. */
  double uc = uo[idx], kc = kp[idx];
  double s = (kp[idx + 1] + kc) * (uo[idx + 1] - uc) -
             (kc + kp[idx - 1]) * (uc - uo[idx - 1]) +
             (kp[idx + sy] + kc) * (uo[idx + sy] - uc) -
             (kc + kp[idx - sy]) * (uc - uo[idx - sy]) +
             (kp[idx + sz] + kc) * (uo[idx + sz] - uc) -
             (kc + kp[idx - sz]) * (uc - uo[idx - sz]);
  if (rhs) return uc + rhs[idx] + factor * s;
  return uc + factor * s;
}

void orc_step_f64(double *u_new, const double *u_old, const double *rhs,
                  const double *kappa, int nx, int ny, int nz, int sy,
                  long long sz, double factor, int k_start, int k_end)
{
  (void)nz;
  for (int k = k_start; k <= k_end; k++)
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) {
        size_t idx = (size_t)k * (size_t)sz + (size_t)j * (size_t)sy + i;
        u_new[idx] = cell_update(u_old, kappa, rhs, idx, (size_t)sy,
                                 (size_t)sz, factor);
      }
}

int orc_run_f64(double *ua, double *ub, const double *rhs, const double *kappa,
                int nx, int ny, int nz, int sy, long long sz, double factor,
                int steps, int nthreads)
{
  double *uo = ua, *un = ub;
  if (nthreads <= 1) {
    for (int s = 0; s < steps; s++) {
      orc_step_f64(un, uo, rhs, kappa, nx, ny, nz, sy, sz, factor, 1, nz - 2);
      double *t = uo; uo = un; un = t;
    }
  } else {
#ifdef _OPENMP
    /* heat3D_variable_coeff.c:159-195: one parallel region, `omp for` over
     * k, pointer swap in `omp single`. */
#pragma omp parallel num_threads(nthreads) default(shared)
    {
      for (int s = 0; s < steps; s++) {
#pragma omp for schedule(static)
        for (int k = 1; k < nz - 1; k++)
          for (int j = 1; j < ny - 1; j++)
            for (int i = 1; i < nx - 1; i++) {
              size_t idx = (size_t)k * (size_t)sz + (size_t)j * (size_t)sy + i;
              un[idx] = cell_update(uo, kappa, rhs, idx, (size_t)sy,
                                    (size_t)sz, factor);
            }
#pragma omp single
        {
          double *t = uo; uo = un; un = t;
        }
      }
    }
#else
    return orc_run_f64(ua, ub, rhs, kappa, nx, ny, nz, sy, sz, factor, steps, 1);
#endif
  }
  return uo == ua ? 0 : 1;
}

int orc_slab_decompose(int n, int parts, int *nz_d, int *gk_start)
{
  const int ng = 1;
  int nz_int = n - 2;
  if (parts == 1) {
    nz_d[0] = n;
    gk_start[0] = 1;
    return n < 3 ? -1 : 0;
  } else if (parts == 2) {
    /* 2-openmp.c:294-316: the remainder goes to the higher device. */
    int mid = nz_int / 2;
    nz_d[0] = mid + 2 * ng;
    nz_d[1] = (nz_int - mid) + 2 * ng;
  } else if (parts == 4) {
    /* 4-openmp-stream-omp-p2p.c:391-405: remainder to the lower devices. */
    int q = nz_int / 4, r = nz_int % 4;
    nz_d[0] = q + (r > 0 ? 1 : 0) + 2 * ng;
    nz_d[1] = q + (r > 1 ? 1 : 0) + 2 * ng;
    nz_d[2] = q + (r > 2 ? 1 : 0) + 2 * ng;
    nz_d[3] = q + 2 * ng;
  } else {
    return -1;
  }
  gk_start[0] = 1;
  for (int p = 1; p < parts; p++) gk_start[p] = gk_start[p - 1] + (nz_d[p - 1] - 2 * ng);
  for (int p = 0; p < parts; p++)
    if (nz_d[p] < 4) return -1;
  return 0;
}

void orc_pack_face(double *buf, const double *u, int nx, int ny, int sy,
                   long long sz, int k)
{
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++)
      buf[(size_t)j * nx + i] = u[(size_t)k * sz + (size_t)j * sy + i];
}

void orc_unpack_face(double *u, const double *buf, int nx, int ny, int sy,
                     long long sz, int k)
{
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++)
      u[(size_t)k * sz + (size_t)j * sy + i] = buf[(size_t)j * nx + i];
}

/* Neumaier compensated accumulation. */
static inline void neu_add(double *s, double *c, double x)
{
  double t = *s + x;
  if (fabs(*s) >= fabs(x)) *c += (*s - t) + x;
  else *c += (x - t) + *s;
  *s = t;
}

void orc_metrics(const double *ref, const double *cand, size_t n, double atol,
                 double rtol, double *out)
{
  double max_abs = 0.0, sum_abs = 0.0, sum_abs_c = 0.0, sq = 0.0, sq_c = 0.0;
  double rsq = 0.0, rsq_c = 0.0, max_rel = 0.0;
  size_t argmax = 0, n_nan = 0, n_inf = 0, n_fail = 0;
  int first = 1;
  for (size_t i = 0; i < n; i++) {
    double c = cand[i], r = ref[i];
    if (isnan(c)) n_nan++;
    else if (isinf(c)) n_inf++;
    double d = c - r, ad = fabs(d), ar = fabs(r);
    /* np.argmax returns the first maximum; NaN compares as maximal. */
    if (first || ad > max_abs || (isnan(ad) && !isnan(max_abs))) {
      max_abs = ad;
      argmax = i;
      first = 0;
    }
    neu_add(&sum_abs, &sum_abs_c, ad);
    neu_add(&sq, &sq_c, d * d);
    neu_add(&rsq, &rsq_c, r * r);
    if (ar > 1e-12) {
      double rel = ad / (ar > 1e-30 ? ar : 1e-30);
      if (rel > max_rel || isnan(rel)) max_rel = rel;
    }
    if (!isfinite(c) || !(ad <= atol + rtol * ar)) n_fail++;
  }
  double l2d = sqrt(sq + sq_c), l2r = sqrt(rsq + rsq_c);
  out[0] = max_abs;
  out[1] = n ? (sum_abs + sum_abs_c) / (double)n : 0.0;
  out[2] = n ? sqrt((sq + sq_c) / (double)n) : 0.0;
  out[3] = l2r > 0.0 ? l2d / l2r : l2d;
  out[4] = max_rel;
  out[5] = (double)argmax;
  out[6] = (double)n_nan;
  out[7] = (double)n_inf;
  out[8] = (double)n_fail;
}

void orc_stats(const double *u, int nx, int ny, int nz, int sy, long long sz,
               double *out)
{
  double s = 0.0, sc = 0.0, q = 0.0, qc = 0.0, mx = 0.0, bad = 0.0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < nx; i++) {
        double v = u[(size_t)k * sz + (size_t)j * sy + i];
        if (!isfinite(v)) {
          bad += 1.0;
          continue;
        }
        neu_add(&s, &sc, v);
        neu_add(&q, &qc, v * v);
        if (fabs(v) > mx) mx = fabs(v);
      }
  out[0] = s + sc;
  out[1] = q + qc;
  out[2] = mx;
  out[3] = bad;
}

double orc_face_sum_max(const double *kappa, int nx, int ny, int nz, int sy,
                        long long sz)
{
  double best = 0.0;
  for (int k = 1; k < nz - 1; k++)
    for (int j = 1; j < ny - 1; j++)
      for (int i = 1; i < nx - 1; i++) {
        size_t c = (size_t)k * sz + (size_t)j * sy + i;
        double kc = kappa[c];
        double s = (kappa[c + 1] + kc) + (kc + kappa[c - 1]) +
                   (kappa[c + sy] + kc) + (kc + kappa[c - sy]) +
                   (kappa[c + sz] + kc) + (kc + kappa[c - sz]);
        if (s > best) best = s;
      }
  return best;
}

int orc_write_dump(const char *fname, const double *u, int n)
{
  FILE *f = fopen(fname, "w");
  if (!f) return -1;
  for (size_t idx = 0; idx < (size_t)n * n * n; idx++) fprintf(f, "%.17g\n", u[idx]);
  fclose(f);
  return 0;
}

/* ------------------------------------------------------------------ */
/* FP32 storage paths.                                                  */

float orc_bf16_to_f32(uint16_t b)
{
  uint32_t w = (uint32_t)b << 16;
  float f;
  memcpy(&f, &w, 4);
  return f;
}

uint16_t orc_f32_to_bf16(float x)
{
  /* round to nearest even on the upper 16 bits (inputs are finite) */
  uint32_t w;
  memcpy(&w, &x, 4);
  uint32_t lsb = (w >> 16) & 1u;
  w += 0x7fffu + lsb;
  return (uint16_t)(w >> 16);
}

/* IEEE binary16 <-> binary32, round to nearest even, subnormals kept. */
static uint16_t f32_to_f16_bits(float x)
{
  uint32_t w;
  memcpy(&w, &x, 4);
  uint32_t sign = (w >> 16) & 0x8000u;
  uint32_t ex = (w >> 23) & 0xffu;
  uint32_t man = w & 0x7fffffu;
  if (ex == 0xffu) return (uint16_t)(sign | 0x7c00u | (man ? 0x200u : 0u));
  int e = (int)ex - 127 + 15;
  if (e >= 31) return (uint16_t)(sign | 0x7c00u);
  if (e <= 0) {
    if (e < -10) return (uint16_t)sign;
    man |= 0x800000u;
    int shift = 14 - e;                   /* 14..24 */
    uint32_t half = 1u << (shift - 1);
    uint32_t q = man >> shift, rem = man & ((1u << shift) - 1u);
    if (rem > half || (rem == half && (q & 1u))) q++;
    return (uint16_t)(sign | q);
  }
  uint32_t q = man >> 13, rem = man & 0x1fffu;
  uint32_t hv = ((uint32_t)e << 10) | q;
  if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) hv++;
  return (uint16_t)(sign | hv);
}

static float f16_bits_to_f32(uint16_t h)
{
  uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
  uint32_t ex = (h >> 10) & 0x1fu, man = h & 0x3ffu;
  float f;
  if (ex == 0) {
    f = ldexpf((float)man, -24);
    return sign ? -f : f;
  }
  uint32_t w;
  if (ex == 31) w = sign | 0x7f800000u | (man << 13);
  else w = sign | ((ex - 15 + 127) << 23) | (man << 13);
  memcpy(&f, &w, 4);
  return f;
}

/* ulp exponent of the stored value: the residual of a cell holding u is
 * kept as e * 2^(23 - E(u)), E(u) = max(ilogb(u), -126) (-126 for 0),
 * which is O(1): the Kahan residual of a sum is about half an ulp of it. */
static inline int ulp_exp(float u)
{
  if (u == 0.0f || !isfinite(u)) return -126;
  int e = ilogbf(u);
  return e < -126 ? -126 : e;
}

uint16_t orc_comp_encode(float e, float u)
{
  return f32_to_f16_bits(ldexpf(e, 23 - ulp_exp(u)));
}

float orc_comp_decode(uint16_t h, float u)
{
  return ldexpf(f16_bits_to_f32(h), ulp_exp(u) - 23);
}

void orc_run_f32(float *ua, float *ub, const float *kappa_f32,
                 const uint16_t *kappa_bf16, uint16_t *ca, uint16_t *cb,
                 int nx, int ny, int nz, int sy, long long sz, float factor,
                 int steps, int kb, int mode, int *final_index)
{
  size_t total = (size_t)nz * (size_t)sz;
  float *kap = (float *)malloc(total * sizeof(float));
  float *err = (float *)calloc(total, sizeof(float));
  for (size_t c = 0; c < total; c++)
    kap[c] = kappa_f32 ? kappa_f32[c] : orc_bf16_to_f32(kappa_bf16[c]);
  float *uo = ua, *un = ub;
  uint16_t *co = ca, *cn = cb;
  if (kb < 1) kb = 1;
  int done = 0;
  while (done < steps) {
    int nsteps = steps - done < kb ? steps - done : kb;
    /* pass start: residual from nothing (modes 0,1) or from the field */
    for (int k = 1; k < nz - 1; k++)
      for (int j = 1; j < ny - 1; j++)
        for (int i = 1; i < nx - 1; i++) {
          size_t c = (size_t)k * sz + (size_t)j * sy + i;
          err[c] = (mode == 2) ? orc_comp_decode(co[c], uo[c]) : 0.0f;
        }
    for (int s = 0; s < nsteps; s++) {
      for (int k = 1; k < nz - 1; k++)
        for (int j = 1; j < ny - 1; j++)
          for (int i = 1; i < nx - 1; i++) {
            size_t c = (size_t)k * sz + (size_t)j * sy + i;
            float uc = uo[c], kc = kap[c];
            float S = (kap[c + 1] + kc) * (uo[c + 1] - uc) -
                      (kc + kap[c - 1]) * (uc - uo[c - 1]) +
                      (kap[c + sy] + kc) * (uo[c + sy] - uc) -
                      (kc + kap[c - sy]) * (uc - uo[c - sy]) +
                      (kap[c + sz] + kc) * (uo[c + sz] - uc) -
                      (kc + kap[c - sz]) * (uc - uo[c - sz]);
            if (mode == 0) {
              un[c] = fmaf(factor, S, uc);
            } else {
              /* Kahan: keep the part of y the FP32 sum dropped; the
               * increment factor*S joins the residual in one fused
               * multiply-add (a single rounding) */
              float y = fmaf(factor, S, err[c]);
              float t = uc + y;
              err[c] = y - (t - uc);
              un[c] = t;
            }
          }
      /* ring of the destination keeps the Dirichlet values */
      float *t = uo; uo = un; un = t;
    }
    if (mode == 2) {
      for (int k = 1; k < nz - 1; k++)
        for (int j = 1; j < ny - 1; j++)
          for (int i = 1; i < nx - 1; i++) {
            size_t c = (size_t)k * sz + (size_t)j * sy + i;
            cn[c] = orc_comp_encode(err[c], uo[c]);
          }
      uint16_t *t = co; co = cn; cn = t;
    }
    done += nsteps;
  }
  if (final_index) *final_index = (uo == ua ? 0 : 1) | ((co == ca ? 0 : 1) << 1);
  free(kap);
  free(err);
}
