/*
 * h3d-verify -- compare solution dumps against a reference dump, on the GPU.
 *
 * Drop-in for the reference pipeline's checker (HIP/compare.py:78-151 and
 * the PASS rule it shares with verify.sh:239-286): same command line, same
 * printed lines, same exit codes, so a verify step can call either.
 *
 *   h3d-verify [--ref FILE] [--atol A] [--rtol R] [--device D] [FILES...]
 *
 * Every file is one value per line (%.17g), the flat u[k][j][i] dump the
 * solvers write (1-openmp.c:41-51, h3d's WRITE_TXT dump).  Without FILES
 * the known outputs of the reference's GPU binaries and of h3d found in the
 * current directory are compared.  A candidate PASSes when it has no NaN or
 * Inf and |cand - ref| <= atol + rtol |ref| everywhere (numpy.allclose with
 * ref as the second argument).  The metrics (max |d|, rms, ||d||/||ref||,
 * max relative error over |ref| > 1e-12, first argmax) are computed by
 * h3d_compare_f64 on the device.  Dumps hold n^3 values (n = points per
 * axis), which is what the solvers write.
 *
 * Exit status: 0 if every compared file passes, 1 otherwise, 2 if the
 * reference is missing or no candidate exists (compare.py:120-151).
 */
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "h3d.h"

static const char *const KNOWN[][2] = {
    /* the reference's GPU outputs (compare.py:22-33) */
    {"1-GPU OpenMP", "1_openmp_gpu.txt"},
    {"2-GPU", "u_openmp_2gpu.txt"},
    {"2-GPU stream", "u_openmp_stream_2gpu.txt"},
    {"2-GPU stream-omp", "u_openmp_stream_omp_2gpu.txt"},
    {"2-GPU stream-omp-p2p", "u_openmp_stream_omp_p2p_2gpu.txt"},
    {"4-GPU", "u_openmp_4gpu.txt"},
    {"4-GPU stream", "u_openmp_stream_4gpu.txt"},
    {"4-GPU stream-omp", "u_openmp_stream_omp_4gpu.txt"},
    {"4-GPU stream-omp-p2p", "u_openmp_stream_omp_p2p_4gpu.txt"},
    /* h3d's own dumps */
    {"1-GPU h3d", "u_h3d_1gpu.txt"},
};

/* Reads a dump; returns the value count, or -1 with errno-style message in
 * err.  *out is malloc'ed. */
static long long load_dump(const char *path, double **out, char *err, size_t errlen)
{
  FILE *f = fopen(path, "r");
  if (!f) {
    snprintf(err, errlen, "could not read '%s': %s", path, strerror(errno));
    return -1;
  }
  size_t cap = 1 << 20, n = 0;
  double *v = (double *)malloc(cap * sizeof(double));
  char line[128];
  long long lineno = 0;
  while (v && fgets(line, sizeof line, f)) {
    ++lineno;
    char *p = line;
    while (*p == ' ' || *p == '\t') ++p;
    if (*p == '\n' || *p == '\0' || *p == '#') continue;  /* blank / comment */
    char *end;
    const double x = strtod(p, &end);
    while (*end == ' ' || *end == '\t' || *end == '\r' || *end == '\n') ++end;
    if (end == p || *end != '\0') {
      snprintf(err, errlen, "could not read '%s': bad value on line %lld", path, lineno);
      free(v);
      fclose(f);
      return -1;
    }
    if (n == cap) {
      cap *= 2;
      double *w = (double *)realloc(v, cap * sizeof(double));
      if (!w) { free(v); v = NULL; break; }
      v = w;
    }
    v[n++] = x;
  }
  fclose(f);
  if (!v) {
    snprintf(err, errlen, "could not read '%s': out of memory", path);
    return -1;
  }
  *out = v;
  return (long long)n;
}

static int cube_root(long long n)
{
  long long c = (long long)llround(cbrt((double)n));
  for (long long d = c - 1; d <= c + 1; ++d)
    if (d > 0 && d * d * d == n) return (int)d;
  return -1;
}

/* Python prints every NaN as "nan"; C prints the sign bit too. */
static double py_nan(double x) { return isnan(x) ? fabs(x) : x; }

static const char *base_name(const char *p)
{
  const char *s = strrchr(p, '/');
  return s ? s + 1 : p;
}

/* One candidate: 1 pass, 0 fail, -1 not present. */
static int compare_one(const char *label, const char *path, const double *ref,
                       const double *d_ref, long long nref, int n, double atol,
                       double rtol, int dev)
{
  FILE *probe = fopen(path, "r");
  if (!probe) return -1;
  fclose(probe);
  double *cand = NULL;
  char err[512];
  const long long nc = load_dump(path, &cand, err, sizeof err);
  if (nc < 0) {
    printf("  [ERROR] %-22s %-34s %s\n", label, path, err);
    return 0;
  }
  if (nc != nref) {
    printf("  [FAIL ] %-22s %-34s size mismatch: ref=%lld cand=%lld\n", label, path,
           nref, nc);
    free(cand);
    return 0;
  }
  const size_t bytes = (size_t)nc * sizeof(double);
  double *d_cand = NULL;
  double m[9];
  h3d_grid g = {n, n, n, n, (long long)n * n};
  int rc = h3d_target_alloc((void **)&d_cand, bytes, dev);
  if (rc == H3D_OK) rc = h3d_target_memcpy(d_cand, cand, bytes, dev, -1);
  if (rc == H3D_OK) rc = h3d_compare_f64(d_ref, d_cand, &g, atol, rtol, m, NULL);
  if (d_cand) h3d_target_free(d_cand, dev);
  if (rc != H3D_OK) {
    printf("  [ERROR] %-22s %-34s %s\n", label, path, h3d_error_string(rc));
    free(cand);
    return 0;
  }
  const int ok = m[6] == 0 && m[7] == 0 && m[8] == 0;
  printf("  [%s ] %-22s %-34s maxabs=%.3e  rms=%.3e  relL2=%.3e  maxrel=%.3e\n",
         ok ? "PASS" : "FAIL", label, base_name(path), py_nan(m[0]), py_nan(m[2]),
         py_nan(m[3]), py_nan(m[4]));
  if (!ok) {
    const long long i = (long long)m[5];
    printf("           worst @ idx %lld: ref=%.6e  cand=%.6e", i, py_nan(ref[i]),
           py_nan(cand[i]));
    if (m[6] != 0 || m[7] != 0) printf("  NaN=%.0f Inf=%.0f", m[6], m[7]);
    printf("\n");
  }
  free(cand);
  return ok;
}

static void usage(const char *prog)
{
  fprintf(stderr,
          "usage: %s [--ref FILE] [--atol A] [--rtol R] [--device D] [FILES...]\n",
          prog);
}

int main(int argc, char **argv)
{
  const char *ref_path = "u_serial.txt";
  double atol = 1e-6, rtol = 1e-5;
  int dev = 0;
  const char **files = (const char **)calloc((size_t)argc, sizeof(char *));
  int nfiles = 0;
  for (int a = 1; a < argc; ++a) {
    const char *s = argv[a], *v = a + 1 < argc ? argv[a + 1] : NULL;
    char *end = NULL;
    if (!strcmp(s, "--ref") && v) { ref_path = v; ++a; }
    else if (!strcmp(s, "--atol") && v) { atol = strtod(v, &end); ++a; if (*end) { usage(argv[0]); return 2; } }
    else if (!strcmp(s, "--rtol") && v) { rtol = strtod(v, &end); ++a; if (*end) { usage(argv[0]); return 2; } }
    else if (!strcmp(s, "--device") && v) { dev = atoi(v); ++a; }
    else if (s[0] == '-' && s[1] == '-') { usage(argv[0]); return 2; }
    else files[nfiles++] = s;
  }

  FILE *probe = fopen(ref_path, "r");
  if (!probe) {
    printf("ERROR: reference file '%s' not found.\n", ref_path);
    printf("Run the batch script in 'write' mode first to produce u_serial.txt.\n");
    return 2;
  }
  fclose(probe);
  double *ref = NULL;
  char err[512];
  const long long nref = load_dump(ref_path, &ref, err, sizeof err);
  if (nref < 0) {
    printf("ERROR: %s\n", err);
    return 2;
  }
  const int n = cube_root(nref);
  if (n < 3) {
    printf("ERROR: reference '%s' holds %lld values, not n^3 with n >= 3\n", ref_path, nref);
    return 2;
  }
  printf("Reference: %s  (%lld values)\n", ref_path, nref);
  printf("Tolerance: atol=%g, rtol=%g\n\n", atol, rtol);

  double *d_ref = NULL;
  const size_t bytes = (size_t)nref * sizeof(double);
  int rc = h3d_target_alloc((void **)&d_ref, bytes, dev);
  if (rc == H3D_OK) rc = h3d_target_memcpy(d_ref, ref, bytes, dev, -1);
  if (rc != H3D_OK) {
    fprintf(stderr, "ERROR: device %d: %s\n", dev, h3d_error_string(rc));
    return 2;
  }

  int found = 0, total = 0, passed = 0;
  const int nt = nfiles ? nfiles : (int)(sizeof KNOWN / sizeof KNOWN[0]);
  for (int t = 0; t < nt; ++t) {
    const char *label = nfiles ? base_name(files[t]) : KNOWN[t][0];
    const char *path = nfiles ? files[t] : KNOWN[t][1];
    const int r = compare_one(label, path, ref, d_ref, nref, n, atol, rtol, dev);
    if (r < 0) continue;
    ++found;
    ++total;
    passed += r;
  }
  h3d_target_free(d_ref, dev);
  free(ref);
  free(files);
  if (!found) {
    printf("  (no candidate GPU output files found in this directory)\n");
    return 2;
  }
  printf("\nSummary: %d/%d passed (atol=%g, rtol=%g)\n", passed, total, atol, rtol);
  return passed == total ? 0 : 1;
}
