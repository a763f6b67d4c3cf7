/*
 * h3d_oracle.h -- CPU oracle for the 3-D variable-coefficient heat update.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the serial
 * reference algorithm, used as the checker by tests/, by smoke() in
 * __graft_entry__.py and by the cpu_baseline leg of bench.py.  The product
 * (the HIP library and the h3d host program) never links or calls it.
 *
 * Layout everywhere: flat [k][j][i], i fastest, element (k,j,i) at
 *   k*sz + j*sy + i     (sy >= nx row stride, sz >= sy*ny plane stride)
 * which for the dense reference layout is IDX(k,j,i,ny,nx) of
 * OpenMP_Single_Source/1-openmp.c:28-29.
 */
#ifndef H3D_ORACLE_H
#define H3D_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* heat3D_variable_coeff.c:79-84 -- h, dt, factor of reference mode. */
void orc_params_reference(int n, double *h, double *dt, double *factor);

/* heat3D_variable_coeff.c:161-193 -- iterations of `while (t<T) t+=dt`. */
int orc_count_steps(double T, double dt);

/* heat3D_variable_coeff.c:110-150 -- dense n^3 reference fields.
 * u_new may be NULL (the serial code zeroes it, the GPU codes copy u0). */
void orc_init_reference(double *u_old, double *u_new, double *rhs,
                        double *kappa, int n);

/* 4-openmp-stream-omp-p2p.c:208-223 (init_partition) -- one z-slab of the
 * reference fields, nz_local planes, first owned plane at global gk_start. */
void orc_init_partition(double *u, double *rhs, double *kappa, int nz_local,
                        int gk_start, int ng, int n, int ny, int nx, double h,
                        double dt);

/* heat3D_variable_coeff.c:163-181, 1-openmp.c:60-83,
 * 4-openmp-stream-omp-p2p.c:258-285 (interior range form).
 * Updates k in [k_start,k_end], j in [1,ny-2], i in [1,nx-2] in the
 * reference's exact association order.  rhs == NULL means rhs == 0 and the
 * addition is skipped (u + 0.0 is exact up to the sign of zero). */
void orc_step_f64(double *u_new, const double *u_old, const double *rhs,
                  const double *kappa, int nx, int ny, int nz, int sy,
                  long long sz, double factor, int k_start, int k_end);

/* The full time loop of heat3D_variable_coeff.c:155-195: `steps` updates
 * with pointer swaps.  nthreads <= 1 is the serial path, otherwise OpenMP
 * over k (as the reference's `omp for`).  Returns 0 if the final field is
 * in ua, 1 if it is in ub. */
int orc_run_f64(double *ua, double *ub, const double *rhs, const double *kappa,
                int nx, int ny, int nz, int sy, long long sz, double factor,
                int steps, int nthreads);

/* 4-openmp-stream-omp-p2p.c:391-415 (parts == 4, remainder to the lower
 * devices) and 2-openmp.c:294-316 (parts == 2, remainder to the higher
 * device).  Writes nz_d[p] (owned planes + 2 ghosts) and gk_start[p].
 * Returns 0, or -1 if a slab would have fewer than 4 planes or parts is
 * not 1, 2 or 4. */
int orc_slab_decompose(int n, int parts, int *nz_d, int *gk_start);

/* Face pack/unpack: the whole XY plane k including the x/y ring, in [j][i]
 * order, nx*ny values (2-openmp.c:321, 4-openmp-stream-omp-p2p.c:609-616). */
void orc_pack_face(double *buf, const double *u, int nx, int ny, int sy,
                   long long sz, int k);
void orc_unpack_face(double *u, const double *buf, int nx, int ny, int sy,
                     long long sz, int k);

/* compare.py:47-76 and verify.sh:239-286.  out[]:
 *  0 max_abs 1 mean_abs 2 rms 3 rel_l2 4 max_rel(|ref|>1e-12) 5 argmax
 *  6 n_nan 7 n_inf 8 n_fail(|d| > atol + rtol*|ref| or non-finite) */
void orc_metrics(const double *ref, const double *cand, size_t n, double atol,
                 double rtol, double *out);

/* FP64 diagnostics over the whole array (ring included), Neumaier-summed:
 * out[0]=sum, out[1]=sum of squares, out[2]=max|u|, out[3]=non-finite count */
void orc_stats(const double *u, int nx, int ny, int nz, int sy, long long sz,
               double *out);

/* Seeded-mode time-step bound: max over interior cells of
 * sum_{6 faces}(kappa_c + kappa_nb).  dt_max = 2 h^2 / result. */
double orc_face_sum_max(const double *kappa, int nx, int ny, int nz, int sy,
                        long long sz);

/* `%.17g` dump, one value per line, k,j,i order (1-openmp.c:41-51). */
int orc_write_dump(const char *fname, const double *u, int n);

/* ------------------------------------------------------------------ */
/* FP32 storage paths (no reference counterpart: the reference is FP64
 * only).  These restate, step by step and in the same operation order,
 * what the FP32 kernels compute, so that a temporally blocked GPU pass of
 * k steps can be checked bit for bit against k sequential CPU steps.
 *
 * One step of cell c with stored value u, carried residual e:
 *   S  = ((((A - B) + C) - D) + E) - F   (A..F as in orc_step_f64, FP32)
 *   y  = fma(f32factor, S, e) ;  t = u + y ;  e' = y - (t - u) ;  u' = t
 *   (Kahan, with the increment joining the residual in one rounding)
 * plain FP32 (mode 0) uses u' = fma(f32factor, S, u) and no e.
 * kappa is given as float (fp32 mode) or as bf16 bits times scale
 * (bf16 mode: kappa = bf16_to_f32(bits) and f32factor already carries
 * the scale).  Residual modes:
 *   mode 0: plain, no residual
 *   mode 1: residual lives only inside a pass of kb steps (reset to 0
 *           at the start of each pass)
 *   mode 2: residual stored between passes in an FP16 field, scaled by
 *           the cell's ulp exponent (see orc_comp_encode). */
void orc_run_f32(float *ua, float *ub, const float *kappa_f32,
                 const uint16_t *kappa_bf16, uint16_t *ca, uint16_t *cb,
                 int nx, int ny, int nz, int sy, long long sz, float factor,
                 int steps, int kb, int mode, int *final_index);

uint16_t orc_comp_encode(float e, float u);
float orc_comp_decode(uint16_t h, float u);
float orc_bf16_to_f32(uint16_t b);
uint16_t orc_f32_to_bf16(float x);

#ifdef __cplusplus
}
#endif
#endif
