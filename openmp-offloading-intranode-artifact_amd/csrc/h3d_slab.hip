// One GPU's z-slab of a multi-GPU split, at its real shape, with the face
// exchange done in HBM.
//
// The reference's middle devices (4-openmp-stream-omp-p2p.c:655-727, HIP
// twin 4-hip-stream-omp-p2p.cpp:475-500) compute each step as: the two
// boundary planes k = 1 and k = nz-2 (the planes sent to the neighbours),
// the interior 2..nz-3 overlapped with them, then push the boundary planes
// into the neighbours' ghost planes of u_new (k = nz-1 of the lower
// neighbour, k = 0 of the upper one).  On one GPU the neighbours are this
// slab itself: the step packs plane 1 and plane nz-2 of u_new into face
// buffers and unpacks them into ghost planes nz-1 and 0 -- a z-periodic
// stand-in with the same kernels, bytes and ordering as the real exchange
// (pack order: the whole XY plane, ring included, [j][i]).
//
// Overlap: the interior kernel runs on a second stream while the caller's
// stream runs the boundary planes and the face copies, which touch only
// planes 0, 1, nz-2 and nz-1 of u_new; the two streams join at the end of
// the step.  The whole sequence can be captured into a hipGraph.
#include "h3d_internal.h"

namespace {

int slab_step(double *un, const double *uo, const double *rhs,
              const double *kap, const h3d_grid *g, double factor,
              double *face_lo, double *face_hi, const h3d_tune *tune,
              hipStream_t s, hipStream_t s2, hipEvent_t fork, hipEvent_t join) {
  const int nz = g->nz;
  const bool interior = nz >= 5;
  if (interior && s2) {
    H3D_TRY(hipEventRecord(fork, s));
    H3D_TRY(hipStreamWaitEvent(s2, fork, 0));
    int rc = h3d_step_f64(un, uo, rhs, kap, g, factor, 2, nz - 3, tune, s2);
    if (rc != H3D_OK) return rc;
  }
  int rc = h3d_step_f64(un, uo, rhs, kap, g, factor, 1, 1, nullptr, s);
  if (rc == H3D_OK && nz - 2 != 1)
    rc = h3d_step_f64(un, uo, rhs, kap, g, factor, nz - 2, nz - 2, nullptr, s);
  if (rc == H3D_OK) rc = h3d_pack_face_f64(face_lo, un, g, 1, s);
  if (rc == H3D_OK) rc = h3d_pack_face_f64(face_hi, un, g, nz - 2, s);
  if (rc == H3D_OK) rc = h3d_unpack_face_f64(un, face_hi, g, 0, s);
  if (rc == H3D_OK) rc = h3d_unpack_face_f64(un, face_lo, g, nz - 1, s);
  if (rc != H3D_OK) return rc;
  if (interior && !s2) {
    rc = h3d_step_f64(un, uo, rhs, kap, g, factor, 2, nz - 3, tune, s);
    if (rc != H3D_OK) return rc;
  }
  if (interior && s2) {
    H3D_TRY(hipEventRecord(join, s2));
    H3D_TRY(hipStreamWaitEvent(s, join, 0));
  }
  return H3D_OK;
}

bool slab_args_ok(const void *un, const void *uo, const void *kap,
                  const h3d_grid *g, const void *lo, const void *hi) {
  return un && uo && kap && lo && hi && un != uo && lo != hi && h3d_grid_ok(g) &&
         g->nz >= 4;
}

}  // namespace

extern "C" int h3d_slab_step_f64(double *u_new, const double *u_old,
                                 const double *rhs, const double *kappa,
                                 const h3d_grid *g, double factor,
                                 double *face_lo, double *face_hi,
                                 const h3d_tune *tune, void *stream) {
  if (!slab_args_ok(u_new, u_old, kappa, g, face_lo, face_hi)) return H3D_EARG;
  return slab_step(u_new, u_old, rhs, kappa, g, factor, face_lo, face_hi, tune,
                   h3d_stream(stream), nullptr, nullptr, nullptr);
}

extern "C" int h3d_run_slab_f64(double *u_a, double *u_b, const double *rhs,
                                const double *kappa, const h3d_grid *g,
                                double factor, double *face_lo,
                                double *face_hi, int steps,
                                const h3d_tune *tune, int graph, void *stream,
                                float *ms, int *final_in_b) {
  if (!slab_args_ok(u_a, u_b, kappa, g, face_lo, face_hi) || steps < 0)
    return H3D_EARG;
  hipStream_t s = h3d_stream(stream);
  if (!s) return H3D_EARG;  // the overlap needs a created stream to fork from
  hipStream_t s2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr, e0 = nullptr, e1 = nullptr;
  hipGraph_t gr = nullptr;
  hipGraphExec_t ge = nullptr;
  int rc = H3D_OK;
  hipError_t he = hipStreamCreateWithFlags(&s2, hipStreamNonBlocking);
  if (he == hipSuccess) he = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
  if (he == hipSuccess) he = hipEventCreateWithFlags(&join, hipEventDisableTiming);
  if (he == hipSuccess) he = hipEventCreate(&e0);
  if (he == hipSuccess) he = hipEventCreate(&e1);
  if (he != hipSuccess) rc = (int)he;
  if (rc == H3D_OK && graph && steps >= 2) {
    he = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (he != hipSuccess) rc = (int)he;
    if (rc == H3D_OK) {
      int r1 = slab_step(u_b, u_a, rhs, kappa, g, factor, face_lo, face_hi, tune, s, s2, fork, join);
      int r2 = r1 == H3D_OK ? slab_step(u_a, u_b, rhs, kappa, g, factor, face_lo, face_hi, tune, s, s2, fork, join) : r1;
      he = hipStreamEndCapture(s, &gr);
      if (r2 != H3D_OK) rc = r2;
      else if (he != hipSuccess) rc = (int)he;
    }
    if (rc == H3D_OK) {
      he = hipGraphInstantiate(&ge, gr, nullptr, nullptr, 0);
      if (he != hipSuccess) rc = (int)he;
    }
  }
  if (rc == H3D_OK) {
    he = hipEventRecord(e0, s);
    if (he != hipSuccess) rc = (int)he;
  }
  int done = 0;
  if (rc == H3D_OK && ge) {
    for (; done + 2 <= steps; done += 2) {
      he = hipGraphLaunch(ge, s);
      if (he != hipSuccess) { rc = (int)he; break; }
    }
  }
  for (; rc == H3D_OK && done < steps; ++done) {
    const bool even = done % 2 == 0;
    rc = slab_step(even ? u_b : u_a, even ? u_a : u_b, rhs, kappa, g, factor,
                   face_lo, face_hi, tune, s, s2, fork, join);
  }
  if (rc == H3D_OK) {
    float t = 0.f;
    he = hipEventRecord(e1, s);
    if (he == hipSuccess) he = hipEventSynchronize(e1);
    if (he == hipSuccess) he = hipEventElapsedTime(&t, e0, e1);
    if (he != hipSuccess) rc = (int)he;
    else {
      if (ms) *ms = t;
      if (final_in_b) *final_in_b = steps % 2;
    }
  }
  if (ge) (void)hipGraphExecDestroy(ge);
  if (gr) (void)hipGraphDestroy(gr);
  if (fork) (void)hipEventDestroy(fork);
  if (join) (void)hipEventDestroy(join);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (s2) {
    (void)hipStreamSynchronize(s2);
    (void)hipStreamDestroy(s2);
  }
  return rc;
}
