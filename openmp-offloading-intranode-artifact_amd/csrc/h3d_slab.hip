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
//
// Deep ghosts (no reference counterpart): a slab with gh ghost planes per
// side runs fused passes of up to gh steps (h3d_pass_f64) and exchanges
// bands of gh planes once per pass -- the same bytes per step as the
// one-plane exchange, in 1/gh of the messages.  With gh = 1 this is the
// reference's step above.
#include "h3d_internal.h"

namespace {

// One pass of `ns` steps (1 <= ns <= gh) of a slab with gh ghost planes per
// side: owned planes gh .. nz-1-gh.  The bands sent to the neighbours, owned
// planes gh .. 2gh-1 and nz-2gh .. nz-1-gh, are computed first on the
// caller's stream and pushed into the ghosts of u_new: the upper band into
// ghosts 0 .. gh-1, the lower band into ghosts nz-gh .. nz-1.  The owned
// planes in between run on s2 meanwhile.  A fused pass of ns steps reads
// the planes within ns of its output planes, all of which the slab holds,
// so its owned planes are exactly ns steps of the update; ghost values the
// pass computes on the way are not kept (the exchange overwrites them).
int slab_pass(double *un, const double *uo, const double *rhs,
              const double *kap, const h3d_grid *g, double factor, int gh,
              int ns, double *face_lo, double *face_hi, const h3d_tune *tune,
              hipStream_t s, hipStream_t s2, hipEvent_t fork, hipEvent_t join) {
  const int nz = g->nz;
  const int lo_end = 2 * gh - 1;            // lower band gh .. 2gh-1
  const int hi_beg = max(nz - 2 * gh, 2 * gh);  // upper band .. nz-1-gh
  const int in_beg = 2 * gh, in_end = nz - 2 * gh - 1;
  const bool interior = in_end >= in_beg;
  auto pass = [&](int ks, int ke, const h3d_tune *t, hipStream_t st) {
    return h3d_pass_f64(un, uo, rhs, kap, g, factor, ns, ks, ke, t, st);
  };
  if (interior && s2) {
    H3D_TRY(hipEventRecord(fork, s));
    H3D_TRY(hipStreamWaitEvent(s2, fork, 0));
    int rc = pass(in_beg, in_end, tune, s2);
    if (rc != H3D_OK) return rc;
  }
  int rc = pass(gh, lo_end, nullptr, s);
  if (rc == H3D_OK && hi_beg <= nz - 1 - gh) rc = pass(hi_beg, nz - 1 - gh, nullptr, s);
  if (rc == H3D_OK) rc = h3d_pack_band_f64(face_lo, un, g, gh, gh, s);
  if (rc == H3D_OK) rc = h3d_pack_band_f64(face_hi, un, g, nz - 2 * gh, gh, s);
  if (rc == H3D_OK) rc = h3d_unpack_band_f64(un, face_hi, g, 0, gh, s);
  if (rc == H3D_OK) rc = h3d_unpack_band_f64(un, face_lo, g, nz - gh, gh, s);
  if (rc != H3D_OK) return rc;
  if (interior && !s2) {
    rc = pass(in_beg, in_end, tune, s);
    if (rc != H3D_OK) return rc;
  }
  if (interior && s2) {
    H3D_TRY(hipEventRecord(join, s2));
    H3D_TRY(hipStreamWaitEvent(s, join, 0));
  }
  return H3D_OK;
}

// Ghost depth 1..4 (the fused pass's depths), at least gh owned planes
// (the bands must be owned), and one ghost slab at least 4 planes deep.
bool slab_args_ok(const void *un, const void *uo, const void *kap,
                  const h3d_grid *g, int gh, const void *lo, const void *hi) {
  return un && uo && kap && lo && hi && un != uo && lo != hi && h3d_grid_ok(g) &&
         gh >= 1 && gh <= 4 && g->nz >= 3 * gh && g->nz >= 4;
}

int run_slab(double *u_a, double *u_b, const double *rhs, const double *kappa,
             const h3d_grid *g, double factor, int gh, double *face_lo,
             double *face_hi, int steps, const h3d_tune *tune, int graph,
             void *stream, float *ms, int *final_in_b) {
  if (!slab_args_ok(u_a, u_b, kappa, g, gh, face_lo, face_hi) || steps < 0)
    return H3D_EARG;
  hipStream_t s = h3d_stream(stream);
  if (!s) return H3D_EARG;  // the overlap needs a created stream to fork from
  hipStream_t s2 = nullptr;
  hipEvent_t fork = nullptr, join = nullptr, e0 = nullptr, e1 = nullptr;
  hipGraph_t gr = nullptr;
  hipGraphExec_t ge = nullptr;
  int rc = H3D_OK;
  // passes of gh steps, the remainder as one shorter pass; the buffers
  // swap once per pass
  const int full = steps / gh, rem = steps % gh;
  const int npass = full + (rem ? 1 : 0);
  hipError_t he = hipStreamCreateWithFlags(&s2, hipStreamNonBlocking);
  if (he == hipSuccess) he = hipEventCreateWithFlags(&fork, hipEventDisableTiming);
  if (he == hipSuccess) he = hipEventCreateWithFlags(&join, hipEventDisableTiming);
  if (he == hipSuccess) he = hipEventCreate(&e0);
  if (he == hipSuccess) he = hipEventCreate(&e1);
  if (he != hipSuccess) rc = (int)he;
  if (rc == H3D_OK && graph && full >= 2) {
    he = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
    if (he != hipSuccess) rc = (int)he;
    if (rc == H3D_OK) {
      int r1 = slab_pass(u_b, u_a, rhs, kappa, g, factor, gh, gh, face_lo, face_hi, tune, s, s2, fork, join);
      int r2 = r1 == H3D_OK ? slab_pass(u_a, u_b, rhs, kappa, g, factor, gh, gh, face_lo, face_hi, tune, s, s2, fork, join) : r1;
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
    for (; done + 2 <= full; done += 2) {
      he = hipGraphLaunch(ge, s);
      if (he != hipSuccess) { rc = (int)he; break; }
    }
  }
  for (; rc == H3D_OK && done < npass; ++done) {
    const bool even = done % 2 == 0;
    rc = slab_pass(even ? u_b : u_a, even ? u_a : u_b, rhs, kappa, g, factor, gh,
                   done < full ? gh : rem, face_lo, face_hi, tune, s, s2, fork, join);
  }
  if (rc == H3D_OK) {
    float t = 0.f;
    he = hipEventRecord(e1, s);
    if (he == hipSuccess) he = hipEventSynchronize(e1);
    if (he == hipSuccess) he = hipEventElapsedTime(&t, e0, e1);
    if (he != hipSuccess) rc = (int)he;
    else {
      if (ms) *ms = t;
      if (final_in_b) *final_in_b = npass % 2;
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

}  // namespace

extern "C" int h3d_slab_step_f64(double *u_new, const double *u_old,
                                 const double *rhs, const double *kappa,
                                 const h3d_grid *g, double factor,
                                 double *face_lo, double *face_hi,
                                 const h3d_tune *tune, void *stream) {
  if (!slab_args_ok(u_new, u_old, kappa, g, 1, face_lo, face_hi)) return H3D_EARG;
  return slab_pass(u_new, u_old, rhs, kappa, g, factor, 1, 1, face_lo, face_hi, tune,
                   h3d_stream(stream), nullptr, nullptr, nullptr);
}

extern "C" int h3d_run_slab_f64(double *u_a, double *u_b, const double *rhs,
                                const double *kappa, const h3d_grid *g,
                                double factor, double *face_lo,
                                double *face_hi, int steps,
                                const h3d_tune *tune, int graph, void *stream,
                                float *ms, int *final_in_b) {
  return run_slab(u_a, u_b, rhs, kappa, g, factor, 1, face_lo, face_hi, steps, tune,
                  graph, stream, ms, final_in_b);
}

extern "C" int h3d_slab_pass_f64(double *u_new, const double *u_old,
                                 const double *rhs, const double *kappa,
                                 const h3d_grid *g, double factor, int ghost,
                                 int nsteps, double *face_lo, double *face_hi,
                                 const h3d_tune *tune, void *stream) {
  if (!slab_args_ok(u_new, u_old, kappa, g, ghost, face_lo, face_hi)) return H3D_EARG;
  if (nsteps < 1 || nsteps > ghost) return H3D_EARG;
  return slab_pass(u_new, u_old, rhs, kappa, g, factor, ghost, nsteps, face_lo, face_hi,
                   tune, h3d_stream(stream), nullptr, nullptr, nullptr);
}

extern "C" int h3d_run_slab_pass_f64(double *u_a, double *u_b, const double *rhs,
                                     const double *kappa, const h3d_grid *g,
                                     double factor, int ghost, double *face_lo,
                                     double *face_hi, int steps,
                                     const h3d_tune *tune, int graph,
                                     void *stream, float *ms, int *final_in_b) {
  return run_slab(u_a, u_b, rhs, kappa, g, factor, ghost, face_lo, face_hi, steps, tune,
                  graph, stream, ms, final_in_b);
}
