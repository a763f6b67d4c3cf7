// Ensembles: many independent heat problems of uneven sizes, advanced one
// time step per launch by a persistent kernel over a device-side work list.
//
// Each member is cut into tiles (x segments of 64 lanes, y strips, z
// chunks of `zchunk` planes: the z-march depth, chosen at creation or by
// h3d_ensemble_autotune).  The work list holds one int4 {member, tx, ty,
// tz} per tile, largest members first, and a fixed grid of 256-thread
// workgroups -- as many as are resident on the device at once -- walks it
// with a grid stride.  FP64 members run the FP64 tile (2 rows per lane, 4
// waves: 64 x 8 points per z step), bit-identical to h3d_step_f64.  FP32
// members run the FP32 pass tile with one fused step and the FP16 residual
// field (mode H3D_F32_COMP16), so the compensation carries across steps;
// that tile computes the same bits as h3d_pass_f32 with nsteps = 1 (which
// runs the wave-independent single-step tile: inside this persistent
// kernel the LDS tile measured 8% faster over the 64-member ensemble).
// Both kinds share one launch, so a step of the whole ensemble is one
// kernel, and `steps` launches are captured into one hipGraph.
#include <vector>
#include <algorithm>

#include "h3d_tiles.h"

namespace {

constexpr int ENS_WAVES = 4;    // 256-thread workgroups
constexpr int ENS_F64_ROWS = 2;
constexpr int ENS_F32_RY = 4;
constexpr int ENS_ZCHUNK = 64;  // default z-march depth

struct EMember {
  void *ua, *ub;
  const void *kap;
  unsigned short *ca, *cb;  // FP16 residual fields (FP32 members)
  int nx, ny, nz, sy;
  long long sz;
  double factor;
  float factor32;
  int precision;
};

__global__ void __launch_bounds__(ENS_WAVES * 64)
ensemble_step(const EMember *__restrict__ mem, const int4 *__restrict__ work,
              int nwork, int parity, int zchunk) {
  extern __shared__ float lds[];
  for (int t = blockIdx.x; t < nwork; t += gridDim.x) {
    const int4 d = work[t];
    const EMember m = mem[d.x];
    if (m.precision == 0) {
      const double *in = static_cast<const double *>(parity ? m.ub : m.ua);
      double *out = static_cast<double *>(parity ? m.ua : m.ub);
      march_f64_tile<ENS_F64_ROWS, false>(out, in, nullptr,
                                          static_cast<const double *>(m.kap),
                                          m.nx, m.ny, m.sy, m.sz, m.factor, 1,
                                          m.nz - 2, zchunk, d.y, d.z, d.w);
    } else {
      const float *in = static_cast<const float *>(parity ? m.ub : m.ua);
      float *out = static_cast<float *>(parity ? m.ua : m.ub);
      const unsigned short *cin = parity ? m.cb : m.ca;
      unsigned short *cout = parity ? m.ca : m.cb;
      const int zb = 1 + d.w * zchunk;
      pass_f32_tile<1, ENS_F32_RY, ENS_WAVES, 2, false>(
          lds, out, in, static_cast<const float *>(m.kap), nullptr, cout, cin,
          m.nx, m.ny, m.nz, m.sy, m.sz, m.factor32, zb,
          min(m.nz - 2, zb + zchunk - 1), d.y, d.z);
    }
    // the next tile's FP32 set-up rewrites LDS
    __syncthreads();
  }
}

constexpr size_t ENS_LDS =
    sizeof(float) * (size_t)pass_lds_floats(1, ENS_F32_RY, ENS_WAVES);

}  // namespace

struct h3d_ensemble {
  int device;
  int count;
  int nwork;
  int nblocks;
  int zchunk;
  int work_cap;  // int4 slots allocated at d_work
  struct Tiles { int tiles_x, tiles_y, planes, precision; };
  std::vector<Tiles> tiles;  // per member
  long long done;  // steps taken so far (parity of the next launch)
  EMember *d_mem;
  int4 *d_work;
  std::vector<unsigned short *> comp;  // owned FP16 residual fields
  // cached graph of `graph_steps` launches starting at `graph_parity`
  hipGraphExec_t graph = nullptr;
  hipGraph_t graph_def = nullptr;
  int graph_steps = -1, graph_parity = -1;
};

static void ensemble_free(h3d_ensemble *e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  if (e->graph) (void)hipGraphExecDestroy(e->graph);
  if (e->graph_def) (void)hipGraphDestroy(e->graph_def);
  if (e->d_mem) (void)hipFree(e->d_mem);
  if (e->d_work) (void)hipFree(e->d_work);
  for (auto *c : e->comp) (void)hipFree(c);
  delete e;
}

// (Re)builds the work list for z chunks of `zchunk` planes: one int4
// {member, tx, ty, tz} per tile, costliest first (stable, so a member's
// tiles stay together), and sizes the persistent grid.  Drops any cached
// graph, whose launches carry the old list length.
static int build_work(h3d_ensemble *e, int zchunk) {
  if (zchunk < 1) return H3D_EARG;
  struct Job { long long cost; int4 d; };
  std::vector<Job> jobs;
  for (int m = 0; m < e->count; ++m) {
    const h3d_ensemble::Tiles &t = e->tiles[m];
    const int chunks = (t.planes + zchunk - 1) / zchunk;
    for (int tz = 0; tz < chunks; ++tz) {
      const int planes = std::min(zchunk, t.planes - tz * zchunk);
      for (int ty = 0; ty < t.tiles_y; ++ty)
        for (int tx = 0; tx < t.tiles_x; ++tx)
          jobs.push_back({(long long)planes * (t.precision ? 2 : 3), {m, tx, ty, tz}});
    }
  }
  if (jobs.size() > 0x7fffffffu) return H3D_ELIMIT;
  std::stable_sort(jobs.begin(), jobs.end(),
                   [](const Job &a, const Job &b) { return a.cost > b.cost; });
  std::vector<int4> work(jobs.size());
  for (size_t j = 0; j < jobs.size(); ++j) work[j] = jobs[j].d;
  hipError_t he = hipSuccess;
  if ((int)work.size() > e->work_cap) {
    if (e->d_work) (void)hipFree(e->d_work);
    e->d_work = nullptr;
    e->work_cap = 0;
    he = hipMalloc(&e->d_work, sizeof(int4) * work.size());
    if (he == hipSuccess) e->work_cap = (int)work.size();
  }
  if (he == hipSuccess)
    he = hipMemcpy(e->d_work, work.data(), sizeof(int4) * work.size(), hipMemcpyHostToDevice);
  int cus = 0, per_cu = 0;
  if (he == hipSuccess)
    he = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device);
  if (he == hipSuccess)
    he = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void *)ensemble_step,
                                                      ENS_WAVES * 64, ENS_LDS);
  if (he != hipSuccess) return (int)he;
  e->nwork = (int)work.size();
  e->nblocks = std::max(1, std::min(e->nwork, cus * std::max(per_cu, 1)));
  e->zchunk = zchunk;
  if (e->graph) { (void)hipGraphExecDestroy(e->graph); e->graph = nullptr; }
  if (e->graph_def) { (void)hipGraphDestroy(e->graph_def); e->graph_def = nullptr; }
  e->graph_steps = e->graph_parity = -1;
  return H3D_OK;
}

extern "C" int h3d_ensemble_create(h3d_ensemble **out,
                                   const h3d_member *members, int count,
                                   const h3d_tune *tune, int device_id) {
  if (!out || !members || count < 1) return H3D_EARG;
  *out = nullptr;
  int ndev = 0;
  H3D_TRY(hipGetDeviceCount(&ndev));
  if (device_id < 0 || device_id >= ndev) return H3D_ENODEV;
  H3D_TRY(hipSetDevice(device_id));
  for (int m = 0; m < count; ++m) {
    const h3d_member &x = members[m];
    if (!x.u_a || !x.u_b || !x.kappa || x.u_a == x.u_b || !h3d_grid_ok(&x.g))
      return H3D_EARG;
    if (x.precision != 0 && x.precision != 1) return H3D_EARG;
  }
  h3d_ensemble *e = new h3d_ensemble();
  e->device = device_id;
  e->count = count;
  std::vector<EMember> host(count);
  int rc = H3D_OK;
  for (int m = 0; m < count && rc == H3D_OK; ++m) {
    const h3d_member &x = members[m];
    EMember em{};
    em.ua = x.u_a;
    em.ub = x.u_b;
    em.kap = x.kappa;
    em.nx = x.g.nx; em.ny = x.g.ny; em.nz = x.g.nz; em.sy = x.g.sy; em.sz = x.g.sz;
    em.factor = x.factor;
    em.factor32 = (float)x.factor;
    em.precision = x.precision;
    h3d_ensemble::Tiles t{};
    t.planes = x.g.nz - 2;
    t.precision = x.precision;
    if (x.precision == 0) {
      t.tiles_x = (x.g.nx - 2 + 63) / 64;
      t.tiles_y = (x.g.ny - 2 + ENS_F64_ROWS * ENS_WAVES - 1) / (ENS_F64_ROWS * ENS_WAVES);
    } else {
      const int OX = 62, OY = ENS_F32_RY * ENS_WAVES - 2;
      t.tiles_x = (x.g.nx - 2 + OX - 1) / OX;
      t.tiles_y = (x.g.ny - 2 + OY - 1) / OY;
      // residual fields start at zero; they live as long as the ensemble
      const size_t bytes = h3d_grid_elems(&x.g) * sizeof(unsigned short);
      for (int k = 0; k < 2 && rc == H3D_OK; ++k) {
        void *p = nullptr;
        hipError_t he = hipMalloc(&p, bytes);
        if (he == hipSuccess) he = hipMemset(p, 0, bytes);
        if (he != hipSuccess) rc = (int)he;
        else e->comp.push_back(static_cast<unsigned short *>(p));
        if (he == hipSuccess) (k == 0 ? em.ca : em.cb) = static_cast<unsigned short *>(p);
      }
    }
    e->tiles.push_back(t);
    host[m] = em;
  }
  if (rc != H3D_OK) { ensemble_free(e); return rc; }
  hipError_t he = hipMalloc(&e->d_mem, sizeof(EMember) * count);
  if (he == hipSuccess)
    he = hipMemcpy(e->d_mem, host.data(), sizeof(EMember) * count, hipMemcpyHostToDevice);
  if (he != hipSuccess) { ensemble_free(e); return (int)he; }
  const int zc = (tune && tune->zchunk > 0) ? tune->zchunk : ENS_ZCHUNK;
  rc = build_work(e, zc);
  if (rc != H3D_OK) { ensemble_free(e); return rc; }
  *out = e;
  return H3D_OK;
}

static int ensemble_launch(h3d_ensemble *e, int parity, hipStream_t s) {
  hipLaunchKernelGGL(ensemble_step, dim3(e->nblocks), dim3(ENS_WAVES * 64),
                     ENS_LDS, s, e->d_mem, e->d_work, e->nwork, parity, e->zchunk);
  return h3d_launch_status();
}

extern "C" int h3d_ensemble_run(h3d_ensemble *e, int steps, int graph,
                                void *stream, float *ms) {
  if (!e || steps < 0) return H3D_EARG;
  hipStream_t s = h3d_stream(stream);
  if (graph && !s) return H3D_EARG;  // capture needs a created stream
  H3D_TRY(hipSetDevice(e->device));
  const int parity0 = (int)(e->done % 2);
  if (graph && steps > 0 &&
      (e->graph_steps != steps || e->graph_parity != parity0 || !e->graph)) {
    if (e->graph) { (void)hipGraphExecDestroy(e->graph); e->graph = nullptr; }
    if (e->graph_def) { (void)hipGraphDestroy(e->graph_def); e->graph_def = nullptr; }
    H3D_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = H3D_OK;
    for (int k = 0; k < steps && rc == H3D_OK; ++k)
      rc = ensemble_launch(e, (parity0 + k) % 2, s);
    hipGraph_t gdef = nullptr;
    hipError_t ec = hipStreamEndCapture(s, &gdef);
    if (rc != H3D_OK) { if (gdef) (void)hipGraphDestroy(gdef); return rc; }
    H3D_TRY(ec);
    e->graph_def = gdef;
    H3D_TRY(hipGraphInstantiate(&e->graph, gdef, nullptr, nullptr, 0));
    e->graph_steps = steps;
    e->graph_parity = parity0;
  }
  hipEvent_t e0, e1;
  H3D_TRY(hipEventCreate(&e0));
  H3D_TRY(hipEventCreate(&e1));
  int rc = H3D_OK;
  hipError_t he = hipEventRecord(e0, s);
  if (he != hipSuccess) rc = (int)he;
  if (rc == H3D_OK && steps > 0) {
    if (graph) {
      he = hipGraphLaunch(e->graph, s);
      if (he != hipSuccess) rc = (int)he;
    } else {
      for (int k = 0; k < steps && rc == H3D_OK; ++k)
        rc = ensemble_launch(e, (parity0 + k) % 2, s);
    }
  }
  if (rc == H3D_OK) {
    he = hipEventRecord(e1, s);
    if (he == hipSuccess) he = hipEventSynchronize(e1);
    float t = 0.f;
    if (he == hipSuccess) he = hipEventElapsedTime(&t, e0, e1);
    if (he != hipSuccess) rc = (int)he;
    else if (ms) *ms = t;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc == H3D_OK) e->done += steps;
  return rc;
}

// Times each candidate z-march depth over a few launches and keeps the
// fastest.  Every timed launch has the parity of the ensemble's next step,
// so it reads the current fields and rewrites the other buffer with the
// step the next run computes again: member state is unchanged.
extern "C" int h3d_ensemble_autotune(h3d_ensemble *e, void *stream, h3d_tune *out) {
  if (!e) return H3D_EARG;
  hipStream_t s = h3d_stream(stream);
  H3D_TRY(hipSetDevice(e->device));
  const int parity = (int)(e->done % 2);
  hipEvent_t e0, e1;
  H3D_TRY(hipEventCreate(&e0));
  H3D_TRY(hipEventCreate(&e1));
  int best_zc = e->zchunk, rc = H3D_OK;
  float best = 1e30f;
  for (int zc : {32, 64, 128, 256}) {
    rc = build_work(e, zc);
    if (rc != H3D_OK) break;
    rc = ensemble_launch(e, parity, s);  // warm-up
    float t = 1e30f;
    for (int r = 0; r < 3 && rc == H3D_OK; ++r) {
      float ms = 0.f;
      hipError_t he = hipEventRecord(e0, s);
      if (he == hipSuccess) rc = ensemble_launch(e, parity, s);
      if (rc != H3D_OK) break;
      if (he == hipSuccess) he = hipEventRecord(e1, s);
      if (he == hipSuccess) he = hipEventSynchronize(e1);
      if (he == hipSuccess) he = hipEventElapsedTime(&ms, e0, e1);
      if (he != hipSuccess) { rc = (int)he; break; }
      t = std::min(t, ms);
    }
    if (rc != H3D_OK) break;
    if (t < best) { best = t; best_zc = zc; }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != H3D_OK) return rc;
  rc = build_work(e, best_zc);
  if (rc != H3D_OK) return rc;
  if (out) {
    out->rows = 0;
    out->waves = ENS_WAVES;
    out->zchunk = best_zc;
    out->kblock = 1;
  }
  return H3D_OK;
}

extern "C" int h3d_ensemble_info(const h3d_ensemble *e, int *nwork,
                                 int *nblocks, long long *steps_done) {
  if (!e) return H3D_EARG;
  if (nwork) *nwork = e->nwork;
  if (nblocks) *nblocks = e->nblocks;
  if (steps_done) *steps_done = e->done;
  return H3D_OK;
}

extern "C" int h3d_ensemble_destroy(h3d_ensemble *e) {
  if (!e) return H3D_EARG;
  ensemble_free(e);
  return H3D_OK;
}
