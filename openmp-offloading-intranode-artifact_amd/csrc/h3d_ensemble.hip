// Ensembles: many independent heat problems of uneven sizes, advanced by
// persistent kernels over device-side work lists, one group of launches per
// step (or per pass of kblock steps).
//
// Each member is cut into tiles (x segments of 64 lanes, y strips, z
// chunks of `zchunk` planes: the z-march depth).  A work list holds one
// int4 {member, tx, ty, tz} per tile, longest chunks first, split into
// one queue per XCD, and a fixed grid of workgroups -- as many as are
// resident on the device at once -- takes tiles from the queues until all
// are done (ensemble_step).  One persistent launch per pass serves FP64
// and FP32 members alike, each tile with its precision's code and shape:
//  - FP64 members: 1024-thread workgroups of 32 rows (16 waves x 2 rows),
//    the FP64 step tile for single steps (bit-identical to h3d_step_f64)
//    and the FP64 pass tile for passes (bit-identical to h3d_pass_f64);
//  - FP32 members: 1024-thread workgroups of 48 rows (16 waves x 3 rows),
//    the FP32 pass tile with the FP16 residual field (mode
//    H3D_F32_COMP16, precision 1), so the compensation carries across
//    passes, or with the residual carried in registers through each
//    pass's fused steps (mode H3D_F32_CARRY, precision 2); a pass
//    fuses `kblock` steps, bit-identical to h3d_pass_f32 with the same
//    nsteps (any shape: the tile changes the schedule, not the
//    arithmetic).  Against 512-thread workgroups of 32 rows (8 waves x 4
//    rows) the ensemble ran 512.5 vs 491.6 GLUP/s (16 x 2: 505.5; 16 x 4
//    spills 19 VGPRs with the residual field).  Inside a persistent kernel
//    the LDS tile measured 8% faster than the wave-independent single-step
//    tile, so it is used for kblock = 1 too.
// Both tiles are 1024-thread workgroups that fill a CU alone (the FP64
// pass tile's 128 KB of LDS), so one kernel holds either at the occupancy
// it would have alone, and the FP64 and FP32 tiles share the queues: one
// launch, one tail per pass.  As two kernels, one per precision, one after
// the other, the 64-member ensemble ran 509 GLUP/s against 518 (and 500
// against 514 on another box, with the tiles ordered by weighted cost).
//
// With kblock = K > 1 a group of K steps is one launch in which every tile
// takes a K-step pass.  Steps beyond the last full group
// are single-step launches over every tile.  Every member flips its
// ping-pong buffers once per group (h3d_ensemble_where).  The launches of
// a run are captured into one hipGraph on request.
#include <vector>
#include <algorithm>

#include "h3d_tiles.h"

namespace {

// Every tile is a workgroup of ENS_WAVES waves: FP64 members 2 rows per
// lane (32 rows), FP32 members 3 (48 rows).
constexpr int ENS_WAVES = 16;
constexpr int ENS64_ROWS = 2;   // rows per lane of the FP64 step tile
constexpr int ENS64_PRY = 2;    // rows per lane of the FP64 pass tile
constexpr int ENS32_RY = 3;     // rows per lane of the FP32 pass tile
constexpr int ENS_ZCHUNK = 64;  // default z-march depth
constexpr int ENS_MAXKB = 3;    // deepest pass fused inside the ensemble

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

// LDS of the depth-KB kernel in bytes: the larger of the two tiles',
// then one int for the tile index the workgroup has taken from its queue
// (kept at the end, so the tile's doubles stay 8-byte aligned).
template <int KB>
constexpr size_t ens_tile_lds() {
  const size_t f32 = sizeof(float) * (size_t)pass_lds_floats(KB, ENS32_RY, ENS_WAVES);
  const size_t f64 =
      KB > 1 ? sizeof(double) * (size_t)pass64_lds_doubles(KB, ENS64_PRY, ENS_WAVES) : 0;
  return f32 > f64 ? f32 : f64;
}
template <int KB>
constexpr size_t ens_lds() { return ens_tile_lds<KB>() + 16; }

// XCD queues of a work list: workgroup b runs on XCD b mod 8 and takes
// its tiles from that XCD's queue first.
constexpr int ENS_QUEUES = 8;

// Tiles of every member at pass depth KB; parity: which buffer every
// member reads this launch.
//
// The list is split into ENS_QUEUES queues, work[qoff[x] .. qoff[x+1]),
// one per XCD, each costliest tile first.  A workgroup takes the next tile
// of its own XCD's queue (an atomic counter per queue) whenever it has
// finished one, and once that queue is empty takes tiles from the other
// queues, so every workgroup stays busy until the whole list is done
// (list scheduling; a fixed grid-stride assignment left workgroups idle
// behind slower neighbours).  Each queue holds runs of consecutive tiles
// as long as its XCD has workgroups, so neighbouring tiles of a member
// chunk march together under one L2 and share their halo lines there.
// The last workgroup to finish zeroes the counters for the next launch,
// so a launch needs no reset of its own (none inside a captured graph).
template <int KB>
__global__ void __launch_bounds__(ENS_WAVES * 64)
ensemble_step(const EMember *__restrict__ mem, const int4 *__restrict__ work,
              const int *__restrict__ qoff, int *qcnt, int parity, int zchunk) {
  extern __shared__ float lds[];
  int *grabbed = reinterpret_cast<int *>(reinterpret_cast<char *>(lds) + ens_tile_lds<KB>());
  const int home = (int)(blockIdx.x % ENS_QUEUES);
  for (int q = 0; q < ENS_QUEUES; ++q) {
    const int x = (home + q) % ENS_QUEUES;
    const int lo = qoff[x], n = qoff[x + 1] - lo;
    for (;;) {
      if (threadIdx.x == 0) *grabbed = atomicAdd(&qcnt[x], 1);
      __syncthreads();
      // the same value in every lane: read as a scalar, so that the tile
      // descriptor and the member record load into scalar registers
      const int t = __builtin_amdgcn_readfirstlane(*grabbed);
      __syncthreads();  // every thread has read it before it is rewritten
      if (t >= n) break;  // workgroup-uniform
      const int4 d = work[lo + t];
      const EMember m = mem[d.x];
      if (m.precision == 0) {  // workgroup-uniform
        const double *in = static_cast<const double *>(parity ? m.ub : m.ua);
        double *out = static_cast<double *>(parity ? m.ua : m.ub);
        if constexpr (KB > 1) {
          const int zb = 1 + d.w * zchunk;
          pass_f64_tile<KB, ENS64_PRY, ENS_WAVES, false>(
              reinterpret_cast<double *>(lds), out, in, nullptr,
              static_cast<const double *>(m.kap), m.nx, m.ny, m.nz, m.sy, m.sz, m.factor, zb,
              min(m.nz - 2, zb + zchunk - 1), d.y, d.z);
        } else {
          march_f64_tile<ENS64_ROWS, false>(out, in, nullptr,
                                            static_cast<const double *>(m.kap),
                                            m.nx, m.ny, m.sy, m.sz, m.factor, 1,
                                            m.nz - 2, zchunk, 1, d.y, d.z, d.w);
        }
      } else if (m.precision == 1) {
        const float *in = static_cast<const float *>(parity ? m.ub : m.ua);
        float *out = static_cast<float *>(parity ? m.ua : m.ub);
        const unsigned short *cin = parity ? m.cb : m.ca;
        unsigned short *cout = parity ? m.ca : m.cb;
        const int zb = 1 + d.w * zchunk;
        pass_f32_tile<KB, ENS32_RY, ENS_WAVES, 2, false>(
            lds, out, in, static_cast<const float *>(m.kap), nullptr, cout, cin,
            m.nx, m.ny, m.nz, m.sy, m.sz, m.factor32, zb,
            min(m.nz - 2, zb + zchunk - 1), d.y, d.z);
      } else {  // precision 2: the residual carried in registers
        const float *in = static_cast<const float *>(parity ? m.ub : m.ua);
        float *out = static_cast<float *>(parity ? m.ua : m.ub);
        const int zb = 1 + d.w * zchunk;
        pass_f32_tile<KB, ENS32_RY, ENS_WAVES, 1, false>(
            lds, out, in, static_cast<const float *>(m.kap), nullptr, nullptr, nullptr,
            m.nx, m.ny, m.nz, m.sy, m.sz, m.factor32, zb,
            min(m.nz - 2, zb + zchunk - 1), d.y, d.z);
      }
      // the next tile's set-up rewrites LDS
      __syncthreads();
    }
  }
  // Every workgroup counts itself out; the last one (all others have
  // finished taking tiles) zeroes the queue counters and the count.
  if (threadIdx.x == 0) {
    __threadfence();
    if (atomicAdd(&qcnt[ENS_QUEUES], 1) == (int)gridDim.x - 1) {
      for (int x = 0; x <= ENS_QUEUES; ++x) atomicExch(&qcnt[x], 0);
      __threadfence();
    }
  }
}


}  // namespace

struct h3d_ensemble {
  int device;
  int count;
  int zchunk;
  int kb;           // steps per pass
  long long done;   // steps taken so far
  long long passes; // passes taken so far (every member flips once per pass)
  EMember *d_mem;
  struct Tiles {
    // tiles per axis at pass depth d (FP64 members: [1] is the step tile)
    int tiles_x[ENS_MAXKB + 1], tiles_y[ENS_MAXKB + 1], planes, precision;
  };
  std::vector<Tiles> tiles;  // per member
  // Work lists in one device buffer (0: every tile at depth kb, 1: every
  // tile at depth 1, for the steps past the last full group), with the
  // persistent grid of each list's kernel.
  int4 *d_work;
  int work_cap;
  int n[2], blocks[2];
  // per list: the ENS_QUEUES + 1 queue offsets into d_work (d_qoff +
  // list * 16) and the queue counters plus the finished-workgroup count
  // (d_qcnt, same layout; zero between launches)
  int *d_qoff;
  int *d_qcnt;
  std::vector<unsigned short *> comp;  // owned FP16 residual fields
  // cached graph of `graph_steps` steps starting at buffer parity graph_p
  hipGraphExec_t graph = nullptr;
  hipGraph_t graph_def = nullptr;
  int graph_steps = -1, graph_p = -1;
};

static void drop_graph(h3d_ensemble *e) {
  if (e->graph) (void)hipGraphExecDestroy(e->graph);
  if (e->graph_def) (void)hipGraphDestroy(e->graph_def);
  e->graph = nullptr;
  e->graph_def = nullptr;
  e->graph_steps = e->graph_p = -1;
}

static void ensemble_free(h3d_ensemble *e) {
  if (!e) return;
  (void)hipSetDevice(e->device);
  drop_graph(e);
  if (e->d_mem) (void)hipFree(e->d_mem);
  if (e->d_work) (void)hipFree(e->d_work);
  if (e->d_qoff) (void)hipFree(e->d_qoff);
  if (e->d_qcnt) (void)hipFree(e->d_qcnt);
  for (auto *c : e->comp) (void)hipFree(c);
  delete e;
}

// Kernels whose LDS exceeds the 64 KB default need the attribute raised on
// each device before they are sized or launched.
template <int KB>
static int allow_lds() {
  if (ens_lds<KB>() <= 64 * 1024) return H3D_OK;
  static std::atomic<unsigned long long> done{0};
  return h3d_allow_big_lds((const void *)ensemble_step<KB>, &done);
}

// Persistent grid of the depth-kb kernel for nwork tiles: as many
// workgroups as are resident on the device at once.
template <int KB>
static int occupancy_one(int cus, int nwork, int *blocks) {
  const int lrc = allow_lds<KB>();
  if (lrc != H3D_OK) return lrc;
  int per_cu = 0;
  H3D_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, (const void *)ensemble_step<KB>, ENS_WAVES * 64, ens_lds<KB>()));
  *blocks = std::max(1, std::min(std::max(nwork, 1), cus * std::max(per_cu, 1)));
  return H3D_OK;
}

static int occupancy_blocks(int cus, int kb, int nwork, int *blocks) {
  return kb == 1   ? occupancy_one<1>(cus, nwork, blocks)
         : kb == 2 ? occupancy_one<2>(cus, nwork, blocks)
                   : occupancy_one<3>(cus, nwork, blocks);
}

// (Re)builds the work lists for z chunks of `zchunk` planes and passes of
// `kb` steps, costliest tiles first (stable, so a member's tiles stay
// together), and sizes the persistent grids.  Drops any cached graph.
static int build_work(h3d_ensemble *e, int zchunk, int kb) {
  if (zchunk < 1 || kb < 1 || kb > ENS_MAXKB) return H3D_EARG;
  struct Job { long long cost; int4 d; };
  auto list = [&](int depth) {
    std::vector<Job> jobs;
    for (int m = 0; m < e->count; ++m) {
      const h3d_ensemble::Tiles &t = e->tiles[m];
      const int tx_n = t.tiles_x[depth], ty_n = t.tiles_y[depth];
      const int chunks = (t.planes + zchunk - 1) / zchunk;
      for (int tz = 0; tz < chunks; ++tz) {
        const int planes = std::min(zchunk, t.planes - tz * zchunk);
        // a tile's planes, weighted by the measured time of one plane
        // iteration: 2.4 us for the 48-row FP32 pass tile with the FP16
        // residual field against 1.7 us for the 32-row FP64 tile (depth 3,
        // 64-member ensemble, MI355X); the FP32 carry tile, without the
        // residual field, is weighted as the FP64 tile
        const long long cost = (long long)planes * (t.precision == 1 ? 7 : 5);
        for (int ty = 0; ty < ty_n; ++ty)
          for (int tx = 0; tx < tx_n; ++tx) jobs.push_back({cost, {m, tx, ty, tz}});
      }
    }
    std::stable_sort(jobs.begin(), jobs.end(),
                     [](const Job &a, const Job &b) { return a.cost > b.cost; });
    std::vector<int4> w(jobs.size());
    for (size_t j = 0; j < jobs.size(); ++j) w[j] = jobs[j].d;
    return w;
  };
  std::vector<int4> lists[2];
  lists[0] = list(kb);
  lists[1] = kb > 1 ? list(1) : lists[0];
  int cus = 0;
  H3D_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, e->device));
  int n[2], blocks[2];
  for (int li = 0; li < 2; ++li) {
    if (lists[li].size() > 0x3fffffffu) return H3D_ELIMIT;
    n[li] = (int)lists[li].size();
    const int rc = occupancy_blocks(cus, li == 0 ? kb : 1, n[li], &blocks[li]);
    if (rc != H3D_OK) return rc;
  }
  // Each list dealt into the XCD queues in runs of as many consecutive
  // tiles as an XCD has workgroups, in snake order over the queues (0..7,
  // 7..0, ...) so that the costlier runs of each sweep spread evenly.
  std::vector<int4> all;
  std::vector<int> qoff(2 * 16, 0);
  for (int li = 0; li < 2; ++li) {
    const std::vector<int4> &w = lists[li];
    const size_t run = (size_t)std::max(1, (blocks[li] + ENS_QUEUES - 1) / ENS_QUEUES);
    std::vector<std::vector<int4>> q(ENS_QUEUES);
    for (size_t r = 0; r * run < w.size(); ++r) {
      const size_t sweep = r / ENS_QUEUES, pos = r % ENS_QUEUES;
      const size_t x = sweep % 2 == 0 ? pos : ENS_QUEUES - 1 - pos;
      const size_t end = std::min(w.size(), (r + 1) * run);
      q[x].insert(q[x].end(), w.begin() + r * run, w.begin() + end);
    }
    int *o = &qoff[li * 16];
    for (int x = 0; x < ENS_QUEUES; ++x) {
      o[x] = (int)all.size();
      all.insert(all.end(), q[x].begin(), q[x].end());
    }
    o[ENS_QUEUES] = (int)all.size();
  }
  if (all.size() > 0x7fffffffu) return H3D_ELIMIT;
  const size_t total = all.size();
  if ((int)total > e->work_cap) {
    if (e->d_work) (void)hipFree(e->d_work);
    e->d_work = nullptr;
    e->work_cap = 0;
    H3D_TRY(hipMalloc(&e->d_work, sizeof(int4) * std::max<size_t>(total, 1)));
    e->work_cap = (int)total;
  }
  if (total)
    H3D_TRY(hipMemcpy(e->d_work, all.data(), sizeof(int4) * total, hipMemcpyHostToDevice));
  H3D_TRY(hipMemcpy(e->d_qoff, qoff.data(), sizeof(int) * qoff.size(), hipMemcpyHostToDevice));
  for (int li = 0; li < 2; ++li) {
    e->n[li] = n[li];
    e->blocks[li] = blocks[li];
  }
  e->zchunk = zchunk;
  e->kb = kb;
  drop_graph(e);
  return H3D_OK;
}

template <int KB>
static int launch_one(h3d_ensemble *e, int li, int parity, hipStream_t s) {
  if (e->n[li] == 0) return H3D_OK;
  const int lrc = allow_lds<KB>();
  if (lrc != H3D_OK) return lrc;
  hipLaunchKernelGGL(ensemble_step<KB>, dim3(e->blocks[li]), dim3(ENS_WAVES * 64),
                     ens_lds<KB>(), s, e->d_mem, e->d_work, e->d_qoff + li * 16,
                     e->d_qcnt + li * 16, parity, e->zchunk);
  return h3d_launch_status();
}

// Passes (launches) in a run of `steps` steps: full groups plus single
// steps.
static int ens_passes(const h3d_ensemble *e, int steps) {
  const int full = e->kb > 1 ? steps / e->kb : 0;
  return full + (steps - full * e->kb);
}

// One group on s: the kernel over list 0 (every tile, passes of kb steps)
// or list 1 (every tile, single steps).
static int ens_launch(h3d_ensemble *e, int list, int parity, hipStream_t s) {
  const int depth = list == 0 ? e->kb : 1;
  return depth == 1   ? launch_one<1>(e, list, parity, s)
         : depth == 2 ? launch_one<2>(e, list, parity, s)
                      : launch_one<3>(e, list, parity, s);
}

// Issues the launches of `steps` steps from buffer parity p.
static int ens_issue(h3d_ensemble *e, int steps, int p, hipStream_t s) {
  int rc = H3D_OK, np = 0, k = 0;
  if (e->kb > 1)
    for (; rc == H3D_OK && k + e->kb <= steps; k += e->kb, ++np)
      rc = ens_launch(e, 0, (p + np) % 2, s);
  for (; rc == H3D_OK && k < steps; ++k, ++np) rc = ens_launch(e, 1, (p + np) % 2, s);
  return rc;
}

extern "C" int h3d_ensemble_create(h3d_ensemble **out,
                                   const h3d_member *members, int count,
                                   const h3d_tune *tune, int device_id) {
  if (!out || !members || count < 1) return H3D_EARG;
  *out = nullptr;
  const int kb = (tune && tune->kblock > 0) ? tune->kblock : 1;
  if (kb > ENS_MAXKB) return H3D_ELIMIT;
  int ndev = 0;
  H3D_TRY(hipGetDeviceCount(&ndev));
  if (device_id < 0 || device_id >= ndev) return H3D_ENODEV;
  H3D_TRY(hipSetDevice(device_id));
  for (int m = 0; m < count; ++m) {
    const h3d_member &x = members[m];
    if (!x.u_a || !x.u_b || !x.kappa || x.u_a == x.u_b || !h3d_grid_ok(&x.g))
      return H3D_EARG;
    if (x.precision < 0 || x.precision > 2) return H3D_EARG;
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
    for (int d = 1; d <= ENS_MAXKB; ++d) {
      // FP64 single steps: the step tile, 64 x (rows x waves) cells; every
      // pass tile (FP32 at any depth, FP64 from depth 2) keeps
      // (64 - 2d) x (RY x waves - 2d) of its cells
      const bool step64 = x.precision == 0 && d == 1;
      const int rows = (x.precision ? ENS32_RY : ENS64_PRY) * ENS_WAVES;
      const int OX = step64 ? 64 : 64 - 2 * d;
      const int OY = step64 ? ENS64_ROWS * ENS_WAVES : rows - 2 * d;
      t.tiles_x[d] = (x.g.nx - 2 + OX - 1) / OX;
      t.tiles_y[d] = (x.g.ny - 2 + OY - 1) / OY;
    }
    if (x.precision == 1) {
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
  // queue offsets and counters; the counters start at zero and every
  // launch leaves them at zero
  hipError_t he = hipMalloc(&e->d_qoff, sizeof(int) * 2 * 16);
  if (he == hipSuccess) he = hipMalloc(&e->d_qcnt, sizeof(int) * 2 * 16);
  if (he == hipSuccess) he = hipMemset(e->d_qcnt, 0, sizeof(int) * 2 * 16);
  if (he == hipSuccess) he = hipMalloc(&e->d_mem, sizeof(EMember) * count);
  if (he == hipSuccess)
    he = hipMemcpy(e->d_mem, host.data(), sizeof(EMember) * count, hipMemcpyHostToDevice);
  if (he != hipSuccess) { ensemble_free(e); return (int)he; }
  const int zc = (tune && tune->zchunk > 0) ? tune->zchunk : ENS_ZCHUNK;
  rc = build_work(e, zc, kb);
  if (rc != H3D_OK) { ensemble_free(e); return rc; }
  *out = e;
  return H3D_OK;
}

extern "C" int h3d_ensemble_run(h3d_ensemble *e, int steps, int graph,
                                void *stream, float *ms) {
  if (!e || steps < 0) return H3D_EARG;
  hipStream_t s = h3d_stream(stream);
  if (graph && !s) return H3D_EARG;  // capture needs a created stream
  H3D_TRY(hipSetDevice(e->device));
  const int p = (int)(e->passes % 2);
  if (graph && steps > 0 && (e->graph_steps != steps || e->graph_p != p || !e->graph)) {
    drop_graph(e);
    H3D_TRY(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int rc = ens_issue(e, steps, p, s);
    hipGraph_t gdef = nullptr;
    hipError_t ec = hipStreamEndCapture(s, &gdef);
    if (rc != H3D_OK) { if (gdef) (void)hipGraphDestroy(gdef); return rc; }
    H3D_TRY(ec);
    e->graph_def = gdef;
    H3D_TRY(hipGraphInstantiate(&e->graph, gdef, nullptr, nullptr, 0));
    e->graph_steps = steps;
    e->graph_p = p;
  }
  h3d_run_res R;
  H3D_TRY(hipEventCreate(&R.e0));
  H3D_TRY(hipEventCreate(&R.e1));
  H3D_TRY(hipEventRecord(R.e0, s));
  if (steps > 0) {
    if (graph) {
      H3D_TRY(hipGraphLaunch(e->graph, s));
    } else {
      const int rc = ens_issue(e, steps, p, s);
      if (rc != H3D_OK) return rc;
    }
  }
  H3D_TRY(hipEventRecord(R.e1, s));
  H3D_TRY(hipEventSynchronize(R.e1));
  float t = 0.f;
  H3D_TRY(hipEventElapsedTime(&t, R.e0, R.e1));
  if (ms) *ms = t;
  e->done += steps;
  e->passes += ens_passes(e, steps);
  return H3D_OK;
}

// Times candidate z-march depths and pass depths 1..3, and keeps the
// fastest per step.  A depth-K group is one launch over every tile, timed
// with the buffer parity of the ensemble's next step, so it reads the
// current fields and rewrites only the idle buffers with values the next
// run recomputes: member state is unchanged.
extern "C" int h3d_ensemble_autotune(h3d_ensemble *e, void *stream, h3d_tune *out) {
  if (!e) return H3D_EARG;
  hipStream_t s = h3d_stream(stream);
  H3D_TRY(hipSetDevice(e->device));
  const int p = (int)(e->passes % 2);
  h3d_run_res R;
  H3D_TRY(hipEventCreate(&R.e0));
  H3D_TRY(hipEventCreate(&R.e1));
  auto time_list = [&](int list, float *best) {
    int rc = ens_launch(e, list, p, s);  // warm-up
    *best = 1e30f;
    for (int r = 0; r < 3 && rc == H3D_OK; ++r) {
      float ms = 0.f;
      hipError_t he = hipEventRecord(R.e0, s);
      if (he == hipSuccess) rc = ens_launch(e, list, p, s);
      if (rc != H3D_OK) break;
      if (he == hipSuccess) he = hipEventRecord(R.e1, s);
      if (he == hipSuccess) he = hipEventSynchronize(R.e1);
      if (he == hipSuccess) he = hipEventElapsedTime(&ms, R.e0, R.e1);
      if (he != hipSuccess) return (int)he;
      *best = std::min(*best, ms);
    }
    return rc;
  };
  int best_zc = e->zchunk, best_kb = e->kb, rc = H3D_OK;
  float best = 1e30f;
  // every depth is a candidate: FP64-only ensembles gain from depth 2 too
  for (int kb = 1; kb <= ENS_MAXKB && rc == H3D_OK; ++kb)
    for (int zc : {32, 64, 96, 128, 192, 256}) {
      rc = build_work(e, zc, kb);
      if (rc != H3D_OK) break;
      float tg = 0.f;
      rc = time_list(0, &tg);
      if (rc != H3D_OK) break;
      const float per_step = tg / kb;
      if (per_step < best) { best = per_step; best_zc = zc; best_kb = kb; }
    }
  if (rc != H3D_OK) return rc;
  rc = build_work(e, best_zc, best_kb);
  if (rc != H3D_OK) return rc;
  if (out) {
    out->rows = 0;
    out->waves = 0;  // each precision runs its own workgroup shape
    out->zchunk = best_zc;
    out->kblock = best_kb;
  }
  return H3D_OK;
}

extern "C" int h3d_ensemble_where(const h3d_ensemble *e, int member, int *in_b) {
  if (!e || !in_b || member < 0 || member >= e->count) return H3D_EARG;
  *in_b = (int)(e->passes % 2);
  return H3D_OK;
}

extern "C" int h3d_ensemble_info(const h3d_ensemble *e, int *nwork,
                                 int *nblocks, long long *steps_done) {
  if (!e) return H3D_EARG;
  if (nwork) *nwork = e->n[0];
  if (nblocks) *nblocks = e->blocks[0];
  if (steps_done) *steps_done = e->done;
  return H3D_OK;
}

extern "C" int h3d_ensemble_destroy(h3d_ensemble *e) {
  if (!e) return H3D_EARG;
  ensemble_free(e);
  return H3D_OK;
}
