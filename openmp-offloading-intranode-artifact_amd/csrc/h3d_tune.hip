// Autotuner: times candidate launch shapes of the FP64 step or the FP32
// pass on the caller's problem and keeps the fastest per problem shape.
//
// Candidates are timed with HIP events on the caller's stream, through
// two scratch fields seeded with a copy of u_a (the caller's fields are
// only read): launch j reads u_a (j = 0) or the previous launch's output
// and writes the other scratch field, as the steps of a run do.  Every
// candidate gets one warm-up launch and the best of three samples of two
// launches; the eight fastest are then timed again over samples of ten
// launches, and the best of those wins (two-launch samples alone favoured
// a z chunk that runs 8% slower over a 1000-step run).  The figure of
// merit is time per time step (pass time / kblock).
#include <algorithm>
#include <mutex>
#include <vector>

#include "h3d_internal.h"

namespace {

struct Key {
  int kind, nx, ny, nz, sy, mode, bf16, fixed_kb;
  long long sz;
  bool operator==(const Key &o) const {
    return kind == o.kind && nx == o.nx && ny == o.ny && nz == o.nz &&
           sy == o.sy && sz == o.sz && mode == o.mode && bf16 == o.bf16 &&
           fixed_kb == o.fixed_kb;
  }
};

// The winner per problem key, with its device time per pass.
struct Entry {
  Key key;
  h3d_tune t;
  float ms_per_pass;
};
std::mutex cache_mu;
std::vector<Entry> cache;

// Best of three samples of `nl` launches (ms per launch), after one
// warm-up launch; launch(j) is the j-th launch of a sample.
template <typename F>
int time_best(hipStream_t s, F launch, int nl, float *best_ms) {
  hipEvent_t e0, e1;
  H3D_TRY(hipEventCreate(&e0));
  H3D_TRY(hipEventCreate(&e1));
  int rc = launch(0);  // warm-up (also rejects invalid shapes)
  float best = 1e30f;
  for (int r = 0; rc == H3D_OK && r < 3; ++r) {
    float ms = 0;
    hipError_t e = hipEventRecord(e0, s);
    for (int j = 0; e == hipSuccess && rc == H3D_OK && j < nl; ++j) rc = launch(j);
    if (rc != H3D_OK) break;
    if (e == hipSuccess) e = hipEventRecord(e1, s);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e != hipSuccess) { rc = (int)e; break; }
    if (ms / nl < best) best = ms / nl;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *best_ms = best;
  return rc;
}

}  // namespace

int h3d_round_zchunks(int planes, int kblock, long long tiles, int wg_per_cu, int *zc, int n,
                      int *got) {
  if (planes < 1 || tiles < 1 || !zc || n < 1 || !got) return H3D_EARG;
  *got = 0;
  int dev = 0, cus = 0;
  H3D_TRY(hipGetDevice(&dev));
  H3D_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  const long long slots = (long long)(wg_per_cu > 0 ? wg_per_cu : 1) * (cus > 0 ? cus : 1);
  // A workgroup marches z + 2*kblock planes plus a start-up worth about
  // 8 more (its LDS clear and first loads).  The launch costs its total
  // march over all slots, but no less than 0.85 of the makespan of whole
  // rounds: a round with fewer workgroups than slots runs them faster
  // (they share the bandwidth), yet not in proportion.  Measured on MI355X
  // at 512^3 (3-step FP64 pass, 180 tiles): z chunks of 128, 103, 86 and
  // 73 planes run in the ratios 1 : 1.02 : 1.06 : 1.07 this predicts, and
  // 2 chunks (1.4 rounds) far slower.  Chunks are at most 256 planes,
  // except one chunk of the whole z range when every tile then has a
  // resident slot (one round, no warm-up planes re-read): at 1024^3 marches
  // of 512 and 1024 planes over several rounds ran 8-22% slower than the
  // model predicts, while at 512^3 the 4-step pass on 8 waves x 4 rows
  // (220 tiles) ran 985 us per pass as one whole-z round against 1009 us in
  // two chunks of 256 and 1058 us in eight of 64, as predicted.
  struct C {
    double cost;
    long long rounds;
    int z;
  };
  std::vector<C> v;
  int last = -1;
  for (int c = 1; c <= planes; ++c) {
    const int z = (planes + c - 1) / c;
    if (z == last) continue;
    last = z;
    if (c > 1 && z < 4 * kblock) break;  // warm-up would exceed half the work
    if (z > 256 && !(c == 1 && tiles <= slots)) continue;
    const long long wgs = tiles * ((planes + z - 1) / z);
    const long long rounds = (wgs + slots - 1) / slots;
    const double march = (double)(z + 2 * kblock + 8);
    const double cost = std::max((double)wgs * march / (double)slots, 0.85 * rounds * march);
    v.push_back({cost, rounds, z});
  }
  // cheapest first; equal cost: fewer rounds (neighbouring tiles march
  // together in one round, so their shared halo lines meet in L2)
  std::stable_sort(v.begin(), v.end(), [](const C &a, const C &b) {
    return a.cost != b.cost ? a.cost < b.cost : a.rounds < b.rounds;
  });
  int k = 0;
  for (size_t i = 0; i < v.size() && k < n; ++i) zc[k++] = v[i].z;
  *got = k;
  return H3D_OK;
}

extern "C" int h3d_autotune(int kind, void *u_a, void *u_b, const void *kappa,
                            int kappa_bf16, const h3d_grid *g, int mode,
                            int fixed_kblock, void *stream, h3d_tune *out) {
  (void)u_b;
  if (!u_a || !kappa || !out || !h3d_grid_ok(g)) return H3D_EARG;
  if (kind != 0 && kind != 1) return H3D_EARG;
  if (kind == 1 && (mode < 0 || mode > 2)) return H3D_EARG;
  if (kind == 0 && (mode < 0 || mode > 1)) return H3D_EARG;
  if (fixed_kblock < 0 || fixed_kblock > (kind ? 8 : 4)) return H3D_ELIMIT;
  const Key key{kind, g->nx, g->ny, g->nz, g->sy, mode,
                kind ? kappa_bf16 : 0, fixed_kblock, g->sz};
  {
    std::lock_guard<std::mutex> lk(cache_mu);
    for (auto &e : cache)
      if (e.key == key) { *out = e.t; return H3D_OK; }
  }
  hipStream_t s = h3d_stream(stream);
  const size_t cells = h3d_grid_elems(g);
  const size_t esz = kind == 0 ? sizeof(double) : sizeof(float);
  // scratch fields at u_a's offset within 256 bytes: the kernels' tile
  // origins and line crossings follow the field's address, so candidates
  // are timed on the alignment the caller's run will have
  void *base[2] = {nullptr, nullptr}, *scr[2] = {nullptr, nullptr};
  void *c_in = nullptr, *c_out = nullptr;
  const uintptr_t align_off = (uintptr_t)u_a % 256;
  hipError_t e = hipMalloc(&base[0], cells * esz + 256);
  if (e == hipSuccess) e = hipMalloc(&base[1], cells * esz + 256);
  for (int i = 0; i < 2 && e == hipSuccess; ++i) {
    scr[i] = (void *)((uintptr_t)base[i] + align_off);  // allocations are 256-B aligned
    e = hipMemcpyAsync(scr[i], u_a, cells * esz, hipMemcpyDeviceToDevice, s);
  }
  if (e == hipSuccess && kind == 1 && mode == 2) {
    e = hipMalloc(&c_in, cells * 2);
    if (e == hipSuccess) e = hipMalloc(&c_out, cells * 2);
    if (e == hipSuccess) e = hipMemsetAsync(c_in, 0, cells * 2, s);
  }
  h3d_tune best{};
  float best_ms = 1e30f;
  int rc = (e == hipSuccess) ? H3D_OK : (int)e;
  // launch j of a sample: reads u_a (j = 0) or the previous output,
  // writes the other scratch field (residual fields alternate likewise)
  const double *rhs = kind == 0 && mode == 1 ? (const double *)u_a : nullptr;  // u_a stands in
  const float *kf = kappa_bf16 ? nullptr : (const float *)kappa;
  const uint16_t *kb16 = kappa_bf16 ? (const uint16_t *)kappa : nullptr;
  auto launcher = [&](const h3d_tune &t) {
    return [&, t](int j) {
      void *out = scr[j % 2];
      const void *in = j == 0 ? u_a : scr[(j + 1) % 2];
      if (kind == 0)
        return h3d_pass_f64((double *)out, (const double *)in, rhs, (const double *)kappa,
                            g, 1e-3, t.kblock, 1, g->nz - 2, &t, s);
      return h3d_pass_f32((float *)out, (const float *)in, kf, kb16,
                          (uint16_t *)(j % 2 ? c_in : c_out),
                          (const uint16_t *)(j % 2 ? c_out : c_in), g, 1e-3f, t.kblock, mode,
                          &t, s);
    };
  };
  std::vector<std::pair<float, h3d_tune>> cand;  // (ms per step, shape)
  if (rc == H3D_OK && kind == 0) {
    const int planes = g->nz - 2;
    // depths 1..4 when the tuner chooses; a fixed depth alone
    for (int kbk = 1; kbk <= (fixed_kblock ? fixed_kblock : 4) && rc == H3D_OK; ++kbk) {
      if (fixed_kblock && kbk != fixed_kblock) continue;
      // single steps: whole z range first (on large grids the chunk-edge
      // re-reads cost more than the extra parallelism of short chunks
      // buys); fused passes: the LDS-shared tiles with short z chunks
      const std::vector<int> rws = kbk == 1 ? std::vector<int>{1, 2, 4} : std::vector<int>{2, 4};
      const std::vector<int> wvs =
          kbk == 1 ? std::vector<int>{2, 4, 8} : std::vector<int>{4, 8, 16};
      for (int rows : rws)
        for (int waves : wvs) {
          // fused passes: 32..128-plane chunks and the chunks that fill
          // whole rounds of resident workgroups (h3d_round_zchunks)
          std::vector<int> zcs = kbk == 1 ? std::vector<int>{planes, 256, 128, 64, 32}
                                          : std::vector<int>{32, 64, 128};
          if (kbk > 1) {
            int occ = 0, z[3], nz3 = 0;
            long long tl = 0;
            if (h3d_pass_f64_geometry(g, kbk, rows, waves, mode, &occ, &tl) != H3D_OK) continue;
            if (h3d_round_zchunks(planes, kbk, tl, occ, z, 3, &nz3) != H3D_OK) nz3 = 0;
            for (int i = 0; i < nz3; ++i)
              if (std::find(zcs.begin(), zcs.end(), z[i]) == zcs.end()) zcs.push_back(z[i]);
          }
          for (int zc : zcs) {
            if (rc != H3D_OK) break;
            if (zc > planes && zc != zcs.front()) continue;
            h3d_tune t{rows, waves, zc, kbk};
            float ms;
            int r = time_best(s, launcher(t), 2, &ms);
            if (r != H3D_OK && r != H3D_EARG && r != H3D_ELIMIT) { rc = r; break; }
            if (r == H3D_OK) cand.push_back({ms / kbk, t});
          }
        }
    }
  } else if (rc == H3D_OK) {
    const int planes = g->nz - 2;
    // depths 1..4 when the tuner chooses; a fixed depth (up to 8) alone
    for (int kbk = 1; kbk <= (fixed_kblock ? fixed_kblock : 4) && rc == H3D_OK; ++kbk) {
      if (fixed_kblock && kbk != fixed_kblock) continue;
      // single-step passes use the wave-independent tiles (rows 1-4, any
      // wave count, long z marches); fused passes the LDS-shared tiles in
      // every shape the depth offers, each with the z chunks that fill
      // whole rounds of resident workgroups (h3d_round_zchunks), 64 and 128
      std::vector<std::pair<int, int>> shapes;
      if (kbk == 1) {
        for (int rows : {1, 2, 4})
          for (int waves : {2, 4, 8}) shapes.push_back({rows, waves});
      } else {
        shapes = {{2, 4}, {2, 8}, {4, 4}, {4, 8}, {2, 16}, {3, 16},
                  {4, 16}, {5, 16}, {6, 16}, {3, 8}};
      }
      for (const auto &sh : shapes) {
        std::vector<int> zcs;
        if (kbk == 1) {
          zcs = {planes, 256, 128, 64, 32};
        } else {
          int occ = 0;
          long long tl = 0;
          if (h3d_pass_f32_geometry(g, kbk, sh.first, sh.second, mode, kappa_bf16, &occ, &tl) !=
              H3D_OK)
            continue;  // shape not offered at this depth
          int z[3], nz3 = 0;
          if (h3d_round_zchunks(planes, kbk, tl, occ, z, 3, &nz3) != H3D_OK) nz3 = 0;
          for (int i = 0; i < nz3; ++i) zcs.push_back(z[i]);
          for (int zf : {64, 128})
            if (std::find(zcs.begin(), zcs.end(), zf) == zcs.end() && planes > zf)
              zcs.push_back(zf);
        }
        for (int zc : zcs) {
          if (zc > planes && zc != zcs.front()) continue;
          h3d_tune t{sh.first, sh.second, zc, kbk};
          float ms;
          int r = time_best(s, launcher(t), 2, &ms);
          if (r != H3D_OK && r != H3D_EARG && r != H3D_ELIMIT) { rc = r; break; }
          if (r == H3D_OK) cand.push_back({ms / kbk, t});
        }
        if (rc != H3D_OK) break;
      }
    }
  }
  // the eight fastest again, over ten launches per sample (with four, the
  // n = 512 reference case with rhs kept a z chunk 1.5% slower than the
  // best: two-launch samples order shapes within a few percent by chance)
  std::stable_sort(cand.begin(), cand.end(),
                   [](const auto &a, const auto &b) { return a.first < b.first; });
  for (size_t c = 0; rc == H3D_OK && c < cand.size() && c < 8; ++c) {
    float ms;
    const int r = time_best(s, launcher(cand[c].second), 10, &ms);
    if (r != H3D_OK) { rc = r; break; }
    if (ms / cand[c].second.kblock < best_ms) {
      best_ms = ms / cand[c].second.kblock;
      best = cand[c].second;
    }
  }
  (void)hipStreamSynchronize(s);
  for (void *p : base)
    if (p) (void)hipFree(p);
  if (c_in) (void)hipFree(c_in);
  if (c_out) (void)hipFree(c_out);
  if (rc != H3D_OK) return rc;
  if (best_ms >= 1e30f) return H3D_ELIMIT;  // no candidate ran
  {
    std::lock_guard<std::mutex> lk(cache_mu);
    cache.push_back({key, best, best_ms * (float)best.kblock});
  }
  *out = best;
  // FP64 without rhs at a free depth of 2 or 3: the next depth as well, so
  // that h3d_run_f64 can weigh deeper passes against a short remainder
  // pass on measured times (h3d_tuned_f64)
  if (kind == 0 && mode == 0 && fixed_kblock == 0 && (best.kblock == 2 || best.kblock == 3)) {
    h3d_tune deeper;
    const int r2 = h3d_autotune(kind, u_a, u_b, kappa, kappa_bf16, g, mode, best.kblock + 1,
                                stream, &deeper);
    if (r2 != H3D_OK && r2 != H3D_ELIMIT) return r2;
  }
  return H3D_OK;
}

int h3d_tuned_f64(const h3d_grid *g, int mode, int kblock, h3d_tune *t, float *ms_per_pass) {
  if (!g) return H3D_EARG;
  std::lock_guard<std::mutex> lk(cache_mu);
  for (auto &e : cache) {
    const Key &k = e.key;
    if (k.kind != 0 || k.nx != g->nx || k.ny != g->ny || k.nz != g->nz || k.sy != g->sy ||
        k.sz != g->sz || k.mode != mode || e.t.kblock != kblock)
      continue;
    if (k.fixed_kb != 0 && k.fixed_kb != kblock) continue;
    if (t) *t = e.t;
    if (ms_per_pass) *ms_per_pass = e.ms_per_pass;
    return H3D_OK;
  }
  return H3D_ELIMIT;
}
