// Batched solves and continuous batching (the ausm_batch_* entry points).
//
// No counterpart in the reference, which runs one grid at a time; each
// solve here runs exactly the single-grid step (the same tile code and the
// same per-solve dt rule), so a solve's state is bitwise the same alone and
// inside any batch.
//
// Runtime:
//  * one device arena (first-fit, coalescing) holds every solve's buffers;
//    a queued solve is admitted when a table slot is free and its buffers
//    fit in the arena (admission by HBM budget);
//  * the batch advances in chunks of `check_every` steps; a chunk is one
//    captured graph of (dt, stage 1, stage 2, stage 3) launches per step;
//    solves join and leave only at chunk boundaries, so every solve's
//    convergence checks fall on multiples of check_every of its own steps;
//  * after each chunk one kernel reduces every solve's FP64 residual norm;
//    converged / capped / non-finite solves retire (forces computed in
//    FP64), queued solves are admitted into the freed slots and memory;
//  * the tile shape (rows per wavefront) and the persistent grid size
//    (workgroups per CU) are picked per size mix by timing candidates on
//    scratch buffers, and cached per mix class.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

#include "aux_kernels.hpp"
#include "ausm.h"
#include "batch_kernels.hpp"
#include "ctx.hpp"
#include "kernels.hpp"

using namespace ausm;

namespace ausm {

static size_t rup(size_t x, size_t a) { return (x + a - 1) / a * a; }

void make_freestream(DevParams& p, double rho, double pr, double vx, double vy, double gamma);

// First-fit allocator over one device allocation.
class Arena {
 public:
  void init(char* base, size_t size) {
    base_ = base;
    size_ = size;
    free_.clear();
    free_[0] = size;
  }
  bool alloc(size_t n, size_t& off) {
    n = rup(n, 256);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        off = it->first;
        const size_t rem = it->second - n, noff = it->first + n;
        free_.erase(it);
        if (rem) free_[noff] = rem;
        return true;
      }
    }
    return false;
  }
  void release(size_t off, size_t n) {
    n = rup(n, 256);
    auto it = free_.emplace(off, n).first;
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  size_t free_bytes() const {
    size_t s = 0;
    for (auto& kv : free_) s += kv.second;
    return s;
  }
  char* base() const { return base_; }
  size_t size() const { return size_; }

 private:
  char* base_ = nullptr;
  size_t size_ = 0;
  std::map<size_t, size_t> free_;
};

struct Slot {
  uint64_t id = 0;
  ausm_solve_desc d{};
  std::vector<float> verts;
  int state = 0;  // 0 queued, 1 running, 2 converged, 3 capped, 4 failed
  long long steps = 0;
  double res = -1.0, res0 = -1.0, cl = 0.0, cd = 0.0;
  size_t off = 0, bytes = 0;
  int index = -1;
  bool frozen = false;  // limiter frozen (freeze_after reached)
  SolveDev dev{};
  DevParams* dpar = nullptr;  // the solve's own parameter copy in the arena
};

struct BGraph {
  int steps, ty, blocks, math, comp, frz;
  hipGraph_t g;
  hipGraphExec_t e;
};

// sizes of one solve's arena block
struct Layout {
  int P, PS, tiles_x, nws_max;
  size_t q, v, ws, res, par, comp, psi, total;
};

static const int kBatchTiles[] = {1, 2, 4};

// comp: two more state-sized plane sets (accumulator, low part); frz: the
// limiter-freezing psi planes
static Layout layout_of(int M, int N, bool comp, bool frz = false) {
  Layout L{};
  L.P = (int)rup((size_t)M, 64);
  L.PS = (int)rup((size_t)N * L.P, 64);
  L.tiles_x = stage_tiles_x(M);
  L.nws_max = 0;
  for (int ty : kBatchTiles) L.nws_max = std::max(L.nws_max, stage_blocks(M, N, ty));
  L.nws_max = std::max(L.nws_max, wave_speed_blocks(M, N));
  L.q = rup(sizeof(float) * 4 * (size_t)L.PS, 256);
  L.v = rup(sizeof(float) * (size_t)(N + 1) * L.P, 256);
  L.ws = rup(sizeof(float) * (size_t)L.nws_max, 256);
  L.res = rup(sizeof(double) * (size_t)N * L.tiles_x, 256);
  L.par = rup(sizeof(DevParams) + sizeof(float), 256);
  L.comp = comp ? 2 * L.q : 0;
  L.psi = frz ? rup(sizeof(float) * 2 * kPsiPerFace * psi_plane(N, L.P), 256) : 0;
  L.total = 3 * L.q + 2 * L.v + L.ws + L.res + L.par + L.comp + L.psi;
  return L;
}

}  // namespace ausm

struct ausm_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int max_solves = 0;
  int cus = 0;
  char* arena_mem = nullptr;
  Arena arena;
  std::map<uint64_t, Slot*> all;
  std::deque<Slot*> queue;
  std::vector<Slot*> table;
  std::vector<SolveDev> h_solves;
  std::vector<BatchItem> h_items;
  SolveDev* d_solves = nullptr;
  BatchItem* d_items = nullptr;
  int* d_counts = nullptr;  // [0] items, [1] table slots
  double* d_out = nullptr;  // 2 per slot: residual norm, max wave speed
  double* h_out = nullptr;  // pinned
  float* d_tmp = nullptr;   // admission scratch (wave-speed partials)
  int tmp_cap = 0;
  double* d_tmpd = nullptr;  // 8 doubles: norms, forces
  double* h_tmpd = nullptr;  // pinned
  float* d_stage = nullptr;  // get_state staging
  size_t stage_cap = 0;
  double* d_seg = nullptr;   // per-segment forces of a retiring solve (2 M doubles)
  size_t seg_cap = 0;
  bool use_graphs = true;    // AUSM_NO_GRAPHS set: chunks are launched directly
  int max_items = 0;
  int n_frozen = 0;  // running solves with the limiter frozen (kernel variant)
  int n_items = 0;
  bool dirty = false;
  int math = AUSM_MATH_FAST;
  bool comp = false;  // AUSM_PREC_FP32C: compensated-state stages
  int fixed_ty = 0, fixed_bpc = 0;
  int ty = 2, bpc = 2;
  std::map<int, std::pair<int, int>> tuned;  // mix class -> (ty, bpc)
  uint64_t next_id = 1;
  std::vector<BGraph> graphs;
};

namespace ausm {

static void drop_bgraphs(ausm_batch* b) {
  // a launched graph may still be running on the stream
  if (!b->graphs.empty()) (void)hipStreamSynchronize(b->stream);
  for (auto& g : b->graphs) {
    (void)hipGraphExecDestroy(g.e);
    (void)hipGraphDestroy(g.g);
  }
  b->graphs.clear();
}

static int n_running(ausm_batch* b) {
  int n = 0;
  for (auto* s : b->table)
    if (s) ++n;
  return n;
}

// rebuild the item list for tile shape ty: running solves sorted by scheme
static int build_items(ausm_batch* b, int ty) {
  b->h_items.clear();
  std::vector<Slot*> run;
  for (auto* s : b->table)
    if (s) run.push_back(s);
  std::stable_sort(run.begin(), run.end(),
                   [](Slot* x, Slot* y) { return x->d.scheme < y->d.scheme; });
  for (auto* s : run) {
    const int nb = stage_blocks(s->d.M, s->d.N, ty);
    b->h_solves[s->index].n_ws = nb;
    for (int t = 0; t < nb; ++t) b->h_items.push_back(BatchItem{s->index, t});
  }
  if ((int)b->h_items.size() > b->max_items) {
    (void)hipFree(b->d_items);
    b->max_items = (int)b->h_items.size() * 2;
    if (hipMalloc(&b->d_items, sizeof(BatchItem) * b->max_items) != hipSuccess)
      return set_error(AUSM_ERR_NOMEM, "item list allocation failed");
    drop_bgraphs(b);  // graphs hold the old item pointer
  }
  b->n_items = (int)b->h_items.size();
  const int counts[2] = {b->n_items, (int)b->table.size()};
  if (b->n_items)
    if (hipMemcpyAsync(b->d_items, b->h_items.data(), sizeof(BatchItem) * b->n_items,
                       hipMemcpyHostToDevice, b->stream) != hipSuccess)
      return set_error(AUSM_ERR_HIP, "item upload failed");
  if (hipMemcpyAsync(b->d_counts, counts, sizeof(counts), hipMemcpyHostToDevice, b->stream) !=
          hipSuccess ||
      hipMemcpyAsync(b->d_solves, b->h_solves.data(), sizeof(SolveDev) * b->table.size(),
                     hipMemcpyHostToDevice, b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess)
    return set_error(AUSM_ERR_HIP, "batch table upload failed");
  return AUSM_OK;
}

// wave-speed partials of a solve's current state, seeded in stage-3 layout
static int seed_ws(ausm_batch* b, Slot* s) {
  const int M = s->d.M, N = s->d.N;
  const int nb = wave_speed_blocks(M, N);
  if (nb > b->tmp_cap) {
    (void)hipFree(b->d_tmp);
    b->tmp_cap = nb;
    if (hipMalloc(&b->d_tmp, sizeof(float) * nb) != hipSuccess)
      return set_error(AUSM_ERR_NOMEM, "scratch allocation failed");
  }
  const Layout L = layout_of(M, N, b->comp);
  if (launch_wave_speed(s->dev.q[0], s->dev.vx, s->dev.vy, s->dpar, M, N, L.P, L.PS, (float*)nullptr,
                        b->d_tmp, b->stream) != hipSuccess ||
      launch_batch_seed_ws(b->d_tmp, nb, s->dev.ws_part, L.nws_max, b->stream) != hipSuccess)
    return set_error(AUSM_ERR_HIP, "wave-speed seeding failed");
  return AUSM_OK;
}

// stage-1 pass of one solve into its scratch Q1: residual partials of the
// current state and maxws_cur; returns the FP64 density-residual norm
static int solve_norm(ausm_batch* b, Slot* s, double* out) {
  StageArgs a = batch_args(s->dev, 1, 0, 1);
  a.par = s->dpar;
  a.fresh = 1;  // reduce the seeded maxima into maxws_cur
  a.n_ws_in = layout_of(s->d.M, s->d.N, b->comp).nws_max;
  a.local_dt = s->d.dt_mode == AUSM_DT_LOCAL;
  if (launch_stage(a, s->d.scheme, b->math, 2, b->stream) != hipSuccess ||
      launch_reduce_sum_f64(s->dev.res_part, s->d.N * s->dev.tiles_x, b->d_tmpd, b->stream) !=
          hipSuccess ||
      hipMemcpyAsync(b->h_tmpd, b->d_tmpd, sizeof(double), hipMemcpyDeviceToHost, b->stream) !=
          hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess)
    return set_error(AUSM_ERR_HIP, "initial residual failed");
  *out = b->h_tmpd[0];
  return AUSM_OK;
}

static int admit(ausm_batch* b) {
  while (!b->queue.empty()) {
    int idx = -1;
    for (int i = 0; i < (int)b->table.size(); ++i)
      if (!b->table[i]) {
        idx = i;
        break;
      }
    if (idx < 0) return AUSM_OK;
    Slot* s = b->queue.front();
    const int M = s->d.M, N = s->d.N;
    const Layout L = layout_of(M, N, b->comp, s->d.freeze_after > 0);
    size_t off;
    if (!b->arena.alloc(L.total, off)) return AUSM_OK;  // wait for memory
    b->queue.pop_front();
    s->off = off;
    s->bytes = L.total;
    s->index = idx;
    char* p = b->arena.base() + off;
    SolveDev& D = s->dev;
    D = SolveDev{};
    for (int k = 0; k < 3; ++k) D.q[k] = (float*)(p + k * L.q);
    float* vx = (float*)(p + 3 * L.q);
    float* vy = (float*)(p + 3 * L.q + L.v);
    D.vx = vx;
    D.vy = vy;
    D.ws_part = (float*)(p + 3 * L.q + 2 * L.v);
    D.res_part = (double*)(p + 3 * L.q + 2 * L.v + L.ws);
    s->dpar = (DevParams*)(p + 3 * L.q + 2 * L.v + L.ws + L.res);
    D.maxws_cur = (float*)((char*)s->dpar + sizeof(DevParams));
    if (b->comp) {
      char* pc = p + 3 * L.q + 2 * L.v + L.ws + L.res + L.par;
      D.acc = (float*)pc;
      D.lo = (float*)(pc + L.q);
    }
    if (L.psi) D.psi = (float*)(p + 3 * L.q + 2 * L.v + L.ws + L.res + L.par + L.comp);
    D.M = M;
    D.N = N;
    D.P = L.P;
    D.PS = L.PS;
    D.tiles_x = L.tiles_x;
    D.scheme = s->d.scheme;
    D.limiter = s->d.limiter;
    D.n_ws = L.nws_max;
    D.active = 1;
    D.local_dt = s->d.dt_mode == AUSM_DT_LOCAL;
    make_freestream(D.par, s->d.rho, s->d.p, s->d.vx, s->d.vy, s->d.gamma);
    D.par.Kp = s->d.K_p;
    D.par.Ku = s->d.K_u;
    D.par.cfl = s->d.cfl;
    D.par.maxdt = s->d.maxdt;
    // vertices -> SoA planes
    std::vector<float> hx((size_t)(N + 1) * L.P, 0.0f), hy((size_t)(N + 1) * L.P, 0.0f);
    for (int j = 0; j <= N; ++j)
      for (int i = 0; i < M; ++i) {
        hx[(size_t)j * L.P + i] = s->verts[((size_t)j * M + i) * 2];
        hy[(size_t)j * L.P + i] = s->verts[((size_t)j * M + i) * 2 + 1];
      }
    if (hipMemcpyAsync(vx, hx.data(), sizeof(float) * hx.size(), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipMemcpyAsync(vy, hy.data(), sizeof(float) * hy.size(), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        hipMemcpyAsync(s->dpar, &D.par, sizeof(DevParams), hipMemcpyHostToDevice, b->stream) != hipSuccess ||
        launch_fill_state(D.q[0], s->dpar, M, N, L.P, L.PS, s->d.init_mode, b->stream) != hipSuccess ||
        (b->comp && hipMemsetAsync(D.lo, 0, L.q, b->stream) != hipSuccess))
      return set_error(AUSM_ERR_HIP, "solve upload failed");
    int r = seed_ws(b, s);
    if (r != AUSM_OK) return r;
    r = solve_norm(b, s, &s->res0);
    if (r != AUSM_OK) return r;
    s->res = s->res0;
    s->state = 1;
    s->steps = 0;
    b->table[idx] = s;
    b->h_solves[idx] = D;
    b->dirty = true;
  }
  return AUSM_OK;
}

// time one step's worth of stage launches over the current items (after
// one untimed launch): stages 1 and 2 as a step runs them, then the stage-2
// arithmetic from Q^n in place of stage 3.  They write only what every step
// rewrites before it reads it (Q1, Q2 and the compensated increment), so
// tuning changes no solve's results.
static float time_schedule(ausm_batch* b, int ty, int bpc) {
  const BatchArgs a1{b->d_solves, b->d_items, b->d_counts, 1, 0, 0};
  const BatchArgs a2{b->d_solves, b->d_items, b->d_counts, 2, 0, 0};
  const BatchArgs a4{b->d_solves, b->d_items, b->d_counts, 4, 0, 0};
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const int blocks = b->cus * bpc;
  (void)launch_batch_stage(a4, b->math, b->comp, b->n_frozen > 0, ty, blocks, b->stream);
  (void)hipEventRecord(e0, b->stream);
  for (const BatchArgs* a : {&a1, &a2, &a4})
    (void)launch_batch_stage(*a, b->math, b->comp, b->n_frozen > 0, ty, blocks, b->stream);
  (void)hipEventRecord(e1, b->stream);
  (void)hipEventSynchronize(e1);
  float ms = 1e30f;
  (void)hipEventElapsedTime(&ms, e0, e1);
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms;
}

static const int kBatchBpc[] = {1, 2, 4, 8};

// choose (rows per wavefront, workgroups per CU) for the current mix: two
// rounds over every candidate, a candidate's time its best round (the
// rounds interleave the candidates, so clock ramps and queue noise do not
// favour the ones timed last)
static int choose_schedule(ausm_batch* b) {
  long long cells = 0;
  for (auto* s : b->table)
    if (s) cells += (long long)s->d.M * s->d.N;
  int cls = 0;
  while ((1LL << cls) < cells) ++cls;
  int ty = b->fixed_ty, bpc = b->fixed_bpc;
  auto it = b->tuned.find(cls);
  if (it != b->tuned.end()) {
    if (!ty) ty = it->second.first;
    if (!bpc) bpc = it->second.second;
  }
  if (!ty || !bpc) {
    // a workgroup count is a candidate only while the one below it still
    // leaves some workgroup more than one item (beyond that the runs are
    // no shorter, only more workgroups find nothing to do)
    std::vector<std::pair<int, int>> cand;
    for (int t : kBatchTiles) {
      if (b->fixed_ty && t != b->fixed_ty) continue;
      if (b->fixed_bpc) {
        cand.emplace_back(t, b->fixed_bpc);
        continue;
      }
      long long items = 0;
      for (auto* s : b->table)
        if (s) items += stage_blocks(s->d.M, s->d.N, t);
      int prev = 0;
      for (int k : kBatchBpc) {
        if (prev && (long long)b->cus * prev >= items) break;
        cand.emplace_back(t, k);
        prev = k;
      }
    }
    std::vector<float> best_ms(cand.size(), 1e30f);
    for (int round = 0; round < 2; ++round) {
      int built = 0;
      for (size_t c = 0; c < cand.size(); ++c) {
        if (cand[c].first != built) {
          const int r = build_items(b, cand[c].first);
          if (r != AUSM_OK) return r;
          built = cand[c].first;
        }
        best_ms[c] = std::min(best_ms[c], time_schedule(b, cand[c].first, cand[c].second));
      }
    }
    size_t bi = 0;
    for (size_t c = 1; c < cand.size(); ++c)
      if (best_ms[c] < best_ms[bi]) bi = c;
    const int bty = cand[bi].first, bbpc = cand[bi].second;
    b->tuned[cls] = {bty, bbpc};
    ty = bty;
    bpc = bbpc;
  }
  if (ty != b->ty) {
    // partial layouts change with the tile shape: reseed every running solve
    for (auto* s : b->table)
      if (s) {
        const int r = seed_ws(b, s);
        if (r != AUSM_OK) return r;
      }
  }
  b->ty = ty;
  b->bpc = bpc;
  return build_items(b, ty);
}

static hipError_t enqueue_steps(ausm_batch* b, int nsteps) {
  const int blocks = b->cus * b->bpc;
  for (int s = 0; s < nsteps; ++s) {
    hipError_t e = launch_batch_dt(b->d_solves, b->d_counts + 1, (int)b->table.size(), b->stream);
    for (int st = 1; st <= 3 && e == hipSuccess; ++st) {
      // the convergence check after the chunk reads the last step's norm
      BatchArgs a{b->d_solves, b->d_items, b->d_counts, st, st == 3, st == 1 && s == nsteps - 1};
      e = launch_batch_stage(a, b->math, b->comp, b->n_frozen > 0, b->ty, blocks, b->stream);
    }
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

static int run_chunk(ausm_batch* b, int nsteps) {
  const int blocks = b->cus * b->bpc;
  if (!b->use_graphs) {
    if (enqueue_steps(b, nsteps) != hipSuccess) return set_error(AUSM_ERR_HIP, "stage launch failed");
    return AUSM_OK;
  }
  BGraph* g = nullptr;
  for (auto& x : b->graphs)
    if (x.steps == nsteps && x.ty == b->ty && x.blocks == blocks && x.math == b->math &&
        x.comp == (int)b->comp && x.frz == (b->n_frozen > 0))
      g = &x;
  if (!g) {
    BGraph n{nsteps, b->ty, blocks, b->math, (int)b->comp, b->n_frozen > 0, nullptr, nullptr};
    if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
      return set_error(AUSM_ERR_HIP, "graph capture failed");
    const hipError_t e = enqueue_steps(b, nsteps);
    const hipError_t e2 = hipStreamEndCapture(b->stream, &n.g);
    if (e != hipSuccess || e2 != hipSuccess) {
      if (n.g) (void)hipGraphDestroy(n.g);
      return set_error(AUSM_ERR_HIP, "graph capture failed");
    }
    if (hipGraphInstantiate(&n.e, n.g, nullptr, nullptr, 0) != hipSuccess) {
      (void)hipGraphDestroy(n.g);
      return set_error(AUSM_ERR_HIP, "graph instantiate failed");
    }
    b->graphs.push_back(n);
    g = &b->graphs.back();
  }
  if (hipGraphLaunch(g->e, b->stream) != hipSuccess) return set_error(AUSM_ERR_HIP, "graph launch failed");
  return AUSM_OK;
}

static int retire(ausm_batch* b, Slot* s, int state) {
  s->state = state;
  // FP64 forces at completion (graphs.js:16-101)
  if (state != 4) {
    // grow-only segment buffer (a retiring solve does not allocate unless it
    // is the largest seen so far)
    const size_t need = 2 * (size_t)s->d.M;
    if (need > b->seg_cap) {
      (void)hipStreamSynchronize(b->stream);
      (void)hipFree(b->d_seg);
      b->d_seg = nullptr;
      b->seg_cap = 0;
      if (hipMalloc(&b->d_seg, sizeof(double) * need) != hipSuccess)
        return set_error(AUSM_ERR_NOMEM, "force scratch allocation failed");
      b->seg_cap = need;
    }
    const double ref = s->d.ref_length > 0 ? s->d.ref_length : 1.0;
    hipError_t e = launch_forces(s->dev.q[0], s->dev.vx, s->dev.vy, s->dpar, s->d.M, s->dev.PS, ref,
                                 b->d_seg, b->d_tmpd, nullptr, b->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(b->h_tmpd, b->d_tmpd, 4 * sizeof(double), hipMemcpyDeviceToHost, b->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(b->stream);
    if (e != hipSuccess) return set_error(AUSM_ERR_HIP, "forces failed");
    s->cl = b->h_tmpd[2];
    s->cd = b->h_tmpd[3];
  }
  if (s->frozen) --b->n_frozen;
  b->h_solves[s->index].active = 0;
  b->table[s->index] = nullptr;
  s->index = -1;
  b->dirty = true;
  return AUSM_OK;
}

// limiter freezing of a running solve: psi recorded from its current
// state by a residual pass of the general stage kernel into the scratch Q2
// (the same record as ausm_freeze_limiter), then the solve's table entry
// switched to the frozen reconstruction
static int freeze_solve(ausm_batch* b, Slot* s) {
  StageArgs a = batch_args(s->dev, 4, 0, 0);  // Q^n in, scratch Q2 out
  a.par = s->dpar;
  a.stage = 0;
  a.freeze = 1;
  a.local_dt = 0;
  if (launch_stage(a, s->d.scheme, b->math, 2, b->stream) != hipSuccess)
    return set_error(AUSM_ERR_HIP, "limiter record failed");
  s->frozen = true;
  ++b->n_frozen;
  s->dev.freeze = 2;
  b->h_solves[s->index].freeze = 2;
  if (hipMemcpyAsync(b->d_solves + s->index, &b->h_solves[s->index], sizeof(SolveDev),
                     hipMemcpyHostToDevice, b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess)
    return set_error(AUSM_ERR_HIP, "limiter record failed");
  return AUSM_OK;
}

static void free_slot(ausm_batch* b, Slot* s) {
  if (s->index >= 0) {
    if (s->frozen) --b->n_frozen;
    b->h_solves[s->index].active = 0;
    b->table[s->index] = nullptr;
    b->dirty = true;
  }
  if (s->bytes) b->arena.release(s->off, s->bytes);
  s->bytes = 0;
}

}  // namespace ausm

extern "C" {

int ausm_batch_create(int device, int max_solves, size_t budget, ausm_batch** out) {
  if (!out) return set_error(AUSM_ERR_ARG, "out is NULL");
  *out = nullptr;
  if (max_solves < 1 || max_solves > 65536) return set_error(AUSM_ERR_ARG, "max_solves out of range");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return set_error(AUSM_ERR_ARG, "device %d not present", device);
  if (hipSetDevice(device) != hipSuccess) return set_error(AUSM_ERR_HIP, "hipSetDevice failed");
  const int dr = device_arithmetic_ok(device);
  if (dr != AUSM_OK) return dr;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return set_error(AUSM_ERR_HIP, "hipMemGetInfo failed");
  if (budget == 0) budget = std::min(fr / 2, (size_t)16 << 30);
  if (budget > fr) return set_error(AUSM_ERR_NOMEM, "HBM budget %zu exceeds free memory %zu", budget, fr);
  ausm_batch* b = new ausm_batch();
  b->device = device;
  b->max_solves = max_solves;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    delete b;
    return set_error(AUSM_ERR_HIP, "device query failed");
  }
  b->cus = prop.multiProcessorCount;
  bool ok = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) == hipSuccess;
  ok = ok && hipMalloc(&b->arena_mem, budget) == hipSuccess;
  ok = ok && hipMalloc(&b->d_solves, sizeof(SolveDev) * max_solves) == hipSuccess;
  ok = ok && hipMalloc(&b->d_counts, sizeof(int) * 2) == hipSuccess;
  ok = ok && hipMalloc(&b->d_out, sizeof(double) * 2 * max_solves) == hipSuccess;
  ok = ok && hipHostMalloc(&b->h_out, sizeof(double) * 2 * max_solves) == hipSuccess;
  ok = ok && hipMalloc(&b->d_tmpd, sizeof(double) * 8) == hipSuccess;
  ok = ok && hipHostMalloc(&b->h_tmpd, sizeof(double) * 8) == hipSuccess;
  b->max_items = 4096;
  ok = ok && hipMalloc(&b->d_items, sizeof(BatchItem) * b->max_items) == hipSuccess;
  if (!ok) {
    ausm_batch_destroy(b);
    return set_error(AUSM_ERR_NOMEM, "batch allocation failed (budget %zu bytes)", budget);
  }
  b->arena.init(b->arena_mem, budget);
  b->use_graphs = getenv("AUSM_NO_GRAPHS") == nullptr;
  b->table.assign(max_solves, nullptr);
  b->h_solves.assign(max_solves, SolveDev{});
  *out = b;
  return AUSM_OK;
}

int ausm_batch_destroy(ausm_batch* b) {
  if (!b) return AUSM_OK;
  (void)hipSetDevice(b->device);
  if (b->stream) (void)hipStreamSynchronize(b->stream);
  drop_bgraphs(b);
  for (auto& kv : b->all) delete kv.second;
  (void)hipFree(b->arena_mem);
  (void)hipFree(b->d_solves);
  (void)hipFree(b->d_items);
  (void)hipFree(b->d_counts);
  (void)hipFree(b->d_out);
  (void)hipFree(b->d_tmp);
  (void)hipFree(b->d_tmpd);
  (void)hipFree(b->d_stage);
  (void)hipFree(b->d_seg);
  if (b->h_out) (void)hipHostFree(b->h_out);
  if (b->h_tmpd) (void)hipHostFree(b->h_tmpd);
  if (b->stream) (void)hipStreamDestroy(b->stream);
  delete b;
  return AUSM_OK;
}

int ausm_batch_stream(ausm_batch* b, void** stream) {
  if (!b || !stream) return set_error(AUSM_ERR_ARG, "NULL argument");
  *stream = (void*)b->stream;
  return AUSM_OK;
}

int ausm_batch_set_math(ausm_batch* b, int math) {
  if (!b) return set_error(AUSM_ERR_ARG, "NULL batch");
  if (math != AUSM_MATH_EXACT && math != AUSM_MATH_FAST) return set_error(AUSM_ERR_ARG, "unknown math mode");
  if (math != b->math) {
    b->math = math;
    b->tuned.clear();
  }
  return AUSM_OK;
}

int ausm_batch_set_precision(ausm_batch* b, int precision) {
  if (!b) return set_error(AUSM_ERR_ARG, "NULL batch");
  if (precision != AUSM_PREC_FP32 && precision != AUSM_PREC_FP32C)
    return set_error(AUSM_ERR_ARG, "batches run FP32 or FP32C precision");
  const bool comp = precision == AUSM_PREC_FP32C;
  if (comp == b->comp) return AUSM_OK;
  if (!b->queue.empty() || n_running(b) > 0)
    return set_error(AUSM_ERR_ORDER, "set the batch precision before submitting solves");
  b->comp = comp;
  b->tuned.clear();
  return AUSM_OK;
}

int ausm_batch_set_schedule(ausm_batch* b, int rows_per_wave, int blocks_per_cu) {
  if (!b) return set_error(AUSM_ERR_ARG, "NULL batch");
  if (rows_per_wave != 0 && rows_per_wave != 1 && rows_per_wave != 2 && rows_per_wave != 4)
    return set_error(AUSM_ERR_ARG, "rows per wavefront must be 0 (auto), 1, 2 or 4");
  if (blocks_per_cu < 0 || blocks_per_cu > 16) return set_error(AUSM_ERR_ARG, "workgroups per CU must be 0..16");
  b->fixed_ty = rows_per_wave;
  b->fixed_bpc = blocks_per_cu;
  b->dirty = true;
  return AUSM_OK;
}

int ausm_batch_submit(ausm_batch* b, const ausm_solve_desc* d, uint64_t* id) {
  if (!b || !d || !id) return set_error(AUSM_ERR_ARG, "NULL argument");
  if (d->M < 4 || d->N < 4 || d->M > (1 << 15) || d->N > (1 << 15) || !d->vertices)
    return set_error(AUSM_ERR_ARG, "bad grid in solve descriptor");
  if (d->scheme < 0 || d->scheme > 2 || d->limiter < 0 || d->limiter > 4)
    return set_error(AUSM_ERR_ARG, "bad scheme or limiter");
  if (!(d->rho > 0) || !(d->p > 0) || !(d->gamma > 1) || !(d->cfl > 0) || d->max_steps < 1)
    return set_error(AUSM_ERR_ARG, "bad flow or run settings");
  if (d->init_mode != AUSM_INIT_REST && d->init_mode != AUSM_INIT_FREESTREAM)
    return set_error(AUSM_ERR_ARG, "bad init mode");
  if (d->dt_mode != AUSM_DT_GLOBAL && d->dt_mode != AUSM_DT_LOCAL)
    return set_error(AUSM_ERR_ARG, "bad dt mode");
  if (d->freeze_after < 0 ||
      (d->freeze_after > 0 && d->limiter != LIM_VANALBADA && d->limiter != LIM_VANALBADA_CLAMPED))
    return set_error(AUSM_ERR_ARG, "limiter freezing needs a van Albada limiter");
  const Layout L = layout_of(d->M, d->N, b->comp, d->freeze_after > 0);
  if (L.total > b->arena.size())
    return set_error(AUSM_ERR_FULL, "a %d x %d solve needs %zu bytes, more than the whole budget", d->M, d->N, L.total);
  Slot* s = new Slot();
  s->id = b->next_id++;
  s->d = *d;
  s->verts.assign(d->vertices, d->vertices + (size_t)(d->N + 1) * d->M * 2);
  s->d.vertices = nullptr;
  b->all[s->id] = s;
  b->queue.push_back(s);
  *id = s->id;
  return AUSM_OK;
}

int ausm_batch_advance(ausm_batch* b, int nsteps, int check_every, int* pending) {
  if (!b) return set_error(AUSM_ERR_ARG, "NULL batch");
  if (nsteps < 0 || check_every < 1) return set_error(AUSM_ERR_ARG, "bad step counts");
  (void)hipSetDevice(b->device);
  int r = admit(b);
  if (r != AUSM_OK) return r;
  int left = nsteps;
  while (left > 0 && n_running(b) > 0) {
    if (b->dirty) {
      r = choose_schedule(b);
      if (r != AUSM_OK) return r;
      b->dirty = false;
    }
    r = run_chunk(b, check_every);
    if (r != AUSM_OK) return r;
    left -= check_every;
    if (launch_batch_norm(b->d_solves, b->d_counts + 1, (int)b->table.size(), b->d_out, b->stream) != hipSuccess ||
        hipMemcpyAsync(b->h_out, b->d_out, sizeof(double) * 2 * b->table.size(), hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
        hipStreamSynchronize(b->stream) != hipSuccess)
      return set_error(AUSM_ERR_HIP, "batch step failed: %s", hipGetErrorString(hipGetLastError()));
    for (int i = 0; i < (int)b->table.size(); ++i) {
      Slot* s = b->table[i];
      if (!s) continue;
      s->steps += check_every;
      s->res = b->h_out[2 * i];
      const double ws = b->h_out[2 * i + 1];
      int st = 0;
      if (!std::isfinite(s->res) || !std::isfinite(ws)) st = 4;
      else if (s->d.res_drop > 0 && s->res <= s->d.res_drop * s->res0) st = 2;
      else if (s->steps >= s->d.max_steps) st = 3;
      if (st) {
        r = retire(b, s, st);
        if (r != AUSM_OK) return r;
      } else if (s->d.freeze_after > 0 && !s->frozen && s->steps >= s->d.freeze_after) {
        r = freeze_solve(b, s);
        if (r != AUSM_OK) return r;
      }
    }
    r = admit(b);
    if (r != AUSM_OK) return r;
  }
  if (pending) *pending = n_running(b) + (int)b->queue.size();
  return AUSM_OK;
}

int ausm_batch_poll(ausm_batch* b, uint64_t id, ausm_solve_status* st) {
  if (!b || !st) return set_error(AUSM_ERR_ARG, "NULL argument");
  auto it = b->all.find(id);
  if (it == b->all.end()) return set_error(AUSM_ERR_NOTFOUND, "unknown solve id %llu", (unsigned long long)id);
  const Slot* s = it->second;
  st->state = s->state;
  st->steps = s->steps;
  st->res_l2 = s->res;
  st->res_l2_0 = s->res0;
  st->cl = s->cl;
  st->cd = s->cd;
  return AUSM_OK;
}

int ausm_batch_get_state(ausm_batch* b, uint64_t id, float* q) {
  if (!b || !q) return set_error(AUSM_ERR_ARG, "NULL argument");
  auto it = b->all.find(id);
  if (it == b->all.end()) return set_error(AUSM_ERR_NOTFOUND, "unknown solve id %llu", (unsigned long long)id);
  Slot* s = it->second;
  if (s->state == 0 || !s->bytes) return set_error(AUSM_ERR_ORDER, "solve %llu has no state on the device", (unsigned long long)id);
  (void)hipSetDevice(b->device);
  const size_t n = 4 * (size_t)s->d.M * (s->d.N + 3);
  if (n > b->stage_cap) {
    (void)hipFree(b->d_stage);
    b->stage_cap = n;
    if (hipMalloc(&b->d_stage, sizeof(float) * n) != hipSuccess) return set_error(AUSM_ERR_NOMEM, "staging allocation failed");
  }
  if (launch_soa_to_aos(s->dev.q[0], s->dev.vx, s->dev.vy, s->dpar, s->d.M, s->d.N, s->dev.P, s->dev.PS, b->d_stage, b->stream) != hipSuccess ||
      hipMemcpyAsync(q, b->d_stage, sizeof(float) * n, hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
      hipStreamSynchronize(b->stream) != hipSuccess)
    return set_error(AUSM_ERR_HIP, "state download failed");
  return AUSM_OK;
}

int ausm_batch_release(ausm_batch* b, uint64_t id) {
  if (!b) return set_error(AUSM_ERR_ARG, "NULL batch");
  auto it = b->all.find(id);
  if (it == b->all.end()) return set_error(AUSM_ERR_NOTFOUND, "unknown solve id %llu", (unsigned long long)id);
  Slot* s = it->second;
  (void)hipSetDevice(b->device);
  (void)hipStreamSynchronize(b->stream);
  auto qi = std::find(b->queue.begin(), b->queue.end(), s);
  if (qi != b->queue.end()) b->queue.erase(qi);
  free_slot(b, s);
  b->all.erase(it);
  delete s;
  return AUSM_OK;
}

}  // extern "C"
