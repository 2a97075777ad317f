// Host orchestration of the triangulation loop and the C ABI (include/wtet.h).
//
// One context owns a HIP stream and device buffers that persist across calls
// and only grow.  A call runs: normalize + Morton sort on the device, then
// insertion rounds (vote, pick, split, point split, flip iterations,
// relocation) until every point is inserted, then flips until none is
// possible, an exact audit of every face, compaction of the live tets, and
// -- when the audit found non-Delaunay faces -- the host repair.
// Reference loop: src/phase1.rs:16-1431; façade: src/lib.rs:160-310.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <functional>
#include <iterator>
#include <atomic>
#include <mutex>
#include <thread>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <sys/mman.h>
#include <string>
#include <vector>

#include "../../include/wtet.h"
#include "kernels.h"
#include "mlp.h"
#include "repair.h"
#include "sort.h"

using namespace wtet;

namespace {

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

}  // namespace

struct wtet_stage_session;

struct wtet_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  wtet_stats stats = {};
  std::vector<DevBuf*> all;

  // point preparation
  DevBuf xyz_in, xyzn, keys, ids, keys2, ids2, sort_tmp, mm6, pts, doff;
  // mesh (per tet slot)
  DevBuf tet, adj_t, tinfo, tfwd, tvote, fvote, tsplit, tpoint, tmsg;
  // lists
  DevBuf uvert[2], utet[2], wlist, act[2], prop, fout, blk, slot2idx, rec;
  // output grouping for many clouds
  DevBuf ord[2], ord_tmp, cstart;
  // learned presizing
  DevBuf presize_params, feat_acc, feat, pred;
  bool presize_loaded = false;
  std::vector<float> h_pred;
  // device-driven flip iterations: sums (checked tets, flips, iterations),
  // persistent grid, and the captured graph of one round's iterations
  DevBuf isums, slow_t, slow_p;
  // neighbourhood of the host repair (sparse download)
  DevBuf nbr_mark, nbr_list, nbr_misc, nbr_pack, nbr_seed;
  uint32_t items = 2;  // list entries per thread the persistent grids are sized for
  uint32_t audit_records = 1u << 20;  // violation records the first audit of a call keeps
  bool use_graphs = true;
  struct GraphEntry {
    std::vector<uint64_t> key;
    hipGraphExec_t exec;
  };
  std::vector<GraphEntry> graphs;        // flip iterations: one per (buffers, grid, iterations)
  std::vector<GraphEntry> round_graphs;  // rounds: one per (buffers, grid buckets, variant)
  // keys launched directly once and not captured yet (a graph is captured
  // the second time its key comes up: a one-off call does not pay for
  // captures it never replays)
  std::vector<std::vector<uint64_t>> graph_keys_seen;
  uint32_t graph_captures = 0;
  // stage-level session (wtet_stage_*), null when none
  wtet_stage_session* stage = nullptr;
  // counters
  DevBuf ctr, gstats, patch;
  uint32_t* h_ctr = nullptr;  // pinned mirror of the counters
  // pinned staging for the mesh when the host needs it (repair / host result)
  uint32_t* h_tets = nullptr;
  uint32_t* h_adj = nullptr;
  float* h_xyzn = nullptr;
  // the normalized points go to h_xyzn on copy_stream while the device
  // phase runs (the host repair reads them); ev_xyzn marks the copy's end
  hipStream_t copy_stream = nullptr;
  hipEvent_t ev_norm = nullptr, ev_xyzn = nullptr;
  uint32_t xyzn_copied = 0;  // points of the copy in flight or done (0: none)
  // per-cloud rows of the last resident batch result
  std::vector<uint64_t> bstart;
  std::vector<uint32_t> bcount;
  uint32_t* h_patch = nullptr;
  size_t h_cap = 0, h_npts = 0, h_patch_cap = 0;
  // allocates h_tets/h_adj while a single cloud's device phase runs (the
  // first call at a new size: pinning and first-touching hundreds of MB);
  // joined before anything reads them
  std::thread pin_thread;
  uint64_t cap_tets = 0;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  // per-kernel timing with HIP events on the context's stream
  // (src/profiler.rs GpuProfiler scopes, here actually recorded)
  bool profiling = false;
  int ring = 1;  // insertion independence radius (face-steps), see is_winner
  // the first rounds insert one point per tet (ring 0); kAutoWarmup: from
  // the largest cloud (warmup_rounds)
  uint32_t ring0_rounds = 0xFFFFFFFFu;
  int repair_rings = 7;       // face-steps of rows staged around each stuck face
  // rows of the host mirror that hold staged rows (sparse repair): those
  // with present[r] == present_mark.  Each call takes the next mark instead
  // of clearing the rows the previous one staged; the array is cleared once
  // every 255 calls, when the mark wraps.
  std::vector<uint8_t> present;
  uint8_t present_mark = 0;
  // a single-cloud repair region above this many tets stops the repair and
  // reruns the device phase with a wider insertion ring (0: never)
  uint32_t region_cap = 1u << 14;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> evpool;
  std::vector<int> evkind;
  size_t evused = 0;
  double ktime[16] = {};
  uint32_t kcount[16] = {};

  wtet_ctx() {
    DevBuf* l[] = {&xyz_in, &xyzn,   &keys,   &ids,   &keys2,  &ids2,     &sort_tmp, &mm6,
                   &pts,    &doff,   &tet,    &adj_t,  &tinfo,  &tfwd,  &tvote,    &fvote,    &tsplit,  &tpoint,
                   &tmsg,   &uvert[0], &uvert[1], &utet[0], &utet[1],
                   &wlist,  &act[0], &act[1], &prop,  &fout,   &blk,      &slot2idx, &rec,
                   &ord[0], &ord[1], &ord_tmp, &cstart, &ctr,  &gstats,   &patch,
                   &presize_params, &feat_acc, &feat, &pred, &isums, &slow_t, &slow_p,
                   &nbr_mark, &nbr_list, &nbr_misc, &nbr_pack, &nbr_seed};
    all.assign(std::begin(l), std::end(l));
  }
};

namespace {

struct HipError {
  hipError_t e;
  const char* what;
};

#define HIP_CHECK(x)                          \
  do {                                        \
    hipError_t e_ = (x);                      \
    if (e_ != hipSuccess) throw HipError{e_, #x}; \
  } while (0)

struct Fail {
  wtet_status st;
  std::string msg;
};

void ensure(wtet_ctx* c, DevBuf& b, size_t bytes, bool keep = false) {
  if (b.bytes >= bytes && b.p) return;
  size_t nb = std::max(bytes, b.bytes + b.bytes / 2);
  void* p = nullptr;
  if (hipMalloc(&p, nb) != hipSuccess) {
    (void)hipGetLastError();
    throw Fail{WTET_ERR_OUT_OF_MEMORY, "device allocation of " + std::to_string(nb) + " bytes failed"};
  }
  if (keep && b.p && b.bytes) HIP_CHECK(hipMemcpyAsync(p, b.p, b.bytes, hipMemcpyDeviceToDevice, c->stream));
  if (b.p) {
    HIP_CHECK(hipStreamSynchronize(c->stream));
    HIP_CHECK(hipFree(b.p));
  }
  b.p = p;
  b.bytes = nb;
}

template <typename T>
T* P(DevBuf& b) {
  return reinterpret_cast<T*>(b.p);
}

void prefault(void* p, size_t bytes) {
  const size_t kChunk = 64u << 20;
  unsigned nt = std::min<unsigned>(8u, (unsigned)((bytes + kChunk - 1) / kChunk));
  if (nt <= 1) {
    std::memset(p, 0, bytes);
    return;
  }
  std::vector<std::thread> th;
  size_t per = (bytes / nt + 4095) & ~(size_t)4095;
  for (unsigned i = 0; i < nt; ++i) {
    size_t a = std::min(bytes, per * i), b = std::min(bytes, per * (i + 1));
    th.emplace_back([p, a, b] { std::memset(static_cast<char*>(p) + a, 0, b - a); });
  }
  for (auto& t : th) t.join();
}

// A result array of `bytes` for the caller (released with free()): 2 MiB
// aligned with huge pages advised when large, so that its first touch
// takes a fault per 2 MiB instead of per 4 KiB
uint32_t* alloc_rows(size_t bytes) {
  const size_t kHuge = (size_t)2 << 20;
  if (bytes < kHuge) return static_cast<uint32_t*>(std::malloc(std::max<size_t>(bytes, 4)));
  size_t nb = (bytes + kHuge - 1) / kHuge * kHuge;
  void* p = std::aligned_alloc(kHuge, nb);
  if (p) (void)madvise(p, nb, MADV_HUGEPAGE);
  return static_cast<uint32_t*>(p);
}

// memcpy on up to 8 threads (64 MiB or more each): a fresh destination's
// first-touch page faults are then taken in parallel
void parallel_copy(void* dst, const void* src, size_t bytes) {
  const size_t kChunk = 64u << 20;
  unsigned nt = std::min<unsigned>(8u, (unsigned)((bytes + kChunk - 1) / kChunk));
  if (nt <= 1) {
    if (bytes) std::memcpy(dst, src, bytes);
    return;
  }
  std::vector<std::thread> th;
  size_t per = (bytes / nt + 4095) & ~(size_t)4095;
  for (unsigned i = 0; i < nt; ++i) {
    size_t a = std::min(bytes, per * i), b = std::min(bytes, per * (i + 1));
    th.emplace_back([dst, src, a, b] {
      std::memcpy(static_cast<char*>(dst) + a, static_cast<const char*>(src) + a, b - a);
    });
  }
  for (auto& t : th) t.join();
}

template <typename T>
void ensure_pinned(T*& p, size_t& cap, size_t need) {
  if (p && cap >= need) return;
  if (p) HIP_CHECK(hipHostFree(p));
  p = nullptr;
  // slack: the rows of one input vary by a few hundred between calls (a
  // non-Delaunay mesh of the same points has a different tet count), and a
  // reallocation pre-faults hundreds of megabytes again
  size_t nc = std::max(need + need / 8, cap + cap / 2);
  if (hipHostMalloc((void**)&p, nc * sizeof(T), hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    p = nullptr;
    cap = 0;
    throw Fail{WTET_ERR_OUT_OF_MEMORY, "pinned host allocation failed"};
  }
  cap = nc;
  // touch every page once: the host maps pinned pages on first write, and
  // the repair's scattered row writes would otherwise fault page by page.
  // Large buffers are touched by several threads (page faults run in
  // parallel).
  prefault(p, nc * sizeof(T));
}

// wait for the mirror allocation pin_ahead started, if any
void pin_join(wtet_ctx* c) {
  if (c->pin_thread.joinable()) c->pin_thread.join();
}

// Start allocating the host side of the repair for a mesh of about `rows`
// rows on a thread of its own, so that it overlaps the device phase: the
// mirror (h_tets, h_adj), the row flags (present) and the staging of
// scattered rows (h_patch, room for a sixteenth of the rows).  A failure is
// left for the synchronous allocation at staging to report.
void pin_ahead(wtet_ctx* c, size_t rows) {
  pin_join(c);
  const size_t need = rows * 4, patch = (rows / 16) * 9 + 9;
  if (c->h_tets && c->h_adj && c->h_cap >= need && c->h_patch && c->h_patch_cap >= patch &&
      c->present.size() >= rows)
    return;
  const int dev = c->device;
  c->pin_thread = std::thread([c, rows, need, patch, dev] {
    try {
      if (hipSetDevice(dev) != hipSuccess) return;
      size_t c1 = c->h_cap, c2 = c->h_cap;
      ensure_pinned(c->h_tets, c1, need);
      ensure_pinned(c->h_adj, c2, need);
      c->h_cap = std::min(c1, c2);
      if (c->present.size() < rows) c->present.assign(rows, 0);
      ensure_pinned(c->h_patch, c->h_patch_cap, patch);
    } catch (...) {
      (void)hipGetLastError();
    }
  });
}

// per-tet-slot arrays sized for `cap` slots; keep=true preserves contents
void ensure_tets(wtet_ctx* c, uint64_t cap, bool keep) {
  if (cap <= c->cap_tets) return;
  if (cap >= (1ull << 30)) throw Fail{WTET_ERR_OUT_OF_MEMORY, "more than 2^30 tet slots required"};
  ensure(c, c->tet, cap * 16, keep);
  ensure(c, c->adj_t, cap * sizeof(AdjRec), keep);
  ensure(c, c->tinfo, cap * 4, keep);
  ensure(c, c->tfwd, cap * 4, keep);
  ensure(c, c->tvote, cap * 8, keep);
  ensure(c, c->fvote, cap * 8, keep);
  ensure(c, c->tsplit, cap * 8, keep);
  ensure(c, c->tpoint, cap * 4, keep);
  ensure(c, c->tmsg, cap * sizeof(FlipMsg), keep);
  ensure(c, c->act[0], cap * 4, keep);
  ensure(c, c->act[1], cap * 4, keep);
  ensure(c, c->slow_t, cap * 4, false);
  ensure(c, c->prop, cap * 16, false);
  ensure(c, c->fout, cap * 32, false);  // two uint4 per flip; also the output adjacency
  ensure(c, c->slot2idx, cap * 4, false);
  ensure(c, c->blk, (cap / kBlock + 2) * 4, false);
  if (keep) c->stats.grow_events++;
  c->cap_tets = std::min<uint64_t>(c->tet.bytes / 16, (1ull << 30) - 1);
}

enum KernelKind : int {
  K_PREP = 0, K_VOTE, K_PICK, K_SPLIT, K_SPLIT_POINTS, K_FLIP_CHECK, K_FLIP_CHECK_DEFERRED,
  K_FLIP_EXEC, K_FLIP_UPDATE, K_RELOCATE, K_AUDIT, K_OUTPUT, K_FLIP_SMALL, K_NKINDS
};
const char* kKernelNames[K_NKINDS] = {"prepare_points", "vote", "pick", "split", "split_points",
                                      "flip_check", "flip_check_deferred", "flip_decide_exec", "flip_update",
                                      "relocate", "audit", "output", "flip_small"};

// brackets the launches issued by fn with an event pair when profiling
template <typename Fn>
void timed(wtet_ctx* c, int kind, Fn fn) {
  if (!c->profiling) {
    fn();
    return;
  }
  if (c->evused == c->evpool.size()) {
    hipEvent_t a, b;
    HIP_CHECK(hipEventCreate(&a));
    HIP_CHECK(hipEventCreate(&b));
    c->evpool.emplace_back(a, b);
    c->evkind.push_back(0);
  }
  auto& pr = c->evpool[c->evused];
  c->evkind[c->evused] = kind;
  ++c->evused;
  HIP_CHECK(hipEventRecord(pr.first, c->stream));
  fn();
  HIP_CHECK(hipEventRecord(pr.second, c->stream));
}

void collect_times(wtet_ctx* c) {
  if (!c->profiling) return;
  HIP_CHECK(hipStreamSynchronize(c->stream));
  for (size_t i = 0; i < c->evused; ++i) {
    float ms = 0;
    HIP_CHECK(hipEventElapsedTime(&ms, c->evpool[i].first, c->evpool[i].second));
    c->ktime[c->evkind[i]] += ms;
    c->kcount[c->evkind[i]] += 1;
  }
  c->evused = 0;
}

Mesh mesh_of(wtet_ctx* c, uint32_t inf, int mode) {
  Mesh M;
  M.pts = P<float4>(c->pts);
  M.tet = P<uint4>(c->tet);
  M.adj = P<AdjRec>(c->adj_t);
  M.tinfo = P<uint32_t>(c->tinfo);
  M.tfwd = P<uint32_t>(c->tfwd);
  M.tvote = P<uint64_t>(c->tvote);
  M.fvote = P<uint64_t>(c->fvote);
  M.tsplit = P<uint2>(c->tsplit);
  M.tpoint = P<uint32_t>(c->tpoint);
  M.tmsg = P<FlipMsg>(c->tmsg);
  M.ctr = P<uint32_t>(c->ctr);
  M.inf = inf;
  M.mode = mode;
  return M;
}

void read_ctr(wtet_ctx* c) {
  HIP_CHECK(hipMemcpyAsync(c->h_ctr, c->ctr.p, C_COUNT * 4, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
}

struct Counters {
  uint64_t sumU = 0, sumA = 0, sumF = 0, peakA = 0;
  uint32_t rounds = 0, iters = 0;
  uint32_t left = 0;  // points not inserted when max_insert_iterations stopped the rounds
  uint64_t flips23 = 0;               // 2-3 flips of sumF
  uint64_t smallA = 0, smallF = 0;    // of sumA, sumF: in one-workgroup iterations
};

void collect_times(wtet_ctx* c);

FlipLists lists_of(wtet_ctx* c) {
  FlipLists L;
  L.act[0] = P<uint32_t>(c->act[0]);
  L.act[1] = P<uint32_t>(c->act[1]);
  L.prop = P<uint4>(c->prop);
  L.out = P<uint4>(c->fout);
  L.slow = P<uint32_t>(c->slow_t);
  return L;
}

void enqueue_flip_iterations(wtet_ctx* c, const Mesh& M, const FlipLists& L, uint64_t* st, int k,
                             uint32_t grid, uint32_t dgrid) {
  hipStream_t s = c->stream;
  uint64_t* sums = P<uint64_t>(c->isums);
  launch_iter_begin(M.ctr, s);
  for (int i = 0; i < k; ++i) {
    timed(c, K_FLIP_CHECK, [&] { launch_flip_check(M, L, grid, st, s); });
    timed(c, K_FLIP_CHECK_DEFERRED, [&] { launch_flip_check_deferred(M, L, dgrid, st, s); });
    timed(c, K_FLIP_EXEC, [&] { launch_flip_decide_exec(M, L, grid, s); });
    timed(c, K_FLIP_UPDATE, [&] { launch_flip_update(M, L, grid, s); });
    if (i + 1 < k) launch_iter_step(M.ctr, sums, s);
    else launch_iter_end(M.ctr, sums, s);
  }
}

// persistent grid for a list of about `a` entries: a power of two of
// workgroups in [lo, 4096]
uint32_t grid_for(const wtet_ctx* c, uint64_t a, uint32_t lo) {
  uint64_t want = (a + (uint64_t)kBlock * c->items - 1) / ((uint64_t)kBlock * c->items);
  uint32_t g = lo;
  while (g < want && g < 4096) g <<= 1;
  return g;
}

// The deferred check's grid: four lanes per tet (one per face), for four
// times the deferred tets an iteration is expected to hold (the first
// iterations of a round hold more than the average, and each deferred tet
// keeps its lanes busy for microseconds), up to eight times the check's
// grid.  Degenerate input defers most tets: a 128^3 lattice's deferred
// checks took 188 ms at one thread per expected tet, 165 ms at four, 154 ms
// with four lanes per tet at four times the expected count.
uint32_t deferred_grid(const wtet_ctx* c, uint64_t dper, uint32_t grid) {
  return std::min(grid_for(c, 16 * dper, 1), std::min(grid * 8, 4096u));
}

// k flip iterations without the host: replays a captured graph (one per
// buffer set, grid and k; captured on first use), or launches directly when
// profiling (per-kernel events) or when graphs are off.
// a_hint: bound on the first list; d_hint: deferred tets per iteration
// expected (from the previous round).  The deferred check runs the full
// predicates, whose exact stage keeps a thread busy for microseconds, so its
// grid follows the deferred count rather than the list.
// Lists this short run in one workgroup (k_flip_small): launches, not
// work, set the time of their iterations.  The list may grow a few times
// within the batch; past kSmallLimit the rest is left to the next batch.
constexpr uint64_t kSmallList = 512;
constexpr uint32_t kSmallLimit = 2048;

// The device buffers a captured sequence reads: a graph of another buffer
// set is stale.
std::vector<uint64_t> buffer_key(wtet_ctx* c, const Mesh& M, const FlipLists& L, uint64_t* st) {
  return {(uint64_t)(uintptr_t)M.pts,    (uint64_t)(uintptr_t)M.tet,    (uint64_t)(uintptr_t)M.tvote,
          (uint64_t)(uintptr_t)M.adj,    (uint64_t)(uintptr_t)M.tinfo,  (uint64_t)(uintptr_t)M.tfwd,
          (uint64_t)(uintptr_t)M.fvote,  (uint64_t)(uintptr_t)M.tsplit, (uint64_t)(uintptr_t)M.tmsg,
          (uint64_t)(uintptr_t)M.tpoint, (uint64_t)(uintptr_t)M.ctr,
          (uint64_t)M.inf,               (uint64_t)M.mode,              (uint64_t)(uintptr_t)L.act[0],
          (uint64_t)(uintptr_t)L.act[1], (uint64_t)(uintptr_t)L.prop,   (uint64_t)(uintptr_t)L.out,
          (uint64_t)(uintptr_t)L.slow,   (uint64_t)(uintptr_t)st,       (uint64_t)(uintptr_t)c->isums.p,
          (uint64_t)(uintptr_t)c->uvert[0].p, (uint64_t)(uintptr_t)c->uvert[1].p,
          (uint64_t)(uintptr_t)c->utet[0].p,  (uint64_t)(uintptr_t)c->utet[1].p,
          (uint64_t)(uintptr_t)c->wlist.p,    (uint64_t)(uintptr_t)c->blk.p,
          (uint64_t)(uintptr_t)c->slow_p.p};
}

// Replays the graph of (buffers, params), capturing `enqueue` into it on
// first use.  Graphs of another buffer set in the same cache are destroyed
// first: every call ends with a stream synchronization, and a growth frees
// the old buffers (which synchronizes), so none of them is still running.
void launch_cached_graph(wtet_ctx* c, std::vector<wtet_ctx::GraphEntry>& cache,
                         const std::vector<uint64_t>& buffers, const std::vector<uint64_t>& params,
                         const std::function<void()>& enqueue) {
  std::vector<uint64_t> key = buffers;
  key.insert(key.end(), params.begin(), params.end());
  hipGraphExec_t exec = nullptr;
  for (auto& g : cache)
    if (g.key == key) exec = g.exec;
  if (!exec) {
    // first sight of the key: launch directly and remember it
    auto& seen = c->graph_keys_seen;
    auto same_buffers = [&](const std::vector<uint64_t>& k) {
      return k.size() >= buffers.size() && std::equal(buffers.begin(), buffers.end(), k.begin());
    };
    seen.erase(std::remove_if(seen.begin(), seen.end(), [&](const std::vector<uint64_t>& k) { return !same_buffers(k); }),
               seen.end());
    if (std::find(seen.begin(), seen.end(), key) == seen.end()) {
      seen.push_back(key);
      enqueue();
      return;
    }
    for (size_t i = 0; i < cache.size();) {
      const std::vector<uint64_t>& k = cache[i].key;
      if (k.size() < buffers.size() || !std::equal(buffers.begin(), buffers.end(), k.begin())) {
        HIP_CHECK(hipGraphExecDestroy(cache[i].exec));
        cache.erase(cache.begin() + (long)i);
      } else {
        ++i;
      }
    }
    hipGraph_t g = nullptr;
    HIP_CHECK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    enqueue();
    HIP_CHECK(hipStreamEndCapture(c->stream, &g));
    hipError_t e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIP_CHECK(e);
    cache.push_back({key, exec});
    c->graph_captures++;
  }
  HIP_CHECK(hipGraphLaunch(exec, c->stream));
}

void run_flip_batch(wtet_ctx* c, uint32_t inf, int mode, uint64_t* st, int k, uint64_t a_hint,
                    uint64_t d_hint) {
  Mesh M = mesh_of(c, inf, mode);
  FlipLists L = lists_of(c);
  if (a_hint <= kSmallList) {
    timed(c, K_FLIP_SMALL, [&] {
      launch_flip_small(M, L, st, P<uint64_t>(c->isums), k, kSmallLimit, c->stream);
    });
    HIP_CHECK(hipGetLastError());
    return;
  }
  uint32_t grid = grid_for(c, a_hint, 16);
  uint32_t dgrid = deferred_grid(c, d_hint, grid);
  if (!c->use_graphs || c->profiling) {
    if (c->evused > 4096) collect_times(c);
    enqueue_flip_iterations(c, M, L, st, k, grid, dgrid);
    HIP_CHECK(hipGetLastError());
    return;
  }
  launch_cached_graph(c, c->graphs, buffer_key(c, M, L, st), {(uint64_t)dgrid, (uint64_t)grid, (uint64_t)k},
                      [&] { enqueue_flip_iterations(c, M, L, st, k, grid, dgrid); });
}

// tet slots after growth by `need`; the flip iterations read the capacity
void grow_for(wtet_ctx* c, uint64_t need) {
  if (need > c->cap_tets) ensure_tets(c, need + need / 2, true);
  launch_set_ctr(P<uint32_t>(c->ctr), C_CAP, (uint32_t)c->cap_tets, C_STOP, 0u, c->stream);
}

// Morton bits per axis and the bit where the cloud index starts in a key
// Key layout, low to high: Morton code (kbits per axis), then in Hilbert
// mode the 24-bit Hilbert index, then the cloud index.
void key_layout(uint32_t C, bool hilbert, int* kbits, int* cshift, int* end_bit) {
  int cb = 0;
  while ((1ull << cb) < C) ++cb;
  int room = 63 - cb - (hilbert ? 24 : 0);
  *kbits = std::min(hilbert ? 13 : 21, room / 3);
  *cshift = 3 * *kbits + (hilbert ? 24 : 0);
  *end_bit = *cshift + cb;
}

void end_stage(wtet_ctx* c);

// State of one device phase over clouds off[c]..off[c+1] (one cloud: {0, n}):
// run_device drives it start to end; the stage entry points drive it one
// step at a time.
struct Run {
  wtet_config cfg{};
  std::vector<uint32_t> off;
  uint32_t C = 0, n = 0, V = 0;  // clouds, points, infinity vertex
  int mode = PRED_FILTERED;
  bool hilbert = false, learned = false, flipping = true;
  int kbits = 21, cshift = 63, end_bit = 63;
  Clouds CL{};
  uint64_t* st = nullptr;  // predicate stage counters (collect_stats)
  uint32_t cur = 0;        // uninserted list buffer in use
  uint32_t round = 0;
  uint32_t U = 0;          // uninserted points
  uint32_t W = 0;          // insertions of the current round
  uint64_t dsum_prev = 0;  // deferred-check total at the previous round
  int max_fi = 10;
  uint32_t warmup = 0;     // rounds that insert one point per tet
  Counters K;
};

// Rounds that insert one point per tet before the ring applies.  While a
// cloud is small a round costs about the same whatever it inserts, and ring
// 0 inserts several times as many points as ring 1 there; later, adjacent
// insertions leave faces the flips cannot fix.  Measured best: 6 rounds for
// clouds of 1M-2M points, 4 for clouds of at most 50,000 (about log4(n) - 4).
uint32_t warmup_rounds(const std::vector<uint32_t>& off) {
  uint32_t big = 0;
  for (size_t k = 0; k + 1 < off.size(); ++k) big = std::max(big, off[k + 1] - off[k]);
  int l4 = (int)std::lround(std::log((double)std::max(big, 1u)) / std::log(4.0));
  return (uint32_t)std::min(6, std::max(0, l4 - 4));
}

void run_setup(wtet_ctx* c, Run& R, const std::vector<uint32_t>& off, const wtet_config& cfg) {
  R = Run{};
  R.cfg = cfg;
  R.off = off;
  R.mode = cfg.predicate_mode == 1 ? PRED_EXACT_ONLY : PRED_FILTERED;
  R.C = (uint32_t)off.size() - 1;
  R.n = off[R.C];
  R.V = R.n + 4 * R.C;
  R.hilbert = cfg.enable_hilbert_sorting != 0;
  R.learned = cfg.presize_mode == 1;
  R.flipping = cfg.enable_flipping != 0;
  R.max_fi = (int)std::min<uint32_t>(std::max<uint32_t>(cfg.max_flip_iterations, 1u), 64u);
  R.warmup = c->ring0_rounds == 0xFFFFFFFFu ? warmup_rounds(off) : c->ring0_rounds;
  key_layout(R.C, R.hilbert, &R.kbits, &R.cshift, &R.end_bit);
  R.st = cfg.collect_stats ? P<uint64_t>(c->gstats) : nullptr;
}

// Buffers, normalization, keys, sort, vertex layout, uninserted list,
// presizing and the start tets (src/gpu/mod.rs:48-204, src/lib.rs:313-337,
// init.wgsl:87-148).
void dev_prepare(wtet_ctx* c, Run& R, const float* d_xyz) {
  hipStream_t s = c->stream;
  const uint32_t C = R.C, n = R.n, V = R.V;
  const std::vector<uint32_t>& off = R.off;
  // the previous call's copy of the normalized points may still read xyzn
  // (a call that needed no repair never waits for it) and write h_xyzn:
  // both may be replaced below
  HIP_CHECK(hipStreamSynchronize(c->copy_stream));
  c->xyzn_copied = 0;
  ensure(c, c->xyzn, (size_t)n * 12);
  ensure(c, c->keys, (size_t)n * 8);
  ensure(c, c->ids, (size_t)n * 4);
  ensure(c, c->keys2, (size_t)n * 8);
  ensure(c, c->ids2, (size_t)n * 4);
  ensure(c, c->sort_tmp, std::max<size_t>(std::max(sort_temp_bytes(n, R.end_bit), sort_temp_bytes(n, 64)), 16));
  ensure(c, c->mm6, (size_t)C * 24);
  ensure(c, c->doff, ((size_t)C + 1) * 4);
  ensure(c, c->cstart, ((size_t)C + 1) * 4);
  ensure(c, c->pts, ((size_t)V + 1) * 16);
  for (int b = 0; b < 2; ++b) {
    ensure(c, c->uvert[b], (size_t)n * 4);
    ensure(c, c->utet[b], (size_t)n * 4);
  }
  ensure(c, c->wlist, (size_t)n * 8);
  ensure(c, c->slow_p, (size_t)n * 4 + 16);
  ensure(c, c->ctr, C_COUNT * 4);
  ensure(c, c->gstats, 64);
  ensure(c, c->isums, S_COUNT * 8);
  ensure(c, c->rec, (size_t)c->audit_records * 4);
  ensure(c, c->blk, ((size_t)round_grid(n) + 2) * 4);
  if (R.learned) {
    if (!c->presize_loaded) throw Fail{WTET_ERR_INVALID_ARGUMENT, "presize_mode 1 needs a model (wtet_set_presize_model)"};
    ensure(c, c->feat_acc, (size_t)C * (kHistBins / 2 + 9) * 8 + 64);
    ensure(c, c->feat, (size_t)C * kFeat * 4);
    ensure(c, c->pred, (size_t)C * kOut * 4);
    c->h_pred.resize((size_t)C * kOut);
  }
  R.st = R.cfg.collect_stats ? P<uint64_t>(c->gstats) : nullptr;
  ensure_pinned(c->h_xyzn, c->h_npts, (size_t)n * 3);

  HIP_CHECK(hipMemcpyAsync(c->doff.p, off.data(), ((size_t)C + 1) * 4, hipMemcpyHostToDevice, s));
  Clouds& CL = R.CL;
  CL.off = P<uint32_t>(c->doff);
  CL.C = C;
  CL.V = V;
  CL.kbits = R.kbits;
  CL.cshift = R.cshift;
  CL.hilbert = R.hilbert ? 1 : 0;

  HIP_CHECK(hipEventRecord(c->ev0, s));
  launch_fill_u32(P<uint32_t>(c->ctr), 0u, C_COUNT, s);
  launch_fill_u64(P<uint64_t>(c->gstats), 0ull, 8, s);
  timed(c, K_PREP, [&] {
    // normalize each cloud (f32, as the reference), (cloud, Morton) keys,
    // sort, gather into the vertex layout
    launch_prepare(d_xyz, n, CL, P<uint32_t>(c->mm6), P<float>(c->xyzn), P<uint64_t>(c->keys),
                   P<uint32_t>(c->ids), s);
    // the normalized points are final: their copy to the host overlaps
    // the rest of the device phase
    HIP_CHECK(hipEventRecord(c->ev_norm, s));
    HIP_CHECK(hipStreamWaitEvent(c->copy_stream, c->ev_norm, 0));
    HIP_CHECK(hipMemcpyAsync(c->h_xyzn, c->xyzn.p, (size_t)n * 12, hipMemcpyDeviceToHost, c->copy_stream));
    HIP_CHECK(hipEventRecord(c->ev_xyzn, c->copy_stream));
    c->xyzn_copied = n;
    HIP_CHECK(sort_pairs(P<uint64_t>(c->keys), P<uint32_t>(c->ids), P<uint64_t>(c->keys2),
                         P<uint32_t>(c->ids2), n, R.end_bit, c->sort_tmp.p, c->sort_tmp.bytes, s));
    launch_gather_points(P<float>(c->xyzn), P<uint32_t>(c->ids2), P<uint64_t>(c->keys2), n, CL,
                         P<float4>(c->pts), s);
    // uninserted list: every point (minus redundant duplicates) in its
    // cloud's super tet
    launch_init_uninserted(P<float4>(c->pts), P<uint64_t>(c->keys2), n, CL, P<uint32_t>(c->blk),
                           P<uint32_t>(c->uvert[0]), P<uint32_t>(c->utet[0]), P<uint32_t>(c->ctr),
                           s);
    if (R.learned) {
      // features from the raw points and this call's per-cloud boxes, then
      // the network's predictions (log space, margin included)
      launch_features(d_xyz, n, CL.off, C, P<uint32_t>(c->mm6), P<double>(c->feat_acc),
                      P<float>(c->feat), s);
      launch_mlp_predict(P<float>(c->presize_params), P<float>(c->feat), C, P<float>(c->pred), s);
      HIP_CHECK(hipMemcpyAsync(c->h_pred.data(), c->pred.p, (size_t)C * kOut * 4,
                               hipMemcpyDeviceToHost, s));
    }
  });
  HIP_CHECK(hipGetLastError());
  read_ctr(c);
  R.U = c->h_ctr[C_NPTS];
  c->stats.duplicates = n - R.U;
  if (R.hilbert) {
    timed(c, K_PREP, [&] {
      HIP_CHECK(launch_stratify(P<uint32_t>(c->uvert[0]), P<uint32_t>(c->utet[0]), R.U, CL,
                                P<uint64_t>(c->keys), P<uint64_t>(c->keys2), P<uint32_t>(c->ids2),
                                c->sort_tmp.p, c->sort_tmp.bytes, s));
    });
    HIP_CHECK(hipGetLastError());
  }
  // tet slots for the whole call: the reference's fixed formula, the
  // learned prediction, or a small start that the grow-and-retry path extends
  uint64_t slots;
  if (R.learned) {
    double sum = 0.0;
    for (uint32_t cl = 0; cl < C; ++cl) {
      double lp = std::min<double>(c->h_pred[(size_t)cl * kOut], 4.0);  // <= e^4 slots per point
      sum += (double)(off[cl + 1] - off[cl]) * std::exp(lp);
    }
    for (int k = 0; k < 3; ++k) c->stats.presize_pred[k] = C == 1 ? c->h_pred[k] : 0.0f;
    slots = (uint64_t)std::ceil(sum) + 5ull * C + 1024;
  } else if (R.cfg.presize_mode == 2) {
    slots = 2ull * n + 4096 + 5ull * C;
  } else {
    // the reference's fixed formula per cloud (src/gpu/mod.rs:101-108)
    slots = 0;
    for (uint32_t cl = 0; cl < C; ++cl) {
      uint64_t nc = off[cl + 1] - off[cl];
      slots += std::max<uint64_t>(std::max<uint64_t>((nc + 5) * 8 + 8, 12 * nc), 64);
    }
  }
  c->stats.presize_slots = slots;
  ensure_tets(c, slots, false);
  c->stats.tet_capacity = c->cap_tets;
  launch_start_tets(mesh_of(c, V, R.mode), CL, s);
  launch_set_ctr(P<uint32_t>(c->ctr), C_TETS, 5u * C, -1, 0u, s);
  HIP_CHECK(hipGetLastError());
  launch_fill_u64(P<uint64_t>(c->isums), 0ull, S_COUNT, s);
  launch_set_ctr(P<uint32_t>(c->ctr), C_CAP, (uint32_t)c->cap_tets, -1, 0u, s);
}

// After a round's pick, on the host: the one read of the round (winners,
// reserved slots, and whether the previous round's flips stopped for
// capacity), its checks and the growth of the tet slots.  U: the round's
// uninserted count before the pick.
void after_pick(wtet_ctx* c, Run& R, uint32_t U) {
  read_ctr(c);
  uint32_t W = c->h_ctr[C_WIN];
  if (c->h_ctr[C_BADTET]) throw Fail{WTET_ERR_INTERNAL, "a point was located outside the super tet"};
  if (W == 0) throw Fail{WTET_ERR_INTERNAL, "no insertion possible in a round"};
  uint64_t need = (uint64_t)c->h_ctr[C_TETS] + 64 + (c->h_ctr[C_STOP] ? c->h_ctr[C_A] : 0u);
  if (need > c->cap_tets || c->h_ctr[C_STOP]) grow_for(c, need);
  R.K.sumU += U;
  R.K.rounds++;
  R.W = W;
  R.U = U - W;
  R.cur ^= 1;
}

// One insertion round's vote, winner pick and compaction of the survivors
// (vote.wgsl:89-344, compact_vertex_arrays_*.wgsl), launched one by one:
// returns the round's insertions; the survivors become the uninserted list.
uint32_t dev_vote(wtet_ctx* c, Run& R) {
  hipStream_t s = c->stream;
  ++R.round;
  if (R.round > (1u << 16)) throw Fail{WTET_ERR_NOT_CONVERGED, "insertion did not finish"};
  Mesh M = mesh_of(c, R.V, R.mode);
  const uint32_t U = R.U, cur = R.cur, round = R.round;
  launch_set_ctr(M.ctr, C_ULEN, U, C_ROUND, round, s);
  timed(c, K_VOTE, [&] { launch_vote(M, P<uint32_t>(c->uvert[cur]), P<uint32_t>(c->utet[cur]), U, s); });
  timed(c, K_PICK, [&] {
    launch_pick(M, P<uint32_t>(c->uvert[cur]), P<uint32_t>(c->utet[cur]), U, P<uint2>(c->wlist),
                P<uint32_t>(c->uvert[cur ^ 1]), P<uint32_t>(c->utet[cur ^ 1]), P<uint32_t>(c->blk),
                round <= R.warmup ? 0 : c->ring, s);
  });
  HIP_CHECK(hipGetLastError());
  after_pick(c, R, U);
  return R.W;
}

// 1 -> 4 splits of the round's winners (split.wgsl:104-327); the children
// join the active list.
void dev_split(wtet_ctx* c, Run& R) {
  Mesh M = mesh_of(c, R.V, R.mode);
  timed(c, K_SPLIT, [&] { launch_split(M, P<uint2>(c->wlist), R.W, lists_of(c), c->stream); });
  HIP_CHECK(hipGetLastError());
}

// each uninserted point of a split tet into its child (split_points.wgsl:222-335)
void dev_split_points(wtet_ctx* c, Run& R) {
  Mesh M = mesh_of(c, R.V, R.mode);
  timed(c, K_SPLIT_POINTS, [&] {
    launch_split_points(M, P<uint32_t>(c->uvert[R.cur]), P<uint32_t>(c->utet[R.cur]), R.U, R.st,
                        P<uint32_t>(c->slow_p), c->stream);
  });
  HIP_CHECK(hipGetLastError());
}

// the round's flip iterations (src/phase1.rs, flip loop), from device counters
void dev_round_flips(wtet_ctx* c, Run& R) {
  // the first list holds the entries carried over plus four per split;
  // the deferred count per iteration follows the previous round's
  uint64_t dsum = c->h_ctr[C_SLOW_SUM];
  uint64_t dper = (2 * (dsum - R.dsum_prev) + R.max_fi - 1) / (uint64_t)R.max_fi;
  R.dsum_prev = dsum;
  run_flip_batch(c, R.V, R.mode, R.st, R.max_fi, (uint64_t)c->h_ctr[C_A] + 4ull * R.W, dper);
}

// uninserted points of flipped tets to their new tets (relocate_points_fast.wgsl)
void dev_relocate(wtet_ctx* c, Run& R) {
  if (R.U == 0) return;
  Mesh M = mesh_of(c, R.V, R.mode);
  timed(c, K_RELOCATE, [&] {
    launch_relocate(M, P<uint32_t>(c->uvert[R.cur]), P<uint32_t>(c->utet[R.cur]), R.U, R.st,
                    P<uint32_t>(c->slow_p), c->stream);
  });
  HIP_CHECK(hipGetLastError());
}

// One round after its pick as one captured graph: the 1-4 splits, the
// points into the children, the round's flip iterations and the
// relocation, then (unless it is the last round) the next round's vote and
// pick.  Every count comes from ctr and every grid is a power of two over
// the host's bound, so the graph of one (buffer set, grids, variant)
// replays for each round and call that fits it; the host's one read per
// round follows the pick.  When profiling, or with graphs off, the same
// sequence is launched directly.
void run_round(wtet_ctx* c, Run& R, bool last) {
  hipStream_t s = c->stream;
  Mesh M = mesh_of(c, R.V, R.mode);
  FlipLists L = lists_of(c);
  const uint32_t cur = R.cur, W = R.W, U = R.U;
  const int ring_next = R.round + 1 <= R.warmup ? 0 : c->ring;
  // the round's flips, sized as dev_round_flips sizes them
  uint64_t dsum = c->h_ctr[C_SLOW_SUM];
  uint64_t dper = (2 * (dsum - R.dsum_prev) + R.max_fi - 1) / (uint64_t)R.max_fi;
  R.dsum_prev = dsum;
  const uint64_t a_hint = (uint64_t)c->h_ctr[C_A] + 4ull * W;
  const int variant = !R.flipping ? 0 : (a_hint <= kSmallList ? 1 : 2);
  const uint32_t grid = variant == 2 ? grid_for(c, a_hint, 16) : 0u;
  const uint32_t dgrid = variant == 2 ? deferred_grid(c, dper, grid) : 0u;
  const int k = R.max_fi;
  uint32_t* uv = P<uint32_t>(c->uvert[cur]);
  uint32_t* ut = P<uint32_t>(c->utet[cur]);
  uint32_t* slow_p = P<uint32_t>(c->slow_p);
  uint64_t* sums = P<uint64_t>(c->isums);
  auto enqueue = [&] {
    timed(c, K_SPLIT, [&] { launch_split(M, P<uint2>(c->wlist), W, L, s); });
    timed(c, K_SPLIT_POINTS, [&] { launch_split_points(M, uv, ut, U, R.st, slow_p, s); });
    if (variant == 0) {
      launch_set_ctr(M.ctr, C_A, 0u, -1, 0u, s);
    } else {
      if (variant == 1)
        timed(c, K_FLIP_SMALL, [&] { launch_flip_small(M, L, R.st, sums, k, kSmallLimit, s); });
      else
        enqueue_flip_iterations(c, M, L, R.st, k, grid, dgrid);
      timed(c, K_RELOCATE, [&] { launch_relocate(M, uv, ut, U, R.st, slow_p, s); });
    }
    if (!last) {
      launch_round_next(M.ctr, s);
      timed(c, K_VOTE, [&] { launch_vote(M, uv, ut, U, s); });
      timed(c, K_PICK, [&] {
        launch_pick(M, uv, ut, U, P<uint2>(c->wlist), P<uint32_t>(c->uvert[cur ^ 1]),
                    P<uint32_t>(c->utet[cur ^ 1]), P<uint32_t>(c->blk), ring_next, s);
      });
    }
  };
  if (!c->use_graphs || c->profiling) {
    if (c->evused > 4096) collect_times(c);
    enqueue();
  } else {
    launch_cached_graph(c, c->round_graphs, buffer_key(c, M, L, R.st),
                        {(uint64_t)cur, (uint64_t)round_grid(W), (uint64_t)round_grid(U), (uint64_t)(U > 0),
                         (uint64_t)variant, (uint64_t)grid, (uint64_t)dgrid, (uint64_t)k, (uint64_t)last,
                         (uint64_t)ring_next},
                        enqueue);
  }
  HIP_CHECK(hipGetLastError());
}

// batches of flip iterations until the active list is empty
void dev_finish_flips(wtet_ctx* c, Run& R) {
  for (uint32_t guard = 0;; ++guard) {
    if (guard > 20000) throw Fail{WTET_ERR_NOT_CONVERGED, "flipping did not terminate"};
    read_ctr(c);
    if (c->h_ctr[C_STOP]) grow_for(c, (uint64_t)c->h_ctr[C_TETS] + c->h_ctr[C_A] + 64);
    else if (c->h_ctr[C_A] == 0) break;
    run_flip_batch(c, R.V, R.mode, R.st, 16, c->h_ctr[C_A], c->h_ctr[C_A] / 64);
  }
}

// compaction of the live tets into rows grouped by cloud (src/gpu/mod.rs:207-332);
// scratch: ord[0], ord[1] and, for many clouds, ord_tmp plus act[0], act[1]
// (the flip lists, empty once flipping has finished).  out=false computes
// only slot2idx and cstart.
void dev_compact(wtet_ctx* c, Run& R, uint32_t T, bool out, bool keep_lists) {
  OutputBufs OB = {};
  OB.blk = P<uint32_t>(c->blk);
  OB.slot2pos = P<uint32_t>(c->slot2idx);
  OB.cstart = P<uint32_t>(c->cstart);
  if (keep_lists) {
    if (R.C > 1) throw Fail{WTET_ERR_INVALID_ARGUMENT, "a mid-run readback covers one cloud"};
    ensure(c, c->ord[0], (size_t)c->cap_tets * 4);
    OB.olist = P<uint32_t>(c->ord[0]);
  } else {
    OB.olist = P<uint32_t>(c->act[0]);
    OB.okey = P<uint32_t>(c->act[1]);
  }
  if (R.C > 1) {
    ensure(c, c->ord[0], (size_t)c->cap_tets * 4);
    ensure(c, c->ord[1], (size_t)c->cap_tets * 4);
    ensure(c, c->ord_tmp, std::max<size_t>(sort_temp_bytes_u32(T, cloud_key_bits(R.C)), 16));
    OB.olist2 = P<uint32_t>(c->ord[0]);
    OB.okey2 = P<uint32_t>(c->ord[1]);
    OB.temp = c->ord_tmp.p;
    OB.temp_bytes = c->ord_tmp.bytes;
  }
  if (out) {
    OB.out_tets = P<uint32_t>(c->prop);
    OB.out_adj = P<uint32_t>(c->fout);
    // the sort keys are free by now: 2n words hold the n + 4C + 1 ranks
    // (every cloud has at least 4 points)
    ensure(c, c->keys, ((size_t)R.V + 1) * 4 + 16, false);
    OB.vrank = P<uint32_t>(c->keys);
  }
  Mesh M = mesh_of(c, R.V, R.mode);
  timed(c, K_OUTPUT, [&] { HIP_CHECK(launch_output(M, T, R.CL, OB, c->stream)); });
  HIP_CHECK(hipGetLastError());
}

// The audit once more into a record buffer that holds every violation the
// first pass counted (C_FAILREC); leaves the counters in h_ctr.
void reaudit(wtet_ctx* c, Run& R, uint32_t T) {
  uint32_t need = c->h_ctr[C_FAILREC];
  ensure(c, c->rec, ((size_t)need + need / 8 + 1024) * 4);
  Mesh M = mesh_of(c, R.V, R.mode);
  launch_set_ctr(M.ctr, C_VIOL, 0u, C_FAILREC, 0u, c->stream);
  const uint32_t rec_cap = (uint32_t)std::min<size_t>(c->rec.bytes / 4, 0xFFFFFFFFu);
  timed(c, K_AUDIT, [&] {
    launch_audit(M, T, P<uint32_t>(c->rec), rec_cap, nullptr, P<uint32_t>(c->slow_t), c->stream);
  });
  HIP_CHECK(hipGetLastError());
  read_ctr(c);
  if (c->h_ctr[C_FAILREC] > rec_cap) throw Fail{WTET_ERR_INTERNAL, "audit records overflowed twice"};
}

// Audit of every finite face (gather.wgsl:68-151), compaction into rows,
// and the copies the host needs.  Leaves the rows in prop/fout (cloud c
// from cstart[c]), the violation records as rows, and -- unless resident --
// a copy of the rows in the pinned staging buffers.
void dev_output(wtet_ctx* c, Run& R, std::vector<uint32_t>& viol_rec, uint32_t& nviol, bool resident,
                uint32_t& mout, std::vector<uint32_t>& cstart) {
  hipStream_t s = c->stream;
  const uint32_t C = R.C;
  uint64_t sums[S_COUNT] = {};
  HIP_CHECK(hipMemcpyAsync(sums, c->isums.p, S_COUNT * 8, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  R.K.sumA = sums[S_CHECKED];
  R.K.sumF = sums[S_FLIPS];
  R.K.iters = (uint32_t)sums[S_ITERS];
  R.K.flips23 = sums[S_FLIPS23];
  R.K.smallA = sums[S_SMALL_CHECKED];
  R.K.smallF = sums[S_SMALL_FLIPS];
  // audit every finite face
  read_ctr(c);
  uint32_t T = c->h_ctr[C_TETS];
  Mesh M = mesh_of(c, R.V, R.mode);
  launch_set_ctr(M.ctr, C_VIOL, 0u, C_FAILREC, 0u, s);
  const uint32_t rec_cap = (uint32_t)std::min<size_t>(c->rec.bytes / 4, c->audit_records);
  timed(c, K_AUDIT, [&] { launch_audit(M, T, P<uint32_t>(c->rec), rec_cap, R.st, P<uint32_t>(c->slow_t), s); });
  dev_compact(c, R, T, true, false);
  HIP_CHECK(hipEventRecord(c->ev1, s));
  cstart.resize((size_t)C + 1);
  HIP_CHECK(hipMemcpyAsync(cstart.data(), c->cstart.p, ((size_t)C + 1) * 4, hipMemcpyDeviceToHost, s));
  read_ctr(c);
  collect_times(c);
  float ms = 0;
  HIP_CHECK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->stats.ms_device = ms;
  c->stats.walk_failures = c->h_ctr[C_WALK_FAIL];
  R.K.peakA = c->h_ctr[C_PEAKA];
  if (c->h_ctr[C_WALK_FAIL]) throw Fail{WTET_ERR_INTERNAL, "point relocation walk did not terminate"};
  uint32_t Mout = c->h_ctr[C_ALIVE];
  mout = Mout;
  if (cstart[C] != Mout) throw Fail{WTET_ERR_INTERNAL, "output row count mismatch"};
  // more violations than records: a larger record buffer and the audit
  // again (predicate counters already hold the first pass), so that the
  // repair starts from every violated face
  if (c->h_ctr[C_FAILREC] > rec_cap) reaudit(c, R, T);
  nviol = c->h_ctr[C_VIOL];
  // the host needs every row only for a host result; a resident result's
  // repair downloads the neighbourhood it works on
  if (!resident) {
    // room for the repair to append (it rarely adds more than a few tets)
    // and for the tets of the points the insertion cap left out
    pin_join(c);
    size_t cap = (size_t)Mout + Mout / 32 + 65536 + 8ull * R.K.left;
    size_t c1 = c->h_cap, c2 = c->h_cap;
    ensure_pinned(c->h_tets, c1, cap * 4);
    ensure_pinned(c->h_adj, c2, cap * 4);
    c->h_cap = std::min(c1, c2);
    HIP_CHECK(hipMemcpyAsync(c->h_tets, c->prop.p, (size_t)Mout * 16, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(c->h_adj, c->fout.p, (size_t)Mout * 16, hipMemcpyDeviceToHost, s));
  }
  uint32_t nrec = c->h_ctr[C_FAILREC];
  viol_rec.resize(nrec);
  if (nrec) {
    // records are slot-based; map them to rows on the device
    launch_map_records(P<uint32_t>(c->rec), nrec, P<uint32_t>(c->slot2idx), s);
    HIP_CHECK(hipMemcpyAsync(viol_rec.data(), c->rec.p, (size_t)nrec * 4, hipMemcpyDeviceToHost, s));
  }
  if (R.cfg.collect_stats)
    HIP_CHECK(hipMemcpyAsync(c->stats.pred, c->gstats.p, 64, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  c->stats.tet_slots = T;
}

// The device phase for clouds off[c]..off[c+1] of d_xyz (one cloud: {0, n}),
// the loop of src/phase1.rs:16-1431 with one host read per round.
void run_device(wtet_ctx* c, const float* d_xyz, const std::vector<uint32_t>& off,
                const wtet_config& cfg, Counters& K, std::vector<uint32_t>& viol_rec,
                uint32_t& nviol, bool resident, uint32_t& mout, std::vector<uint32_t>& cstart) {
  end_stage(c);  // the call reuses the buffers a stage session holds
  pin_join(c);
  // a single cloud's host work needs a mirror of about every row (the
  // Delaunay mesh of n points in general position has about 6.7n tets):
  // on a first call at this size it is allocated while the device works
  if (off.size() == 2 && 7ull * off[1] >= (1u << 22)) {
    const size_t rows = 7ull * off[1];
    pin_ahead(c, rows + rows / 32 + 65536);
  }
  Run R;
  run_setup(c, R, off, cfg);
  dev_prepare(c, R, d_xyz);
  // at most max_insert_iterations rounds (src/phase1.rs:26); the host phase
  // inserts what they leave.  The first round's vote and pick are launched
  // one by one, every later round runs from a captured graph (run_round).
  if (R.U > 0 && R.round < R.cfg.max_insert_iterations) {
    dev_vote(c, R);
    for (;;) {
      const bool last = R.U == 0 || R.round >= R.cfg.max_insert_iterations;
      const uint32_t U = R.U;
      run_round(c, R, last);
      if (last) break;
      ++R.round;  // the device counted it too (launch_round_next)
      if (R.round > (1u << 16)) throw Fail{WTET_ERR_NOT_CONVERGED, "insertion did not finish"};
      after_pick(c, R, U);
    }
  }
  R.K.left = R.U;
  if (R.flipping) dev_finish_flips(c, R);
  dev_output(c, R, viol_rec, nviol, resident, mout, cstart);
  K = R.K;
}

}  // namespace

// ---------------------------------------------------------------------------
// stage-level session (wtet_stage_*): the state of run_device between calls
// ---------------------------------------------------------------------------
struct wtet_stage_session {
  Run R;
  // 1: between rounds (vote, flip, relocate, gather, readback, finish),
  // 2: voted (split next), 3: split (split_points next)
  int phase = 0;
  bool stale = false;        // flips since the last relocation
  bool audit_valid = false;  // the last gather saw the current mesh
  uint32_t checked = 0, flips = 0, violations = 0, nrec = 0, deferred = 0;
};

namespace {

void end_stage(wtet_ctx* c) {
  delete c->stage;
  c->stage = nullptr;
}

}  // namespace

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
namespace {

wtet_status run_guarded(wtet_ctx* c, void (*fn)(void*), void* arg) {
  try {
    fn(arg);
    return WTET_OK;
  } catch (const Fail& f) {
    if (c) c->err = f.msg;
    return f.st;
  } catch (const HipError& h) {
    if (c) c->err = std::string("HIP error ") + hipGetErrorString(h.e) + " in " + h.what;
    return WTET_ERR_DEVICE;
  } catch (const std::bad_alloc&) {
    if (c) c->err = "host allocation failed";
    return WTET_ERR_OUT_OF_MEMORY;
  } catch (...) {
    if (c) c->err = "unexpected exception";
    return WTET_ERR_INTERNAL;
  }
}

struct DelaunayArgs {
  wtet_ctx* c;
  const float* xyz;
  bool on_device;
  uint32_t n;
  wtet_config cfg;
  wtet_result* out;           // host result, or
  wtet_device_result* dout;   // device-resident result
};

// Rows the host repair will read, for a resident result: those within a
// few face-steps of the recorded faces and the tail rows that fill holes,
// with their neighbours.  They land in the pinned staging at their own row;
// present[] marks them.  rec (rows) is on the device in c->rec.
constexpr int kRepairRingsMax = 16;
constexpr uint32_t kRepairTail = 4096;

// packed rows list[from, to) of the neighbourhood into the host mirror
void scatter_rows(wtet_ctx* c, uint32_t from, uint32_t to) {
  hipStream_t s = c->stream;
  uint32_t n = to - from;
  if (n == 0) return;
  // room for a sixteenth of the mesh at once, so that the staged row count
  // varying from call to call does not reallocate these buffers
  size_t room = std::max<size_t>(n, c->present.size() / 16);
  ensure(c, c->nbr_pack, room * 36 + 16);
  launch_nbr_gather(P<uint32_t>(c->nbr_list) + from, n, P<uint32_t>(c->prop), P<uint32_t>(c->fout),
                    P<uint32_t>(c->nbr_pack), s);
  HIP_CHECK(hipGetLastError());
  ensure_pinned(c->h_patch, c->h_patch_cap, room * 9 + 9);
  HIP_CHECK(hipMemcpyAsync(c->h_patch, c->nbr_pack.p, (size_t)n * 36, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  // rows land at random places of the mirror: up to 8 threads keep more of
  // those cache-missing stores in flight (distinct rows, distinct bytes)
  const uint32_t* hp = c->h_patch;
  uint32_t* ht = c->h_tets;
  uint32_t* ha = c->h_adj;
  uint8_t* pres = c->present.data();
  const uint8_t mark = c->present_mark;
  auto part = [hp, ht, ha, pres, mark](size_t a, size_t b) {
    for (size_t i = a; i < b; ++i) {
      const uint32_t* q = hp + 9 * i;
      uint32_t r = q[0];
      std::memcpy(ht + 4 * (size_t)r, q + 1, 16);
      std::memcpy(ha + 4 * (size_t)r, q + 5, 16);
      pres[r] = mark;
    }
  };
  unsigned nt = n >= (1u << 15) ? 8u : 1u;
  if (nt == 1) {
    part(0, n);
  } else {
    std::vector<std::thread> th;
    size_t per = (n + nt - 1) / nt;
    for (unsigned k = 0; k < nt; ++k)
      th.emplace_back(part, std::min<size_t>(n, per * k), std::min<size_t>(n, per * (k + 1)));
    for (auto& t : th) t.join();
  }
  c->stats.repair_rows_staged += n;
}

// rows listed so far (device counter), checked against the mesh size
uint32_t listed_rows(wtet_ctx* c, uint32_t M) {
  uint32_t n = 0;
  HIP_CHECK(hipMemcpyAsync(&n, c->nbr_misc.p, 4, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  if (n > M) throw Fail{WTET_ERR_INTERNAL, "neighbourhood larger than the mesh"};
  return n;
}

void stage_neighbourhood(wtet_ctx* c, uint32_t nrec, uint32_t M, int rings) {
  hipStream_t s = c->stream;
  size_t cap = (size_t)M + M / 32 + 65536;
  size_t c1 = c->h_cap, c2 = c->h_cap;
  ensure_pinned(c->h_tets, c1, cap * 4);
  ensure_pinned(c->h_adj, c2, cap * 4);
  c->h_cap = std::min(c1, c2);
  uint32_t tail = std::min(M, kRepairTail);
  ensure(c, c->nbr_mark, (size_t)M * 4 + 16);
  ensure(c, c->nbr_list, (size_t)M * 4 + 16);
  ensure(c, c->nbr_misc, 4 * (kRepairRingsMax + 4));
  uint32_t* cnt = P<uint32_t>(c->nbr_misc);
  launch_neighbourhood(P<uint32_t>(c->rec), nrec, M, tail, rings, P<uint32_t>(c->fout),
                       P<uint32_t>(c->nbr_mark), P<uint32_t>(c->nbr_list), cnt, cnt + 1, false, s);
  HIP_CHECK(hipGetLastError());
  uint32_t n = listed_rows(c, M);
  // a few rows more than the last call: the flags are allocated again
  // (all clear) with room to spare, not copied to a vector just as long
  if (c->present.size() < M) c->present.assign((size_t)M + M / 16, 0);
  if (++c->present_mark == 0) {
    std::fill(c->present.begin(), c->present.end(), 0);
    c->present_mark = 1;
  }
  c->stats.repair_rows_staged = 0;
  scatter_rows(c, 0, n);
}

// the repair reached `row`, which the host lacks: the rows within the
// staging rings of it that are not on the host yet.  Each fetch restarts a
// region, so after kMaxFetches of them, or once a quarter of the mesh is on
// the host, it returns false and the repair takes every row instead.  A
// fetch costs a fraction of a millisecond at 1M points, the full download
// and repair 10-15 ms at 1M and about 200 ms at 5M, where 5-6 fetches are
// usual.
constexpr uint32_t kMaxFetches = 32;

bool fetch_rows(wtet_ctx* c, uint32_t row, uint32_t M, uint32_t& fetches) {
  if (row >= M || c->present[row] == c->present_mark || ++fetches > kMaxFetches) return false;
  uint32_t before = listed_rows(c, M);
  if (before > M / 4) return false;
  uint32_t* misc = P<uint32_t>(c->nbr_misc);
  uint32_t* seed = misc + kRepairRingsMax + 3;  // after the count and the ring bounds
  uint32_t key = row << 2;
  HIP_CHECK(hipMemcpyAsync(seed, &key, 4, hipMemcpyHostToDevice, c->stream));
  // wider than the first staging: a region that left it is a large one
  const int rings = std::min(c->repair_rings + 6, (int)kRepairRingsMax);
  launch_neighbourhood(seed, 1, M, 0, rings, P<uint32_t>(c->fout), P<uint32_t>(c->nbr_mark),
                       P<uint32_t>(c->nbr_list), misc, misc + 1, true, c->stream);
  HIP_CHECK(hipGetLastError());
  uint32_t after = listed_rows(c, M);
  scatter_rows(c, before, after);
  return c->present[row] == c->present_mark;
}

// fetch_rows for every row of `rows` the host lacks, in one pass: one
// neighbourhood launch seeded by all of them, one copy (one fetch of the
// kMaxFetches)
bool fetch_rows_many(wtet_ctx* c, const std::vector<uint32_t>& rows, uint32_t M, uint32_t& fetches) {
  std::vector<uint32_t> seeds;
  for (uint32_t r : rows)
    if (r < M && c->present[r] != c->present_mark) seeds.push_back(r << 2);
  if (seeds.empty()) return true;
  if (++fetches > kMaxFetches) return false;
  uint32_t before = listed_rows(c, M);
  if (before > M / 4) return false;
  ensure(c, c->nbr_seed, seeds.size() * 4 + 16);
  HIP_CHECK(hipMemcpyAsync(c->nbr_seed.p, seeds.data(), seeds.size() * 4, hipMemcpyHostToDevice, c->stream));
  uint32_t* misc = P<uint32_t>(c->nbr_misc);
  const int rings = std::min(c->repair_rings + 6, (int)kRepairRingsMax);
  launch_neighbourhood(P<uint32_t>(c->nbr_seed), (uint32_t)seeds.size(), M, 0, rings, P<uint32_t>(c->fout),
                       P<uint32_t>(c->nbr_mark), P<uint32_t>(c->nbr_list), misc, misc + 1, true, c->stream);
  HIP_CHECK(hipGetLastError());
  uint32_t after = listed_rows(c, M);  // synchronizes: seeds stay alive until then
  scatter_rows(c, before, after);
  for (uint32_t r : rows)
    if (r < M && c->present[r] != c->present_mark) return false;
  return true;
}

void download_rows(wtet_ctx* c, uint32_t M, uint32_t left = 0, size_t min_cap = 0) {
  size_t cap = std::max<size_t>((size_t)M + M / 32 + 65536 + 8ull * left, min_cap);
  size_t c1 = c->h_cap, c2 = c->h_cap;
  ensure_pinned(c->h_tets, c1, cap * 4);
  ensure_pinned(c->h_adj, c2, cap * 4);
  c->h_cap = std::min(c1, c2);
  HIP_CHECK(hipMemcpyAsync(c->h_tets, c->prop.p, (size_t)M * 16, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipMemcpyAsync(c->h_adj, c->fout.p, (size_t)M * 16, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  c->stats.repair_rows_staged = M;
}

// the normalized points on the host (h_xyzn): the copy dev_prepare started,
// or one now
void host_points(wtet_ctx* c, uint32_t n) {
  if (c->xyzn_copied == n) {
    HIP_CHECK(hipEventSynchronize(c->ev_xyzn));
    return;
  }
  HIP_CHECK(hipStreamSynchronize(c->copy_stream));
  ensure_pinned(c->h_xyzn, c->h_npts, (size_t)n * 3);
  HIP_CHECK(hipMemcpyAsync(c->h_xyzn, c->xyzn.p, (size_t)n * 12, hipMemcpyDeviceToHost, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  c->xyzn_copied = n;
}

// failed_verts (gather.wgsl:68-151): the vertex opposite each non-Delaunay
// face between two real-only tets, from violation records (row << 2 | face)
std::vector<uint32_t> failed_vertices(const uint32_t* tets, const uint32_t* adj,
                                      const std::vector<uint32_t>& rec, uint32_t n) {
  std::vector<uint32_t> failed;
  for (uint32_t r : rec) {
    uint32_t t = r >> 2, f = r & 3u;
    uint32_t e = adj[4 * (size_t)t + f];
    if (e == kInvalid) continue;
    uint32_t u = e >> 5, g = e & 3u;
    bool real = true;
    for (int k = 0; k < 4; ++k) real = real && tets[4 * (size_t)t + k] < n && tets[4 * (size_t)u + k] < n;
    if (real) failed.push_back(tets[4 * (size_t)u + g]);
  }
  std::sort(failed.begin(), failed.end());
  failed.erase(std::unique(failed.begin(), failed.end()), failed.end());
  return failed;
}

// Host side after the device phase of one cloud: failed_verts, the host
// repair (src/lib.rs:227-307) and the result, as host arrays (out) or as
// device rows (dout).
// Thrown by finish_single when a repair region outgrew region_cap: the
// caller reruns the device phase with a wider insertion ring.
struct WiderRing {};

void finish_single(wtet_ctx* c, uint32_t n, const wtet_config& cfg, const Counters& K,
                   std::vector<uint32_t>& rec, uint32_t nviol, uint32_t mout, wtet_result* out,
                   wtet_device_result* dout, std::chrono::steady_clock::time_point t0,
                   uint32_t region_cap) {
  const bool resident = dout != nullptr;
  const uint32_t left = K.left;
  pin_join(c);
  c->stats.violations_gpu = nviol;
  c->stats.points_left = left;
  auto t1 = std::chrono::steady_clock::now();
  bool sparse = false;
  if (resident && (nviol > 0 || left > 0)) {
    if (left > 0 || rec.size() < nviol) {
      download_rows(c, mout, left);  // point insertion needs every row
    } else {
      // Stuck faces denser than one per 2,000 points come in knots whose
      // regions reach past the staged rings (1M sphere points: 830-1,020
      // faces, 5-15 fetches at 7 rings, none at 10; repair median 24.8 ->
      // 14.5 ms); sparser ones (1M uniform: 120-200) keep the narrower
      // staging, which is 0.7 ms cheaper there.  A rerun with the wider
      // insertion ring (ring 2) leaves few stuck faces in large regions:
      // it stages the wider rings too.
      const bool wide = (uint64_t)nviol * 2000u > n || c->ring >= 2;
      const int rings = std::min(c->repair_rings + (wide ? 3 : 0), kRepairRingsMax);
      stage_neighbourhood(c, (uint32_t)rec.size(), mout, rings);
      sparse = true;
    }
  }
  MeshBuf mb{c->h_tets, c->h_adj, mout, c->h_cap / 4, {}};
  uint32_t fetches = 0;
  if (sparse) {
    mb.present = c->present.data();
    mb.present_mark = c->present_mark;
    mb.known = mout;
    mb.fetch = [c, mout, &fetches](uint32_t row) { return fetch_rows(c, row, mout, fetches); };
    mb.fetch_many = [c, mout, &fetches](const std::vector<uint32_t>& rows) {
      return fetch_rows_many(c, rows, mout, fetches);
    };
  }
  bool on_host = !resident || nviol > 0 || left > 0;
  std::vector<uint32_t> failed;
  if (on_host) failed = failed_vertices(mb.tets, mb.adj, rec, n);
  // points the insertion cap left out (exact duplicates, never inserted,
  // are among them until the host insertion skips them)
  std::vector<uint32_t> missing;
  if (left > 0) missing = missing_points(mb.tets, mb.count, n);
  bool repaired = false;
  if ((nviol > 0 || !missing.empty()) && cfg.enable_splaying) {
    // the repair needs the normalized coordinates of the input ids
    host_points(c, n);
  }
  // The repair and the insertion write into the staging rows.  The Delaunay
  // mesh may need many more tets than the rows hold (a mesh from splits
  // alone has about 3 per point, the Delaunay one about 6.7): when they run
  // out of room, every row is downloaded again into staging twice as large
  // and both start over.
  std::vector<uint32_t> added = missing;
  for (int attempt = 0;; ++attempt) {
    wtet_status rs = WTET_OK;
    std::string msg;
    if (nviol > 0 && cfg.enable_splaying) {
      auto tc = std::chrono::steady_clock::now();
      c->stats.ms_repair_staging = std::chrono::duration<double, std::milli>(tc - t1).count();
      RepairReport rr;
      rs = repair_delaunay(c->h_xyzn, n, mb, rec, cfg.predicate_mode, &rr, region_cap);
      if (rs != WTET_OK && rr.needs_full && !rr.too_large) {
        // the neighbourhood was too small: every row, and again from the start
        download_rows(c, mout, left);
        mb = MeshBuf{c->h_tets, c->h_adj, mout, c->h_cap / 4, {}};
        rr = RepairReport{};
        rs = repair_delaunay(c->h_xyzn, n, mb, rec, cfg.predicate_mode, &rr, region_cap);
        c->stats.repair_full_downloads++;
      }
      c->stats.repair_max_region = rr.max_region;
      if (rs != WTET_OK && rr.too_large) throw WiderRing{};
      c->stats.repair_regions = rr.regions;
      c->stats.repair_fetches = std::min(fetches, kMaxFetches);
      c->stats.ms_repair_compute =
          std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - tc).count();
      msg = "host repair failed: " + rr.message;
      repaired = true;
    }
    if (rs == WTET_OK && !missing.empty() && cfg.enable_splaying) {
      // phase 2 inserts what phase 1 left (src/lib.rs:227-307 rebuilds with
      // every point); failed_verts lists those points
      RepairReport ri;
      rs = insert_points(c->h_xyzn, n, mb, missing, cfg.predicate_mode, &ri);
      msg = "host point insertion failed: " + ri.message;
      if (rs == WTET_OK) {
        std::vector<uint32_t> still = missing_points(mb.tets, mb.count, n);
        added.clear();
        std::set_difference(missing.begin(), missing.end(), still.begin(), still.end(),
                            std::back_inserter(added));
      }
      repaired = true;
    }
    if (rs == WTET_ERR_OUT_OF_MEMORY && attempt < 6) {
      download_rows(c, mout, left, 2 * mb.cap);
      mb = MeshBuf{c->h_tets, c->h_adj, mout, c->h_cap / 4, {}};
      continue;
    }
    if (rs != WTET_OK) throw Fail{rs, msg};
    break;
  }
  if (!missing.empty()) {
    std::vector<uint32_t> u;
    std::set_union(failed.begin(), failed.end(), added.begin(), added.end(), std::back_inserter(u));
    failed.swap(u);
  }
  auto t2 = std::chrono::steady_clock::now();
  c->stats.ms_repair = std::chrono::duration<double, std::milli>(t2 - t1).count();
  c->stats.sum_uninserted = K.sumU;
  c->stats.sum_active = K.sumA;
  c->stats.sum_flips = K.sumF;
  c->stats.flips_23 = K.flips23;
  c->stats.flips_32 = K.sumF - K.flips23;
  c->stats.small_checked = K.smallA;
  c->stats.small_flips = K.smallF;
  c->stats.rounds = K.rounds;
  c->stats.flip_iterations = K.iters;
  c->stats.peak_active = K.peakA;
  c->stats.num_points = n;
  if (resident) {
    if (repaired) {
      // upload only the slots the repair wrote: (slot, tet[4], adj[4])
      size_t np = mb.changed.size();
      if (np) {
        ensure_pinned(c->h_patch, c->h_patch_cap, np * 9);
        for (size_t i = 0; i < np; ++i) {
          uint32_t sl = mb.changed[i];
          uint32_t* q = c->h_patch + 9 * i;
          q[0] = sl;
          std::memcpy(q + 1, mb.tets + 4 * (size_t)sl, 16);
          std::memcpy(q + 5, mb.adj + 4 * (size_t)sl, 16);
        }
        ensure(c, c->patch, np * 36);
        // the repair and the host insertion may add rows
        ensure(c, c->prop, (size_t)mb.count * 16, true);
        ensure(c, c->fout, (size_t)mb.count * 16, true);
        HIP_CHECK(hipMemcpyAsync(c->patch.p, c->h_patch, np * 36, hipMemcpyHostToDevice, c->stream));
        launch_apply_patch(P<uint32_t>(c->patch), (uint32_t)np, P<uint32_t>(c->prop),
                           P<uint32_t>(c->fout), c->stream);
        HIP_CHECK(hipGetLastError());
        HIP_CHECK(hipStreamSynchronize(c->stream));
      }
      mout = (uint32_t)mb.count;
    }
    dout->num_tets = mout;
    dout->d_tets = P<uint32_t>(c->prop);
    dout->d_adjacency = P<uint32_t>(c->fout);
    dout->num_failed = (uint32_t)failed.size();
    c->stats.num_tets_out = mout;
    c->stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return;
  }
  // straight from the staged rows into the caller's arrays
  const size_t words = 4 * mb.count;
  out->num_tets = (uint32_t)mb.count;
  out->tets = alloc_rows(words * 4);
  out->adjacency = alloc_rows(words * 4);
  out->num_failed = (uint32_t)failed.size();
  out->failed_verts = (uint32_t*)std::malloc(std::max<size_t>(failed.size(), 1) * 4);
  if (!out->tets || !out->adjacency || !out->failed_verts) throw std::bad_alloc();
  parallel_copy(out->tets, mb.tets, words * 4);
  parallel_copy(out->adjacency, mb.adj, words * 4);
  if (!failed.empty()) std::memcpy(out->failed_verts, failed.data(), failed.size() * 4);
  c->stats.num_tets_out = out->num_tets;
  c->stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

void delaunay_impl(void* a_) {
  DelaunayArgs& a = *static_cast<DelaunayArgs*>(a_);
  wtet_ctx* c = a.c;
  auto t0 = std::chrono::steady_clock::now();
  HIP_CHECK(hipSetDevice(c->device));
  std::memset(&c->stats, 0, sizeof(c->stats));
  const float* d_xyz = a.xyz;
  if (!a.on_device) {
    ensure(c, c->xyz_in, (size_t)a.n * 12);
    HIP_CHECK(hipMemcpyAsync(c->xyz_in.p, a.xyz, (size_t)a.n * 12, hipMemcpyHostToDevice, c->stream));
    d_xyz = P<float>(c->xyz_in);
  }
  bool resident = a.dout != nullptr;
  // Dense stuck faces (a flip trajectory that left thousands, as happens on
  // co-spherical inputs) can merge the repair's regions into one that spans
  // the mesh, a host re-triangulation of every point.  A region past
  // region_cap instead reruns the device phase with the insertion ring one
  // wider (fewer simultaneous insertions near each other, fewer stuck
  // faces); the rerun's repair has no cap.
  const int ring0 = c->ring;
  struct RingGuard {
    wtet_ctx* c;
    int ring;
    ~RingGuard() { c->ring = ring; }
  } guard{c, ring0};
  uint32_t retries = 0;
  for (;;) {
    std::memset(&c->stats, 0, sizeof(c->stats));
    Counters K;
    std::vector<uint32_t> rec, cstart;
    uint32_t nviol = 0, mout = 0;
    run_device(c, d_xyz, {0u, a.n}, a.cfg, K, rec, nviol, resident, mout, cstart);
    uint32_t cap = c->ring < 2 ? c->region_cap : 0u;
    // A trajectory that left stuck faces by the thousand (over a quarter of
    // a percent of the points) merges them into regions past the cap: the
    // wider ring at once, instead of after a repair that would stop there
    // (1M sphere points: 830-1,020 stuck faces in most runs, 3,100-3,800 in
    // the others, whose repair ran 400-600 ms before its region passed the
    // cap)
    if (cap && a.cfg.enable_splaying && nviol > std::max<uint32_t>(2000u, a.n / 400u)) {
      ++retries;
      c->ring += 1;
      continue;
    }
    try {
      finish_single(c, a.n, a.cfg, K, rec, nviol, mout, a.out, a.dout, t0, cap);
      break;
    } catch (const WiderRing&) {
      ++retries;
      c->ring += 1;
    }
  }
  c->stats.repair_retries = retries;
}


struct BatchArgs {
  wtet_ctx* c;
  const float* xyz;
  bool on_device;
  std::vector<uint32_t> off;
  wtet_config cfg;
  wtet_batch_result* out;
  wtet_batch_device_result* dout;
};

// One cloud's rows after the device phase: a view of the staged rows, or
// the repaired copy.
struct CloudRows {
  std::vector<uint32_t> tets, adj;  // filled when the cloud was repaired
  bool repaired = false;
};

void batch_impl(void* a_) {
  BatchArgs& a = *static_cast<BatchArgs*>(a_);
  wtet_ctx* c = a.c;
  auto t0 = std::chrono::steady_clock::now();
  HIP_CHECK(hipSetDevice(c->device));
  std::memset(&c->stats, 0, sizeof(c->stats));
  const uint32_t C = (uint32_t)a.off.size() - 1;
  const uint32_t n = a.off[C];
  const float* d_xyz = a.xyz;
  if (!a.on_device) {
    ensure(c, c->xyz_in, (size_t)n * 12);
    HIP_CHECK(hipMemcpyAsync(c->xyz_in.p, a.xyz, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
    d_xyz = P<float>(c->xyz_in);
  }
  Counters K;
  std::vector<uint32_t> rec, cstart;
  uint32_t nviol = 0, mout = 0;
  bool resident = a.dout != nullptr;
  run_device(c, d_xyz, a.off, a.cfg, K, rec, nviol, resident, mout, cstart);
  c->stats.violations_gpu = nviol;
  c->stats.points_left = K.left;
  auto t1 = std::chrono::steady_clock::now();

  // violation records (rows) per cloud, as cloud-local (row << 2) | face
  std::vector<std::pair<uint32_t, uint32_t>> byc;  // (cloud, local record)
  byc.reserve(rec.size());
  for (uint32_t r : rec) {
    uint32_t row = r >> 2;
    uint32_t cl = (uint32_t)(std::upper_bound(cstart.begin(), cstart.end(), row) - cstart.begin()) - 1;
    byc.emplace_back(cl, ((row - cstart[cl]) << 2) | (r & 3u));
  }
  std::sort(byc.begin(), byc.end());
  // where each cloud's rows sit in the pinned staging: the whole mesh for a
  // host result; for a resident result only the clouds with violations (every
  // cloud when the insertion cap left points), downloaded here one segment each
  std::vector<uint64_t> hoff(C, 0);
  std::vector<uint8_t> on_host(C, resident ? 0 : 1);
  for (auto& pr : byc) on_host[pr.first] = 1;
  if (K.left > 0) on_host.assign(C, 1);
  if (!resident) {
    for (uint32_t cl = 0; cl < C; ++cl) hoff[cl] = cstart[cl];
  } else if (!byc.empty() || K.left > 0) {
    uint64_t tot = 0;
    for (uint32_t cl = 0; cl < C; ++cl)
      if (on_host[cl]) {
        hoff[cl] = tot;
        tot += cstart[cl + 1] - cstart[cl];
      }
    size_t c1 = c->h_cap, c2 = c->h_cap;
    ensure_pinned(c->h_tets, c1, (size_t)tot * 4 + 4);
    ensure_pinned(c->h_adj, c2, (size_t)tot * 4 + 4);
    c->h_cap = std::min(c1, c2);
    for (uint32_t cl = 0; cl < C; ++cl)
      if (on_host[cl]) {
        size_t cnt = cstart[cl + 1] - cstart[cl];
        HIP_CHECK(hipMemcpyAsync(c->h_tets + 4 * hoff[cl], P<uint32_t>(c->prop) + 4 * (size_t)cstart[cl],
                                 cnt * 16, hipMemcpyDeviceToHost, c->stream));
        HIP_CHECK(hipMemcpyAsync(c->h_adj + 4 * hoff[cl], P<uint32_t>(c->fout) + 4 * (size_t)cstart[cl],
                                 cnt * 16, hipMemcpyDeviceToHost, c->stream));
      }
    HIP_CHECK(hipStreamSynchronize(c->stream));
  }
  // failed_verts per cloud: the vertex opposite each non-Delaunay real-only face
  std::vector<std::vector<uint32_t>> failed(C);
  for (auto& pr : byc) {
    uint32_t cl = pr.first, nc = a.off[cl + 1] - a.off[cl];
    const uint32_t* T = c->h_tets + 4 * (size_t)hoff[cl];
    const uint32_t* A = c->h_adj + 4 * (size_t)hoff[cl];
    uint32_t t = pr.second >> 2, f = pr.second & 3u;
    uint32_t e = A[4 * (size_t)t + f];
    if (e == kInvalid) continue;
    uint32_t u = e >> 5, g = e & 3u;
    bool real = true;
    for (int k = 0; k < 4; ++k) real = real && T[4 * (size_t)t + k] < nc && T[4 * (size_t)u + k] < nc;
    if (real) failed[cl].push_back(T[4 * (size_t)u + g]);
  }
  for (auto& v : failed) {
    std::sort(v.begin(), v.end());
    v.erase(std::unique(v.begin(), v.end()), v.end());
  }
  // points the insertion cap left out, per cloud (cloud-local ids)
  std::vector<std::vector<uint32_t>> missing(C);
  if (K.left > 0)
    for (uint32_t cl = 0; cl < C; ++cl)
      missing[cl] = missing_points(c->h_tets + 4 * (size_t)hoff[cl], cstart[cl + 1] - cstart[cl],
                                   a.off[cl + 1] - a.off[cl]);
  // host repair of each cloud with violations, and insertion of the points
  // the cap left, on a copy of its rows
  std::vector<CloudRows> rows(C);
  if ((nviol > 0 || K.left > 0) && a.cfg.enable_splaying) {
    host_points(c, n);
    // clouds are independent: repair them on several host threads
    struct Job {
      uint32_t cl;
      std::vector<uint32_t> viol;
      wtet_status st = WTET_OK;
      RepairReport rr;
    };
    std::vector<Job> jobs;
    for (size_t i = 0; i < byc.size();) {
      Job j;
      j.cl = byc[i].first;
      for (; i < byc.size() && byc[i].first == j.cl; ++i) j.viol.push_back(byc[i].second);
      jobs.push_back(std::move(j));
    }
    {
      std::vector<uint8_t> has(C, 0);
      for (const Job& jb : jobs) has[jb.cl] = 1;
      for (uint32_t cl = 0; cl < C; ++cl)
        if (!has[cl] && !missing[cl].empty()) {
          Job j;
          j.cl = cl;
          jobs.push_back(std::move(j));
        }
    }
    std::atomic<size_t> next{0};
    auto work = [&]() {
      for (size_t k; (k = next.fetch_add(1)) < jobs.size();) {
        Job& jb = jobs[k];
        uint32_t cl = jb.cl;
        size_t cnt = cstart[cl + 1] - cstart[cl];
        size_t cap = cnt + cnt / 8 + 4096 + 8 * missing[cl].size();
        CloudRows& R = rows[cl];
        const float* xc = c->h_xyzn + 3 * (size_t)a.off[cl];
        const uint32_t nc = a.off[cl + 1] - a.off[cl];
        // out of room (the Delaunay mesh may need far more tets than the
        // rows hold): start over from the cloud's rows with twice the room
        for (int attempt = 0;; ++attempt) {
          R.tets.assign(c->h_tets + 4 * (size_t)hoff[cl], c->h_tets + 4 * (size_t)(hoff[cl] + cnt));
          R.adj.assign(c->h_adj + 4 * (size_t)hoff[cl], c->h_adj + 4 * (size_t)(hoff[cl] + cnt));
          R.tets.resize(4 * cap);
          R.adj.resize(4 * cap);
          MeshBuf mb{R.tets.data(), R.adj.data(), cnt, cap, {}};
          jb.rr = RepairReport{};
          jb.st = WTET_OK;
          if (!jb.viol.empty())
            jb.st = repair_delaunay(xc, nc, mb, jb.viol, a.cfg.predicate_mode, &jb.rr);
          std::vector<uint32_t> added;
          if (jb.st == WTET_OK && !missing[cl].empty()) {
            jb.st = insert_points(xc, nc, mb, missing[cl], a.cfg.predicate_mode, &jb.rr);
            // failed_verts gets the points inserted here (not exact duplicates)
            std::vector<uint32_t> still = missing_points(mb.tets, mb.count, nc);
            std::set_difference(missing[cl].begin(), missing[cl].end(), still.begin(), still.end(),
                                std::back_inserter(added));
          }
          if (jb.st == WTET_ERR_OUT_OF_MEMORY && attempt < 6) {
            cap *= 2;
            continue;
          }
          if (jb.st == WTET_OK && !missing[cl].empty()) missing[cl].swap(added);
          R.tets.resize(4 * mb.count);
          R.adj.resize(4 * mb.count);
          break;
        }
        R.repaired = true;
      }
    };
    unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    size_t nth = std::min<size_t>(std::min<size_t>(hw, 16), jobs.size());
    std::vector<std::thread> pool;
    for (size_t t = 1; t < nth; ++t) pool.emplace_back(work);
    work();
    for (auto& t : pool) t.join();
    for (const Job& jb : jobs) {
      if (jb.st != WTET_OK)
        throw Fail{jb.st, "host repair of cloud " + std::to_string(jb.cl) + " failed: " + jb.rr.message};
      c->stats.repair_regions += jb.rr.regions;
    }
  }
  for (uint32_t cl = 0; cl < C; ++cl)
    if (!missing[cl].empty()) {
      std::vector<uint32_t> u;
      std::set_union(failed[cl].begin(), failed[cl].end(), missing[cl].begin(), missing[cl].end(),
                     std::back_inserter(u));
      failed[cl].swap(u);
    }
  auto t2 = std::chrono::steady_clock::now();
  c->stats.ms_repair = std::chrono::duration<double, std::milli>(t2 - t1).count();
  c->stats.sum_uninserted = K.sumU;
  c->stats.sum_active = K.sumA;
  c->stats.sum_flips = K.sumF;
  c->stats.flips_23 = K.flips23;
  c->stats.flips_32 = K.sumF - K.flips23;
  c->stats.small_checked = K.smallA;
  c->stats.small_flips = K.smallF;
  c->stats.rounds = K.rounds;
  c->stats.flip_iterations = K.iters;
  c->stats.peak_active = K.peakA;
  c->stats.num_points = n;
  uint64_t nfailed = 0;
  for (auto& v : failed) nfailed += v.size();

  if (resident) {
    // repaired clouds move to rows appended after the device phase's rows
    c->bstart.assign(C, 0);
    c->bcount.assign(C, 0);
    uint64_t extra = 0;
    for (uint32_t cl = 0; cl < C; ++cl)
      if (rows[cl].repaired) extra += rows[cl].tets.size() / 4;
    uint64_t total = (uint64_t)mout + extra;
    if (total >= (1ull << 32)) throw Fail{WTET_ERR_OUT_OF_MEMORY, "too many output rows"};
    if (extra) {
      ensure(c, c->prop, total * 16, true);
      ensure(c, c->fout, total * 16, true);
    }
    uint64_t tail = mout;
    for (uint32_t cl = 0; cl < C; ++cl) {
      if (rows[cl].repaired) {
        size_t cnt = rows[cl].tets.size() / 4;
        HIP_CHECK(hipMemcpyAsync(P<uint32_t>(c->prop) + 4 * tail, rows[cl].tets.data(), cnt * 16,
                                 hipMemcpyHostToDevice, c->stream));
        HIP_CHECK(hipMemcpyAsync(P<uint32_t>(c->fout) + 4 * tail, rows[cl].adj.data(), cnt * 16,
                                 hipMemcpyHostToDevice, c->stream));
        c->bstart[cl] = tail;
        c->bcount[cl] = (uint32_t)cnt;
        tail += cnt;
      } else {
        c->bstart[cl] = cstart[cl];
        c->bcount[cl] = cstart[cl + 1] - cstart[cl];
      }
    }
    HIP_CHECK(hipStreamSynchronize(c->stream));
    a.dout->num_clouds = C;
    a.dout->rows = total;
    a.dout->d_tets = P<uint32_t>(c->prop);
    a.dout->d_adjacency = P<uint32_t>(c->fout);
    a.dout->tet_start = c->bstart.data();
    a.dout->tet_count = c->bcount.data();
    a.dout->num_failed = nfailed;
    uint64_t live = 0;
    for (uint32_t cl = 0; cl < C; ++cl) live += c->bcount[cl];
    c->stats.num_tets_out = live;
    c->stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return;
  }
  // host result: clouds in order, contiguous
  wtet_batch_result* out = a.out;
  out->num_clouds = C;
  out->tet_offsets = (uint64_t*)std::malloc(((size_t)C + 1) * 8);
  out->failed_offsets = (uint64_t*)std::malloc(((size_t)C + 1) * 8);
  if (!out->tet_offsets || !out->failed_offsets) throw std::bad_alloc();
  uint64_t total = 0;
  for (uint32_t cl = 0; cl < C; ++cl) {
    out->tet_offsets[cl] = total;
    total += rows[cl].repaired ? rows[cl].tets.size() / 4 : cstart[cl + 1] - cstart[cl];
  }
  out->tet_offsets[C] = total;
  out->num_tets = total;
  out->tets = alloc_rows(total * 16);
  out->adjacency = alloc_rows(total * 16);
  out->num_failed = nfailed;
  out->failed_verts = (uint32_t*)std::malloc(std::max<size_t>(nfailed * 4, 4));
  if (!out->tets || !out->adjacency || !out->failed_verts) throw std::bad_alloc();
  uint64_t fo = 0;
  for (uint32_t cl = 0; cl < C; ++cl) {
    uint64_t r0 = out->tet_offsets[cl], cnt = out->tet_offsets[cl + 1] - r0;
    const uint32_t* st = rows[cl].repaired ? rows[cl].tets.data() : c->h_tets + 4 * (size_t)cstart[cl];
    const uint32_t* sa = rows[cl].repaired ? rows[cl].adj.data() : c->h_adj + 4 * (size_t)cstart[cl];
    std::memcpy(out->tets + 4 * r0, st, cnt * 16);
    std::memcpy(out->adjacency + 4 * r0, sa, cnt * 16);
    out->failed_offsets[cl] = fo;
    if (!failed[cl].empty()) std::memcpy(out->failed_verts + fo, failed[cl].data(), failed[cl].size() * 4);
    fo += failed[cl].size();
  }
  out->failed_offsets[C] = fo;
  c->stats.num_tets_out = total;
  c->stats.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

extern "C" {

static hipError_t resolve_scratch(int device);

wtet_status wtet_create(int device, wtet_ctx** out) {
  if (!out) return WTET_ERR_INVALID_ARGUMENT;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= device || device < 0) {
    (void)hipGetLastError();
    return WTET_ERR_DEVICE;
  }
  wtet_ctx* c = new (std::nothrow) wtet_ctx();
  if (!c) return WTET_ERR_OUT_OF_MEMORY;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess ||
      hipStreamCreate(&c->copy_stream) != hipSuccess ||
      hipHostMalloc((void**)&c->h_ctr, C_COUNT * 4, hipHostMallocDefault) != hipSuccess ||
      hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_norm, hipEventDisableTiming) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_xyzn, hipEventDisableTiming) != hipSuccess ||
      resolve_scratch(device) != hipSuccess) {
    (void)hipGetLastError();
    delete c;
    return WTET_ERR_DEVICE;
  }
  *out = c;
  return WTET_OK;
}

void wtet_destroy(wtet_ctx* c) {
  if (!c) return;
  end_stage(c);
  pin_join(c);
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
  for (auto& g : c->graphs) (void)hipGraphExecDestroy(g.exec);
  for (auto& g : c->round_graphs) (void)hipGraphExecDestroy(g.exec);
  for (DevBuf* b : c->all)
    if (b->p) (void)hipFree(b->p);
  if (c->h_ctr) (void)hipHostFree(c->h_ctr);
  if (c->h_tets) (void)hipHostFree(c->h_tets);
  if (c->h_adj) (void)hipHostFree(c->h_adj);
  if (c->h_xyzn) (void)hipHostFree(c->h_xyzn);
  if (c->h_patch) (void)hipHostFree(c->h_patch);
  for (auto& pr : c->evpool) {
    (void)hipEventDestroy(pr.first);
    (void)hipEventDestroy(pr.second);
  }
  if (c->ev0) (void)hipEventDestroy(c->ev0);
  if (c->ev1) (void)hipEventDestroy(c->ev1);
  if (c->ev_norm) (void)hipEventDestroy(c->ev_norm);
  if (c->ev_xyzn) (void)hipEventDestroy(c->ev_xyzn);
  if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
}

const char* wtet_last_error(const wtet_ctx* c) { return c ? c->err.c_str() : "null context"; }

void wtet_config_default(wtet_config* cfg) {
  if (!cfg) return;
  cfg->insertion_rule = 0;
  cfg->enable_flipping = 1;
  cfg->enable_sorting = 0;
  cfg->enable_hilbert_sorting = 0;
  cfg->enable_splaying = 1;
  cfg->max_insert_iterations = 100;
  cfg->max_flip_iterations = 10;
  cfg->predicate_mode = 0;
  cfg->collect_stats = 0;
  cfg->presize_mode = 0;
}

static wtet_status delaunay_common(wtet_ctx* c, const float* xyz, bool dev, uint32_t n,
                                   const wtet_config* cfg, wtet_result* out,
                                   wtet_device_result* dout) {
  if (!c || (!out && !dout)) return WTET_ERR_INVALID_ARGUMENT;
  if (out) std::memset(out, 0, sizeof(*out));
  if (dout) std::memset(dout, 0, sizeof(*dout));
  if (!xyz) {
    c->err = "points pointer is null";
    return WTET_ERR_INVALID_ARGUMENT;
  }
  if (n < 4) {
    c->err = "Need at least 4 points for 3D Delaunay, got " + std::to_string(n);
    return WTET_ERR_TOO_FEW_POINTS;
  }
  if (n > (1u << 26)) {
    c->err = "too many points for one call";
    return WTET_ERR_INVALID_ARGUMENT;
  }
  DelaunayArgs a;
  a.c = c;
  a.xyz = xyz;
  a.on_device = dev;
  a.n = n;
  if (cfg) a.cfg = *cfg;
  else wtet_config_default(&a.cfg);
  if (a.cfg.insertion_rule != 0 && a.cfg.insertion_rule != 1) {
    c->err = "Invalid insertion_rule. Must be 'circumcenter' or 'centroid'.";
    return WTET_ERR_INVALID_ARGUMENT;
  }
  if (a.cfg.presize_mode < 0 || a.cfg.presize_mode > 2) {
    c->err = "presize_mode must be 0 (fixed), 1 (learned) or 2 (grow)";
    return WTET_ERR_INVALID_ARGUMENT;
  }
  a.out = out;
  a.dout = dout;
  wtet_status s = run_guarded(c, delaunay_impl, &a);
  if (s != WTET_OK) {
    if (out) wtet_result_free(out);
    if (dout) std::memset(dout, 0, sizeof(*dout));
  }
  return s;
}

wtet_status wtet_delaunay(wtet_ctx* c, const float* xyz, uint32_t n, const wtet_config* cfg,
                          wtet_result* out) {
  return delaunay_common(c, xyz, false, n, cfg, out, nullptr);
}

wtet_status wtet_delaunay_device(wtet_ctx* c, const float* d_xyz, uint32_t n,
                                 const wtet_config* cfg, wtet_result* out) {
  return delaunay_common(c, d_xyz, true, n, cfg, out, nullptr);
}

wtet_status wtet_delaunay_resident(wtet_ctx* c, const float* d_xyz, uint32_t n,
                                   const wtet_config* cfg, wtet_device_result* out) {
  if (!out) return WTET_ERR_INVALID_ARGUMENT;
  return delaunay_common(c, d_xyz, true, n, cfg, nullptr, out);
}

static wtet_status batch_common(wtet_ctx* c, const float* xyz, bool dev, const uint64_t* offsets,
                                uint32_t num_clouds, const wtet_config* cfg,
                                wtet_batch_result* out, wtet_batch_device_result* dout) {
  if (!c || (!out && !dout)) return WTET_ERR_INVALID_ARGUMENT;
  if (out) std::memset(out, 0, sizeof(*out));
  if (dout) std::memset(dout, 0, sizeof(*dout));
  if (!xyz || !offsets) {
    c->err = "points or offsets pointer is null";
    return WTET_ERR_INVALID_ARGUMENT;
  }
  if (num_clouds == 0 || num_clouds > (1u << 20)) {
    c->err = "num_clouds must be in 1..2^20, got " + std::to_string(num_clouds);
    return WTET_ERR_INVALID_ARGUMENT;
  }
  if (offsets[0] != 0) {
    c->err = "offsets[0] must be 0";
    return WTET_ERR_INVALID_ARGUMENT;
  }
  if (offsets[num_clouds] > (1ull << 26)) {
    c->err = "too many points for one call (at most 2^26 over all clouds)";
    return WTET_ERR_INVALID_ARGUMENT;
  }
  BatchArgs a;
  a.off.resize((size_t)num_clouds + 1);
  for (uint32_t i = 0; i <= num_clouds; ++i) a.off[i] = (uint32_t)offsets[i];
  for (uint32_t i = 0; i < num_clouds; ++i) {
    if (offsets[i + 1] < offsets[i]) {
      c->err = "offsets must be nondecreasing";
      return WTET_ERR_INVALID_ARGUMENT;
    }
    if (offsets[i + 1] - offsets[i] < 4) {
      c->err = "Need at least 4 points for 3D Delaunay, got " + std::to_string(offsets[i + 1] - offsets[i]) +
               " in cloud " + std::to_string(i);
      return WTET_ERR_TOO_FEW_POINTS;
    }
  }
  a.c = c;
  a.xyz = xyz;
  a.on_device = dev;
  if (cfg) a.cfg = *cfg;
  else wtet_config_default(&a.cfg);
  if (a.cfg.insertion_rule != 0 && a.cfg.insertion_rule != 1) {
    c->err = "Invalid insertion_rule. Must be 'circumcenter' or 'centroid'.";
    return WTET_ERR_INVALID_ARGUMENT;
  }
  if (a.cfg.presize_mode < 0 || a.cfg.presize_mode > 2) {
    c->err = "presize_mode must be 0 (fixed), 1 (learned) or 2 (grow)";
    return WTET_ERR_INVALID_ARGUMENT;
  }
  a.out = out;
  a.dout = dout;
  wtet_status s = run_guarded(c, batch_impl, &a);
  if (s != WTET_OK) {
    if (out) wtet_batch_result_free(out);
    if (dout) std::memset(dout, 0, sizeof(*dout));
  }
  return s;
}

wtet_status wtet_delaunay_batch(wtet_ctx* c, const float* xyz, const uint64_t* offsets,
                                uint32_t num_clouds, const wtet_config* cfg, wtet_batch_result* out) {
  return batch_common(c, xyz, false, offsets, num_clouds, cfg, out, nullptr);
}

wtet_status wtet_delaunay_batch_device(wtet_ctx* c, const float* d_xyz, const uint64_t* offsets,
                                       uint32_t num_clouds, const wtet_config* cfg,
                                       wtet_batch_result* out) {
  return batch_common(c, d_xyz, true, offsets, num_clouds, cfg, out, nullptr);
}

wtet_status wtet_delaunay_batch_resident(wtet_ctx* c, const float* d_xyz, const uint64_t* offsets,
                                         uint32_t num_clouds, const wtet_config* cfg,
                                         wtet_batch_device_result* out) {
  if (!out) return WTET_ERR_INVALID_ARGUMENT;
  return batch_common(c, d_xyz, true, offsets, num_clouds, cfg, nullptr, out);
}

void wtet_batch_result_free(wtet_batch_result* r) {
  if (!r) return;
  std::free(r->tets);
  std::free(r->adjacency);
  std::free(r->tet_offsets);
  std::free(r->failed_verts);
  std::free(r->failed_offsets);
  std::memset(r, 0, sizeof(*r));
}

void wtet_result_free(wtet_result* r) {
  if (!r) return;
  std::free(r->tets);
  std::free(r->adjacency);
  std::free(r->failed_verts);
  std::memset(r, 0, sizeof(*r));
}

wtet_status wtet_get_stats(const wtet_ctx* c, wtet_stats* out) {
  if (!c || !out) return WTET_ERR_INVALID_ARGUMENT;
  *out = c->stats;
  return WTET_OK;
}

uint32_t wtet_encode_adjacency(uint32_t tet, uint32_t face) { return (tet << 5) | (face & 3u); }

void wtet_decode_adjacency(uint32_t packed, uint32_t* tet, uint32_t* face) {
  if (tet) *tet = packed >> 5;
  if (face) *face = packed & 3u;
}

const char* wtet_version(void) { return "0.1.0"; }

wtet_status wtet_device_upload(wtet_ctx* c, const void* host, size_t bytes, void** d_out) {
  if (!c || !host || !d_out || bytes == 0) return WTET_ERR_INVALID_ARGUMENT;
  *d_out = nullptr;
  if (hipSetDevice(c->device) != hipSuccess || hipMalloc(d_out, bytes) != hipSuccess) {
    (void)hipGetLastError();
    c->err = "device allocation failed";
    return WTET_ERR_OUT_OF_MEMORY;
  }
  if (hipMemcpy(*d_out, host, bytes, hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    (void)hipFree(*d_out);
    *d_out = nullptr;
    c->err = "host to device copy failed";
    return WTET_ERR_DEVICE;
  }
  return WTET_OK;
}

wtet_status wtet_device_download(wtet_ctx* c, const void* d_src, void* host_dst, size_t bytes) {
  if (!c || !d_src || !host_dst) return WTET_ERR_INVALID_ARGUMENT;
  if (hipSetDevice(c->device) != hipSuccess ||
      hipMemcpy(host_dst, d_src, bytes, hipMemcpyDeviceToHost) != hipSuccess) {
    (void)hipGetLastError();
    c->err = "device to host copy failed";
    return WTET_ERR_DEVICE;
  }
  return WTET_OK;
}

wtet_status wtet_device_free(wtet_ctx* c, void* d_ptr) {
  if (!c) return WTET_ERR_INVALID_ARGUMENT;
  if (d_ptr && hipFree(d_ptr) != hipSuccess) {
    (void)hipGetLastError();
    return WTET_ERR_DEVICE;
  }
  return WTET_OK;
}


// ---------------------------------------------------------------------------
// learned presizing
// ---------------------------------------------------------------------------
static_assert(kParamCount == WTET_PRESIZE_PARAMS, "parameter layout");
static_assert(kFeat == WTET_PRESIZE_FEATURES && kOut == WTET_PRESIZE_OUTPUTS, "model shape");

struct FeatArgs {
  wtet_ctx* c;
  const float* d_xyz;
  std::vector<uint32_t> off;
  float* feat;
};

static void features_impl(void* a_) {
  FeatArgs& a = *static_cast<FeatArgs*>(a_);
  wtet_ctx* c = a.c;
  HIP_CHECK(hipSetDevice(c->device));
  uint32_t C = (uint32_t)a.off.size() - 1, n = a.off[C];
  hipStream_t s = c->stream;
  ensure(c, c->doff, ((size_t)C + 1) * 4);
  ensure(c, c->mm6, (size_t)C * 24);
  ensure(c, c->feat_acc, (size_t)C * (kHistBins / 2 + 9) * 8 + 64);
  ensure(c, c->feat, (size_t)C * kFeat * 4);
  ensure(c, c->xyzn, (size_t)n * 12 + 16);
  ensure(c, c->keys, (size_t)n * 8 + 16);
  ensure(c, c->ids, (size_t)n * 4 + 16);
  HIP_CHECK(hipMemcpyAsync(c->doff.p, a.off.data(), ((size_t)C + 1) * 4, hipMemcpyHostToDevice, s));
  Clouds CL;
  CL.off = P<uint32_t>(c->doff);
  CL.C = C;
  CL.V = n + 4 * C;
  int kbits, cshift, end_bit;
  key_layout(C, false, &kbits, &cshift, &end_bit);
  CL.kbits = kbits;
  CL.cshift = cshift;
  CL.hilbert = 0;
  launch_prepare(a.d_xyz, n, CL, P<uint32_t>(c->mm6), P<float>(c->xyzn), P<uint64_t>(c->keys),
                 P<uint32_t>(c->ids), s);
  launch_features(a.d_xyz, n, CL.off, C, P<uint32_t>(c->mm6), P<double>(c->feat_acc),
                  P<float>(c->feat), s);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(a.feat, c->feat.p, (size_t)C * kFeat * 4, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// wtet_presize_features_noalloc: the same kernels with every buffer beyond
// the caller's points and features carved from one static device array.
// ---------------------------------------------------------------------------
constexpr uint32_t kNoAllocMaxClouds = 2048;
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kNaOffBytes = align256(((size_t)kNoAllocMaxClouds + 1) * 4);
constexpr size_t kNaMmBytes = align256((size_t)kNoAllocMaxClouds * 24);
constexpr size_t kNaAccBytes = align256((size_t)kNoAllocMaxClouds * (kHistBins / 2 + 9) * 8 + 64);
constexpr size_t kNaFeatBytes = align256((size_t)kNoAllocMaxClouds * kFeat * 4);
constexpr size_t kNoAllocScratchBytes = kNaOffBytes + kNaMmBytes + kNaAccBytes + kNaFeatBytes;
static_assert(kNoAllocScratchBytes == WTET_NOALLOC_SCRATCH_BYTES, "scratch size in wtet.h");
static_assert(kNoAllocMaxClouds == WTET_NOALLOC_MAX_CLOUDS, "cloud limit in wtet.h");

// One copy per GPU, held until the process ends.  In host code this name is
// a host-side stand-in, never a device address: kernels and copies use the
// per-device pointer that hipGetSymbolAddress returns.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices] = {};

// With `device` current: resolves this GPU's copy of the array once.
static hipError_t resolve_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return hipSuccess;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  g_scratch_ptr[device] = ptr;
  return hipSuccess;
}

struct FeatNoAllocArgs {
  wtet_ctx* c;
  const float* d_xyz;
  const uint64_t* offsets;
  uint32_t C;
  float* feat;
};

static void features_noalloc_impl(void* a_) {
  FeatNoAllocArgs& a = *static_cast<FeatNoAllocArgs*>(a_);
  wtet_ctx* c = a.c;
  HIP_CHECK(hipSetDevice(c->device));
  unsigned char* base = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = static_cast<unsigned char*>(g_scratch_ptr[c->device]);
  }
  if (!base) throw Fail{WTET_ERR_DEVICE, "no-alloc scratch was not resolved for this device"};
  uint32_t* doff = reinterpret_cast<uint32_t*>(base);
  uint32_t* mm = reinterpret_cast<uint32_t*>(base + kNaOffBytes);
  double* acc = reinterpret_cast<double*>(base + kNaOffBytes + kNaMmBytes);
  float* feat = reinterpret_cast<float*>(base + kNaOffBytes + kNaMmBytes + kNaAccBytes);
  const uint32_t C = a.C;
  hipStream_t s = c->stream;
  // the cloud offsets as u32 words, from a small stack buffer in chunks;
  // every chunk's copy completes before the buffer is reused
  uint32_t chunk[64];
  for (uint32_t i = 0; i <= C; i += 64) {
    uint32_t k = std::min<uint32_t>(64, C + 1 - i);
    for (uint32_t j = 0; j < k; ++j) chunk[j] = (uint32_t)a.offsets[i + j];
    HIP_CHECK(hipMemcpyAsync(doff + i, chunk, (size_t)k * 4, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
  }
  Clouds CL;
  CL.off = doff;
  CL.C = C;
  CL.V = (uint32_t)a.offsets[C] + 4 * C;
  int kbits, cshift, end_bit;
  key_layout(C, false, &kbits, &cshift, &end_bit);
  CL.kbits = kbits;
  CL.cshift = cshift;
  CL.hilbert = 0;
  const uint32_t n = (uint32_t)a.offsets[C];
  launch_minmax(a.d_xyz, n, CL, mm, s);
  launch_features(a.d_xyz, n, doff, C, mm, acc, feat, s);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(a.feat, feat, (size_t)C * kFeat * 4, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
}

struct PredictArgs {
  wtet_ctx* c;
  const float* params;
  const float* feat;
  uint32_t C;
  float* pred;
};

static void predict_impl(void* a_) {
  PredictArgs& a = *static_cast<PredictArgs*>(a_);
  wtet_ctx* c = a.c;
  HIP_CHECK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  DevBuf p, f, o;
  ensure(c, p, kParamCount * 4);
  ensure(c, f, (size_t)a.C * kFeat * 4);
  ensure(c, o, (size_t)a.C * kOut * 4);
  HIP_CHECK(hipMemcpyAsync(p.p, a.params, kParamCount * 4, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(f.p, a.feat, (size_t)a.C * kFeat * 4, hipMemcpyHostToDevice, s));
  launch_mlp_predict(P<float>(p), P<float>(f), a.C, P<float>(o), s);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(a.pred, o.p, (size_t)a.C * kOut * 4, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
  (void)hipFree(p.p);
  (void)hipFree(f.p);
  (void)hipFree(o.p);
}

// device buffers of one training / gradient call
struct MlpBufs {
  DevBuf params, grad, m, v, xraw, yraw, idx, x, y, h1, h2, out, d3, d2, d1, loss;
  std::vector<DevBuf*> all() {
    return {&params, &grad, &m, &v, &xraw, &yraw, &idx, &x, &y, &h1, &h2, &out, &d3, &d2, &d1, &loss};
  }
  ~MlpBufs() {
    for (DevBuf* b : all())
      if (b->p) (void)hipFree(b->p);
  }
};

static MlpWork mlp_alloc(wtet_ctx* c, MlpBufs& B, uint32_t rows, uint32_t batch) {
  ensure(c, B.params, kParamCount * 4);
  ensure(c, B.grad, kParamCount * 4);
  ensure(c, B.xraw, (size_t)rows * kFeat * 4);
  ensure(c, B.yraw, (size_t)rows * kOut * 4);
  ensure(c, B.idx, (size_t)std::max(rows, batch) * 4);
  ensure(c, B.x, (size_t)batch * kFeat * 4);
  ensure(c, B.y, (size_t)batch * kOut * 4);
  ensure(c, B.h1, (size_t)batch * kHidden * 4);
  ensure(c, B.h2, (size_t)batch * kHidden * 4);
  ensure(c, B.out, (size_t)batch * kOut * 4);
  ensure(c, B.d3, (size_t)batch * kOut * 4);
  ensure(c, B.d2, (size_t)batch * kHidden * 4);
  ensure(c, B.d1, (size_t)batch * kHidden * 4);
  ensure(c, B.loss, 64);
  MlpWork w;
  w.x = P<float>(B.x);
  w.y = P<float>(B.y);
  w.h1 = P<float>(B.h1);
  w.h2 = P<float>(B.h2);
  w.out = P<float>(B.out);
  w.d3 = P<float>(B.d3);
  w.d2 = P<float>(B.d2);
  w.d1 = P<float>(B.d1);
  w.grad = P<float>(B.grad);
  w.loss = P<float>(B.loss);
  return w;
}

struct GradArgs {
  wtet_ctx* c;
  const float* params;
  const float* feat;
  const float* tgt;
  uint32_t rows;
  float* loss;
  float* grad;
};

static void grad_impl(void* a_) {
  GradArgs& a = *static_cast<GradArgs*>(a_);
  wtet_ctx* c = a.c;
  HIP_CHECK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  MlpBufs B;
  MlpWork w = mlp_alloc(c, B, a.rows, a.rows);
  std::vector<uint32_t> idx(a.rows);
  for (uint32_t i = 0; i < a.rows; ++i) idx[i] = i;
  HIP_CHECK(hipMemcpyAsync(B.params.p, a.params, kParamCount * 4, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(B.xraw.p, a.feat, (size_t)a.rows * kFeat * 4, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(B.yraw.p, a.tgt, (size_t)a.rows * kOut * 4, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(B.idx.p, idx.data(), (size_t)a.rows * 4, hipMemcpyHostToDevice, s));
  launch_fill_u32(P<uint32_t>(B.grad), 0u, kParamCount, s);
  const float* Pp = P<float>(B.params);
  launch_gather_standardize(P<float>(B.xraw), P<uint32_t>(B.idx), a.rows, kFeat, Pp + kOffFeatMean,
                            Pp + kOffFeatStd, w.x, s);
  launch_gather_standardize(P<float>(B.yraw), P<uint32_t>(B.idx), a.rows, kOut, Pp + kOffTgtMean,
                            Pp + kOffTgtStd, w.y, s);
  launch_mlp_forward_backward(Pp, w, a.rows, s);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipMemcpyAsync(a.grad, B.grad.p, kParamCount * 4, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipMemcpyAsync(a.loss, B.loss.p, 4, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
}

struct TrainArgs {
  wtet_ctx* c;
  float* params;
  const float* feat;
  const float* tgt;
  uint32_t rows, epochs, batch;
  float lr;
  uint64_t seed;
  float* curve;
};

static void train_impl(void* a_) {
  TrainArgs& a = *static_cast<TrainArgs*>(a_);
  wtet_ctx* c = a.c;
  HIP_CHECK(hipSetDevice(c->device));
  hipStream_t s = c->stream;
  MlpBufs B;
  MlpWork w = mlp_alloc(c, B, a.rows, a.batch);
  ensure(c, B.m, kParamCount * 4);
  ensure(c, B.v, kParamCount * 4);
  HIP_CHECK(hipMemcpyAsync(B.params.p, a.params, kParamCount * 4, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(B.xraw.p, a.feat, (size_t)a.rows * kFeat * 4, hipMemcpyHostToDevice, s));
  HIP_CHECK(hipMemcpyAsync(B.yraw.p, a.tgt, (size_t)a.rows * kOut * 4, hipMemcpyHostToDevice, s));
  launch_fill_u32(P<uint32_t>(B.m), 0u, kParamCount, s);
  launch_fill_u32(P<uint32_t>(B.v), 0u, kParamCount, s);
  launch_fill_u32(P<uint32_t>(B.grad), 0u, kParamCount, s);
  float* Pp = P<float>(B.params);
  std::vector<uint32_t> perm(a.rows);
  std::vector<float> losses;
  uint64_t rng = a.seed * 0x9E3779B97F4A7C15ull + 1;
  auto next = [&]() {
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return rng;
  };
  uint32_t nb = a.rows / a.batch;
  ensure(c, B.loss, (size_t)std::max<uint32_t>(nb, 1) * 4 + 64);
  uint32_t step = 0;
  for (uint32_t e = 0; e < a.epochs; ++e) {
    for (uint32_t i = 0; i < a.rows; ++i) perm[i] = i;
    for (uint32_t i = a.rows - 1; i > 0; --i) std::swap(perm[i], perm[next() % (i + 1)]);
    HIP_CHECK(hipMemcpyAsync(B.idx.p, perm.data(), (size_t)a.rows * 4, hipMemcpyHostToDevice, s));
    for (uint32_t b = 0; b < nb; ++b) {
      const uint32_t* bi = P<uint32_t>(B.idx) + (size_t)b * a.batch;
      launch_gather_standardize(P<float>(B.xraw), bi, a.batch, kFeat, Pp + kOffFeatMean,
                                Pp + kOffFeatStd, w.x, s);
      launch_gather_standardize(P<float>(B.yraw), bi, a.batch, kOut, Pp + kOffTgtMean,
                                Pp + kOffTgtStd, w.y, s);
      w.loss = P<float>(B.loss) + b;  // one slot per minibatch of this epoch
      launch_mlp_forward_backward(Pp, w, a.batch, s);
      launch_adam(Pp, w.grad, P<float>(B.m), P<float>(B.v), a.lr, 0.9f, 0.999f, 1e-8f, ++step, s);
    }
    HIP_CHECK(hipGetLastError());
    if (a.curve && nb) {
      losses.resize(nb);
      HIP_CHECK(hipMemcpyAsync(losses.data(), B.loss.p, (size_t)nb * 4, hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      double sum = 0;
      for (float l : losses) sum += l;
      a.curve[e] = (float)(sum / nb);
    }
  }
  HIP_CHECK(hipMemcpyAsync(a.params, B.params.p, kParamCount * 4, hipMemcpyDeviceToHost, s));
  HIP_CHECK(hipStreamSynchronize(s));
}

static std::vector<uint32_t> offsets32(const uint64_t* offsets, uint32_t C) {
  std::vector<uint32_t> off((size_t)C + 1);
  for (uint32_t i = 0; i <= C; ++i) off[i] = (uint32_t)offsets[i];
  return off;
}

wtet_status wtet_presize_features(wtet_ctx* c, const float* d_xyz, const uint64_t* offsets,
                                  uint32_t num_clouds, float* feat) {
  if (!c || !d_xyz || !offsets || !feat || num_clouds == 0 || offsets[0] != 0 ||
      offsets[num_clouds] > (1ull << 26))
    return WTET_ERR_INVALID_ARGUMENT;
  for (uint32_t i = 0; i < num_clouds; ++i)
    if (offsets[i + 1] <= offsets[i]) return WTET_ERR_INVALID_ARGUMENT;
  FeatArgs a{c, d_xyz, offsets32(offsets, num_clouds), feat};
  return run_guarded(c, features_impl, &a);
}

wtet_status wtet_presize_features_noalloc(wtet_ctx* c, const float* d_xyz, const uint64_t* offsets,
                                          uint32_t num_clouds, float* feat) {
  if (!c) return WTET_ERR_INVALID_ARGUMENT;
  if (!d_xyz || !offsets || !feat || num_clouds == 0 || offsets[0] != 0 ||
      offsets[num_clouds] > (1ull << 26)) {
    c->err = "wtet_presize_features_noalloc: null pointer, no cloud, offsets[0] != 0 or more than 2^26 points";
    return WTET_ERR_INVALID_ARGUMENT;
  }
  if (num_clouds > kNoAllocMaxClouds) {
    c->err = "wtet_presize_features_noalloc: at most " + std::to_string(kNoAllocMaxClouds) +
             " clouds per call (got " + std::to_string(num_clouds) + ")";
    return WTET_ERR_INVALID_ARGUMENT;
  }
  for (uint32_t i = 0; i < num_clouds; ++i)
    if (offsets[i + 1] <= offsets[i]) {
      c->err = "wtet_presize_features_noalloc: every cloud needs at least one point";
      return WTET_ERR_INVALID_ARGUMENT;
    }
  FeatNoAllocArgs a{c, d_xyz, offsets, num_clouds, feat};
  return run_guarded(c, features_noalloc_impl, &a);
}

wtet_status wtet_presize_predict(wtet_ctx* c, const float* params, const float* feat,
                                 uint32_t num_clouds, float* pred) {
  if (!c || !params || !feat || !pred || num_clouds == 0) return WTET_ERR_INVALID_ARGUMENT;
  PredictArgs a{c, params, feat, num_clouds, pred};
  return run_guarded(c, predict_impl, &a);
}

struct SetModelArgs {
  wtet_ctx* c;
  const float* params;
};

static void set_model_impl(void* a_) {
  SetModelArgs& a = *static_cast<SetModelArgs*>(a_);
  HIP_CHECK(hipSetDevice(a.c->device));
  ensure(a.c, a.c->presize_params, kParamCount * 4);
  HIP_CHECK(hipMemcpy(a.c->presize_params.p, a.params, kParamCount * 4, hipMemcpyHostToDevice));
  a.c->presize_loaded = true;
}

wtet_status wtet_set_presize_model(wtet_ctx* c, const float* params) {
  if (!c) return WTET_ERR_INVALID_ARGUMENT;
  if (!params) {
    c->presize_loaded = false;
    return WTET_OK;
  }
  for (int k = 0; k < kFeat; ++k)
    if (!(params[kOffFeatStd + k] > 0.0f)) {
      c->err = "presize model: feature std must be positive";
      return WTET_ERR_INVALID_ARGUMENT;
    }
  SetModelArgs a{c, params};
  return run_guarded(c, set_model_impl, &a);
}

wtet_status wtet_mlp_loss_grad(wtet_ctx* c, const float* params, const float* feat,
                               const float* targets, uint32_t rows, float* loss, float* grad) {
  if (!c || !params || !feat || !targets || !loss || !grad || rows == 0 || rows > 4096)
    return WTET_ERR_INVALID_ARGUMENT;
  GradArgs a{c, params, feat, targets, rows, loss, grad};
  return run_guarded(c, grad_impl, &a);
}

wtet_status wtet_mlp_train(wtet_ctx* c, float* params, const float* feat, const float* targets,
                           uint32_t rows, uint32_t epochs, uint32_t batch, float lr, uint64_t seed,
                           float* loss_curve) {
  if (!c || !params || !feat || !targets || batch == 0 || batch > 4096 || rows < batch ||
      !(lr > 0.0f))
    return WTET_ERR_INVALID_ARGUMENT;
  TrainArgs a{c, params, feat, targets, rows, epochs, batch, lr, seed, loss_curve};
  return run_guarded(c, train_impl, &a);
}

wtet_status wtet_device_copy(wtet_ctx* c, void* d_dst, const void* d_src, size_t bytes) {
  if (!c || !d_dst || !d_src || bytes % 16 || ((uintptr_t)d_dst | (uintptr_t)d_src) % 16)
    return WTET_ERR_INVALID_ARGUMENT;
  if (hipSetDevice(c->device) != hipSuccess) return WTET_ERR_DEVICE;
  launch_copy16(static_cast<uint4*>(d_dst), static_cast<const uint4*>(d_src), bytes / 16, c->stream);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) {
    c->err = "device copy failed";
    return WTET_ERR_DEVICE;
  }
  return WTET_OK;
}

struct SortArgs {
  wtet_ctx* c;
  const void* keys;
  const uint32_t* vals;
  void* keys_out;
  uint32_t* vals_out;
  uint32_t n;
  uint32_t key_bits;
  int end_bit;
};

static void sort_impl(void* a_) {
  SortArgs& a = *static_cast<SortArgs*>(a_);
  wtet_ctx* c = a.c;
  HIP_CHECK(hipSetDevice(c->device));
  size_t tb = a.key_bits == 64 ? radix_temp_bytes_u64(a.n, a.end_bit) : radix_temp_bytes_u32(a.n, a.end_bit);
  ensure(c, c->sort_tmp, std::max<size_t>(tb, 16));
  if (a.key_bits == 64)
    HIP_CHECK(radix_sort_u64(static_cast<const uint64_t*>(a.keys), a.vals, static_cast<uint64_t*>(a.keys_out),
                             a.vals_out, a.n, a.end_bit, c->sort_tmp.p, c->sort_tmp.bytes, c->stream));
  else
    HIP_CHECK(radix_sort_u32(static_cast<const uint32_t*>(a.keys), a.vals, static_cast<uint32_t*>(a.keys_out),
                             a.vals_out, a.n, a.end_bit, c->sort_tmp.p, c->sort_tmp.bytes, c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
}

wtet_status wtet_device_sort_pairs(wtet_ctx* c, const void* d_keys, const uint32_t* d_vals, void* d_keys_out,
                                   uint32_t* d_vals_out, uint32_t n, uint32_t key_bits, int end_bit) {
  if (!c) return WTET_ERR_INVALID_ARGUMENT;
  if ((key_bits != 32 && key_bits != 64) || end_bit < 0 || end_bit > (int)key_bits ||
      (n && (!d_keys || !d_vals || !d_keys_out || !d_vals_out))) {
    c->err = "wtet_device_sort_pairs: key_bits must be 32 or 64, end_bit in 0..key_bits";
    return WTET_ERR_INVALID_ARGUMENT;
  }
  if (n == 0) return WTET_OK;
  SortArgs a{c, d_keys, d_vals, d_keys_out, d_vals_out, n, key_bits, end_bit};
  return run_guarded(c, sort_impl, &a);
}

wtet_status wtet_set_audit_records(wtet_ctx* c, uint32_t records) {
  if (!c || records < 64) return WTET_ERR_INVALID_ARGUMENT;
  c->audit_records = records;
  return WTET_OK;
}

wtet_status wtet_set_flip_graphs(wtet_ctx* c, int on) {
  if (!c) return WTET_ERR_INVALID_ARGUMENT;
  c->use_graphs = on != 0;
  return WTET_OK;
}

wtet_status wtet_set_flip_items(wtet_ctx* c, uint32_t items) {
  if (!c || items == 0 || items > 64) return WTET_ERR_INVALID_ARGUMENT;
  c->items = items;
  return WTET_OK;
}

wtet_status wtet_get_flip_items(const wtet_ctx* c, uint32_t* items, uint32_t* graph_captures) {
  if (!c) return WTET_ERR_INVALID_ARGUMENT;
  if (items) *items = c->items;
  if (graph_captures) *graph_captures = c->graph_captures;
  return WTET_OK;
}

wtet_status wtet_autotune(wtet_ctx* c, const float* d_xyz, const uint64_t* offsets,
                          uint32_t num_clouds, const wtet_config* cfg, uint32_t* best_items,
                          double* best_ms) {
  if (!c || !d_xyz || !offsets || num_clouds == 0) return WTET_ERR_INVALID_ARGUMENT;
  static const uint32_t cand[] = {1, 2, 4, 8};
  uint32_t keep = c->items, best = keep;
  double bms = 1e300;
  for (uint32_t g : cand) {
    c->items = g;
    wtet_batch_device_result r;
    // the first call at a setting launches its new graph keys directly and
    // the second captures them; the median of the next three is its time
    // (one run swings by a few percent)
    double ms[3] = {0, 0, 0};
    for (int rep = 0; rep < 5; ++rep) {
      wtet_status st = wtet_delaunay_batch_resident(c, d_xyz, offsets, num_clouds, cfg, &r);
      if (st != WTET_OK) {
        c->items = keep;
        return st;
      }
      if (rep > 1) ms[rep - 2] = c->stats.ms_device;
    }
    std::sort(ms, ms + 3);
    if (ms[1] < bms) {
      bms = ms[1];
      best = g;
    }
  }
  c->items = best;
  if (best_items) *best_items = best;
  if (best_ms) *best_ms = bms;
  return WTET_OK;
}

uint32_t wtet_initial_stride(uint32_t n) { return initial_stride(n); }

wtet_status wtet_stratified_ranks(uint32_t n, uint32_t stride, uint32_t* rank) {
  if (!rank || n == 0) return WTET_ERR_INVALID_ARGUMENT;
  if (stride == 0) stride = initial_stride(n);
  if (stride & (stride - 1)) return WTET_ERR_INVALID_ARGUMENT;
  for (uint32_t i = 0; i < n; ++i) rank[i] = stratified_rank(i, n, stride);
  return WTET_OK;
}

wtet_status wtet_hilbert_index(const uint32_t* xyz, uint32_t n, uint32_t* out) {
  if (!xyz || !out) return WTET_ERR_INVALID_ARGUMENT;
  for (uint32_t i = 0; i < n; ++i) out[i] = hilbert3_8(xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]);
  return WTET_OK;
}

wtet_status wtet_set_repair_rings(wtet_ctx* c, int rings) {
  if (!c || rings < 1 || rings > kRepairRingsMax) return WTET_ERR_INVALID_ARGUMENT;
  c->repair_rings = rings;
  return WTET_OK;
}

wtet_status wtet_set_repair_region_cap(wtet_ctx* c, uint32_t tets) {
  if (!c) return WTET_ERR_INVALID_ARGUMENT;
  c->region_cap = tets;
  return WTET_OK;
}

wtet_status wtet_set_insertion_warmup(wtet_ctx* c, uint32_t rounds) {
  if (!c) return WTET_ERR_INVALID_ARGUMENT;
  c->ring0_rounds = rounds;
  return WTET_OK;
}

wtet_status wtet_set_insertion_ring(wtet_ctx* c, int ring) {
  if (!c || ring < 0 || ring > 2) return WTET_ERR_INVALID_ARGUMENT;
  c->ring = ring;
  return WTET_OK;
}

wtet_status wtet_set_profiling(wtet_ctx* c, int on) {
  if (!c) return WTET_ERR_INVALID_ARGUMENT;
  c->profiling = on != 0;
  std::memset(c->ktime, 0, sizeof(c->ktime));
  std::memset(c->kcount, 0, sizeof(c->kcount));
  return WTET_OK;
}

wtet_status wtet_kernel_times(const wtet_ctx* c, uint32_t kind, const char** name,
                              uint32_t* launches, double* total_ms) {
  if (!c || kind >= (uint32_t)K_NKINDS) return WTET_ERR_INVALID_ARGUMENT;
  if (name) *name = kKernelNames[kind];
  if (launches) *launches = c->kcount[kind];
  if (total_ms) *total_ms = c->ktime[kind];
  return WTET_OK;
}

wtet_status wtet_device_info(wtet_ctx* c, char* buf, uint32_t len) {
  if (!buf || len == 0) return WTET_ERR_INVALID_ARGUMENT;
  hipDeviceProp_t p;
  int dev = c ? c->device : 0;
  if (hipGetDeviceProperties(&p, dev) != hipSuccess) {
    (void)hipGetLastError();
    std::snprintf(buf, len, "no HIP device");
    return WTET_ERR_DEVICE;
  }
  std::snprintf(buf, len, "GPU Device: %s (%s)\nCompute units: %d\nGlobal memory: %.1f GiB", p.name,
                p.gcnArchName, p.multiProcessorCount, p.totalGlobalMem / 1073741824.0);
  return WTET_OK;
}

wtet_status wtet_predicates_host(int kind, const float* pts, const uint32_t* ids, uint32_t n,
                                 int mode, int32_t* out) {
  if (!pts || !out || (kind < 2 && !ids) || kind < 0 || kind > 3) return WTET_ERR_INVALID_ARGUMENT;
  for (uint32_t i = 0; i < n; ++i) {
    if (kind == 0) {
      const float* q = pts + 12 * (size_t)i;
      const uint32_t* k = ids + 4 * (size_t)i;
      out[i] = orient3d_sos(q, q + 3, q + 6, q + 9, k[0], k[1], k[2], k[3], mode, nullptr);
    } else if (kind == 1) {
      const float* q = pts + 15 * (size_t)i;
      const uint32_t* k = ids + 5 * (size_t)i;
      out[i] = orient4d(q, q + 3, q + 6, q + 9, q + 12, k[0], k[1], k[2], k[3], k[4], mode, nullptr);
    } else if (kind == 2) {
      const float* q = pts + 12 * (size_t)i;
      out[i] = orient3d_raw(q, q + 3, q + 6, q + 9, mode, nullptr);
    } else {
      const float* q = pts + 15 * (size_t)i;
      out[i] = insphere_raw(q, q + 3, q + 6, q + 9, q + 12, mode, nullptr);
    }
  }
  return WTET_OK;
}

struct PredArgs {
  wtet_ctx* c;
  int kind;
  const float* pts;
  const uint32_t* ids;
  uint32_t n;
  int mode;
  int32_t* out;
  uint64_t* stats8;
};

static void pred_device_impl(void* a_) {
  PredArgs& a = *static_cast<PredArgs*>(a_);
  wtet_ctx* c = a.c;
  HIP_CHECK(hipSetDevice(c->device));
  size_t per = (a.kind == 0 || a.kind == 2) ? 4 : 5;
  float* dp = nullptr;
  uint32_t* di = nullptr;
  int32_t* dout = nullptr;
  uint64_t* dst = nullptr;
  HIP_CHECK(hipMalloc(&dp, std::max<size_t>(1, a.n) * per * 12));
  HIP_CHECK(hipMalloc(&di, std::max<size_t>(1, a.n) * per * 4));
  HIP_CHECK(hipMalloc(&dout, std::max<size_t>(1, a.n) * 4));
  HIP_CHECK(hipMalloc(&dst, 64));
  HIP_CHECK(hipMemcpy(dp, a.pts, (size_t)a.n * per * 12, hipMemcpyHostToDevice));
  if (a.ids) HIP_CHECK(hipMemcpy(di, a.ids, (size_t)a.n * per * 4, hipMemcpyHostToDevice));
  launch_fill_u64(dst, 0ull, 8, c->stream);
  launch_pred_batch(a.kind, dp, di, a.n, a.mode, dout, dst, c->stream);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c->stream));
  HIP_CHECK(hipMemcpy(a.out, dout, (size_t)a.n * 4, hipMemcpyDeviceToHost));
  if (a.stats8) HIP_CHECK(hipMemcpy(a.stats8, dst, 64, hipMemcpyDeviceToHost));
  (void)hipFree(dp);
  (void)hipFree(di);
  (void)hipFree(dout);
  (void)hipFree(dst);
}

wtet_status wtet_predicates_device(wtet_ctx* c, int kind, const float* pts, const uint32_t* ids,
                                   uint32_t n, int mode, int32_t* out, uint64_t* stats8) {
  if (!c || !pts || !out || (kind < 2 && !ids) || kind < 0 || kind > 3)
    return WTET_ERR_INVALID_ARGUMENT;
  PredArgs a{c, kind, pts, ids, n, mode, out, stats8};
  return run_guarded(c, pred_device_impl, &a);
}

}  // extern "C"
namespace {

template <typename Fn>
wtet_status guarded(wtet_ctx* c, Fn&& fn) {
  struct A {
    Fn* f;
  } a{&fn};
  return run_guarded(c, [](void* p) { (*static_cast<A*>(p)->f)(); }, &a);
}

wtet_stage_session* session_of(wtet_ctx* c, int phase, const char* what) {
  wtet_stage_session* S = c->stage;
  HIP_CHECK(hipSetDevice(c->device));
  if (!S) throw Fail{WTET_ERR_INVALID_ARGUMENT, std::string(what) + ": no stage session (wtet_stage_init)"};
  if (phase && S->phase != phase) {
    static const char* next[] = {"", "vote", "split", "split_points"};
    throw Fail{WTET_ERR_INVALID_ARGUMENT,
               std::string(what) + " out of order: " + next[S->phase] + " comes next"};
  }
  return S;
}

void stage_info(wtet_ctx* c, const wtet_stage_session* S, wtet_stage_info* info) {
  if (!info) return;
  read_ctr(c);
  info->round = S->R.round;
  info->uninserted = S->R.U;
  info->winners = S->R.W;
  info->tet_slots = c->h_ctr[C_TETS];
  info->active = c->h_ctr[C_A];
  info->checked = S->checked;
  info->flips = S->flips;
  info->violations = S->violations;
  info->deferred = S->deferred;
}

// slots holding tets: between a vote and its split the slots the vote
// reserved (from ctr[C_BASE]) are not written yet
uint32_t written_slots(wtet_ctx* c, const wtet_stage_session* S) {
  read_ctr(c);
  return S->phase == 2 ? c->h_ctr[C_BASE] : c->h_ctr[C_TETS];
}

}  // namespace

wtet_status wtet_stage_init(wtet_ctx* c, const float* xyz, uint32_t n, const wtet_config* cfg,
                            wtet_stage_info* info) {
  if (!c) return WTET_ERR_INVALID_ARGUMENT;
  end_stage(c);
  if (!xyz) {
    c->err = "points pointer is null";
    return WTET_ERR_INVALID_ARGUMENT;
  }
  if (n < 4) {
    c->err = "Need at least 4 points for 3D Delaunay, got " + std::to_string(n);
    return WTET_ERR_TOO_FEW_POINTS;
  }
  if (n > (1u << 26)) {
    c->err = "too many points for one call";
    return WTET_ERR_INVALID_ARGUMENT;
  }
  wtet_config kc;
  if (cfg) kc = *cfg;
  else wtet_config_default(&kc);
  if (kc.insertion_rule != 0 && kc.insertion_rule != 1) {
    c->err = "Invalid insertion_rule. Must be 'circumcenter' or 'centroid'.";
    return WTET_ERR_INVALID_ARGUMENT;
  }
  if (kc.presize_mode < 0 || kc.presize_mode > 2 || kc.predicate_mode < 0 || kc.predicate_mode > 1) {
    c->err = "presize_mode must be 0, 1 or 2 and predicate_mode 0 or 1";
    return WTET_ERR_INVALID_ARGUMENT;
  }
  wtet_status st = guarded(c, [&] {
    HIP_CHECK(hipSetDevice(c->device));
    std::memset(&c->stats, 0, sizeof(c->stats));
    ensure(c, c->xyz_in, (size_t)n * 12);
    HIP_CHECK(hipMemcpyAsync(c->xyz_in.p, xyz, (size_t)n * 12, hipMemcpyHostToDevice, c->stream));
    auto* S = new wtet_stage_session();
    c->stage = S;
    run_setup(c, S->R, {0u, n}, kc);
    dev_prepare(c, S->R, P<float>(c->xyz_in));
    S->phase = 1;
    stage_info(c, S, info);
  });
  if (st != WTET_OK) end_stage(c);
  return st;
}

wtet_status wtet_stage_vote(wtet_ctx* c, wtet_stage_info* info) {
  if (!c) return WTET_ERR_INVALID_ARGUMENT;
  return guarded(c, [&] {
    wtet_stage_session* S = session_of(c, 1, "vote");
    if (S->stale) throw Fail{WTET_ERR_INVALID_ARGUMENT, "vote out of order: relocate after flipping"};
    S->checked = S->flips = 0;
    if (S->R.U == 0) {
      S->R.W = 0;  // every point is in: nothing to vote on
    } else {
      dev_vote(c, S->R);
      S->phase = 2;
    }
    stage_info(c, S, info);
  });
}

wtet_status wtet_stage_split(wtet_ctx* c, wtet_stage_info* info) {
  if (!c) return WTET_ERR_INVALID_ARGUMENT;
  return guarded(c, [&] {
    wtet_stage_session* S = session_of(c, 2, "split");
    dev_split(c, S->R);
    S->phase = 3;
    S->audit_valid = false;
    stage_info(c, S, info);
  });
}

wtet_status wtet_stage_split_points(wtet_ctx* c, wtet_stage_info* info) {
  if (!c) return WTET_ERR_INVALID_ARGUMENT;
  return guarded(c, [&] {
    wtet_stage_session* S = session_of(c, 3, "split_points");
    dev_split_points(c, S->R);
    S->phase = 1;
    stage_info(c, S, info);
  });
}

wtet_status wtet_stage_flip(wtet_ctx* c, int predicate_mode, wtet_stage_info* info) {
  if (!c) return WTET_ERR_INVALID_ARGUMENT;
  if (predicate_mode < -1 || predicate_mode > 1) {
    c->err = "predicate_mode must be -1, 0 or 1";
    return WTET_ERR_INVALID_ARGUMENT;
  }
  return guarded(c, [&] {
    wtet_stage_session* S = session_of(c, 1, "flip");
    Run& R = S->R;
    int mode = predicate_mode < 0 ? R.mode : (predicate_mode == 1 ? PRED_EXACT_ONLY : PRED_FILTERED);
    read_ctr(c);
    if (c->h_ctr[C_STOP]) grow_for(c, (uint64_t)c->h_ctr[C_TETS] + c->h_ctr[C_A] + 64);
    uint32_t a = c->h_ctr[C_A];
    uint64_t before[S_COUNT] = {}, after[S_COUNT] = {};
    HIP_CHECK(hipMemcpyAsync(before, c->isums.p, S_COUNT * 8, hipMemcpyDeviceToHost, c->stream));
    if (a > 0) run_flip_batch(c, R.V, mode, R.st, 1, a, a / 8 + 1);
    HIP_CHECK(hipMemcpyAsync(after, c->isums.p, S_COUNT * 8, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    S->checked = (uint32_t)(after[0] - before[0]);
    S->flips = (uint32_t)(after[1] - before[1]);
    read_ctr(c);
    S->deferred = a > 0 ? c->h_ctr[C_SLOW_T] : 0u;
    if (S->flips) {
      S->audit_valid = false;
      if (R.U > 0) S->stale = true;
    }
    stage_info(c, S, info);
  });
}

wtet_status wtet_stage_relocate(wtet_ctx* c, wtet_stage_info* info) {
  if (!c) return WTET_ERR_INVALID_ARGUMENT;
  return guarded(c, [&] {
    wtet_stage_session* S = session_of(c, 1, "relocate");
    dev_relocate(c, S->R);
    read_ctr(c);
    if (c->h_ctr[C_WALK_FAIL]) throw Fail{WTET_ERR_INTERNAL, "point relocation walk did not terminate"};
    S->stale = false;
    stage_info(c, S, info);
  });
}

wtet_status wtet_stage_gather(wtet_ctx* c, wtet_stage_info* info) {
  if (!c) return WTET_ERR_INVALID_ARGUMENT;
  return guarded(c, [&] {
    wtet_stage_session* S = session_of(c, 0, "gather");
    uint32_t T = written_slots(c, S);
    Mesh M = mesh_of(c, S->R.V, S->R.mode);
    launch_set_ctr(M.ctr, C_VIOL, 0u, C_FAILREC, 0u, c->stream);
    const uint32_t rec_cap = (uint32_t)std::min<size_t>(c->rec.bytes / 4, c->audit_records);
    timed(c, K_AUDIT, [&] {
      launch_audit(M, T, P<uint32_t>(c->rec), rec_cap, S->R.st, P<uint32_t>(c->slow_t), c->stream);
    });
    HIP_CHECK(hipGetLastError());
    read_ctr(c);
    if (c->h_ctr[C_FAILREC] > rec_cap) reaudit(c, S->R, T);
    S->violations = c->h_ctr[C_VIOL];
    S->nrec = c->h_ctr[C_FAILREC];
    S->audit_valid = true;
    stage_info(c, S, info);
  });
}

wtet_status wtet_stage_readback(wtet_ctx* c, wtet_result* out) {
  if (!c || !out) return WTET_ERR_INVALID_ARGUMENT;
  std::memset(out, 0, sizeof(*out));
  wtet_status st = guarded(c, [&] {
    wtet_stage_session* S = session_of(c, 0, "readback");
    hipStream_t s = c->stream;
    uint32_t T = written_slots(c, S);
    // the flip lists stay as they are: the rows go through ord[0]
    dev_compact(c, S->R, T, true, true);
    read_ctr(c);
    uint32_t m = c->h_ctr[C_ALIVE];
    std::vector<uint32_t> tets((size_t)m * 4), adj((size_t)m * 4), rec;
    HIP_CHECK(hipMemcpyAsync(tets.data(), c->prop.p, (size_t)m * 16, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(adj.data(), c->fout.p, (size_t)m * 16, hipMemcpyDeviceToHost, s));
    if (S->audit_valid && S->nrec) {
      // records are slot-based; map a copy of them to rows
      std::vector<uint32_t> slot2row(T);
      rec.resize(S->nrec);
      HIP_CHECK(hipMemcpyAsync(slot2row.data(), c->slot2idx.p, (size_t)T * 4, hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipMemcpyAsync(rec.data(), c->rec.p, (size_t)S->nrec * 4, hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      for (uint32_t& r : rec) r = (slot2row[r >> 2] << 2) | (r & 3u);
    }
    HIP_CHECK(hipStreamSynchronize(s));
    std::vector<uint32_t> failed = failed_vertices(tets.data(), adj.data(), rec, S->R.n);
    out->num_tets = m;
    out->tets = alloc_rows(tets.size() * 4);
    out->adjacency = alloc_rows(adj.size() * 4);
    out->num_failed = (uint32_t)failed.size();
    out->failed_verts = (uint32_t*)std::malloc(std::max<size_t>(failed.size(), 1) * 4);
    if (!out->tets || !out->adjacency || !out->failed_verts) throw std::bad_alloc();
    if (m) {
      std::memcpy(out->tets, tets.data(), tets.size() * 4);
      std::memcpy(out->adjacency, adj.data(), adj.size() * 4);
    }
    if (!failed.empty()) std::memcpy(out->failed_verts, failed.data(), failed.size() * 4);
  });
  if (st != WTET_OK) wtet_result_free(out);
  return st;
}

wtet_status wtet_stage_locations(wtet_ctx* c, uint32_t cap, uint32_t* point_ids, uint32_t* rows,
                                 uint32_t* count) {
  if (!c || !count || (cap && (!point_ids || !rows))) return WTET_ERR_INVALID_ARGUMENT;
  return guarded(c, [&] {
    wtet_stage_session* S = session_of(c, 1, "locations");
    if (S->stale) throw Fail{WTET_ERR_INVALID_ARGUMENT, "locations out of order: relocate after flipping"};
    Run& R = S->R;
    hipStream_t s = c->stream;
    *count = R.U;
    uint32_t k = std::min(cap, R.U);
    if (k == 0) return;
    uint32_t T = written_slots(c, S);
    dev_compact(c, R, T, false, true);  // slot -> row, as a readback numbers them
    std::vector<uint32_t> uv(k), ut(k), slot2row(T);
    std::vector<float4> pts((size_t)R.V + 1);
    HIP_CHECK(hipMemcpyAsync(uv.data(), c->uvert[R.cur].p, (size_t)k * 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(ut.data(), c->utet[R.cur].p, (size_t)k * 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(slot2row.data(), c->slot2idx.p, (size_t)T * 4, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(pts.data(), c->pts.p, pts.size() * 16, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    for (uint32_t i = 0; i < k; ++i) {
      if (uv[i] >= R.V || ut[i] >= T) throw Fail{WTET_ERR_INTERNAL, "uninserted list entry out of range"};
      point_ids[i] = pt_rank(pts[uv[i]]);
      rows[i] = slot2row[ut[i]];
    }
  });
}

wtet_status wtet_stage_finish(wtet_ctx* c, wtet_result* out) {
  if (!c || !out) return WTET_ERR_INVALID_ARGUMENT;
  std::memset(out, 0, sizeof(*out));
  wtet_status st = guarded(c, [&] {
    auto t0 = std::chrono::steady_clock::now();
    wtet_stage_session* S = session_of(c, 1, "finish");
    Run& R = S->R;
    if (R.U > 0)
      throw Fail{WTET_ERR_INVALID_ARGUMENT, "finish: " + std::to_string(R.U) + " points are not inserted yet"};
    if (R.flipping) dev_finish_flips(c, R);
    std::vector<uint32_t> rec, cstart;
    uint32_t nviol = 0, mout = 0;
    dev_output(c, R, rec, nviol, false, mout, cstart);
    Counters K = R.K;
    uint32_t n = R.n;
    wtet_config cfg = R.cfg;
    end_stage(c);
    finish_single(c, n, cfg, K, rec, nviol, mout, out, nullptr, t0, 0u);
  });
  if (st != WTET_OK) {
    wtet_result_free(out);
    if (st != WTET_ERR_INVALID_ARGUMENT) end_stage(c);  // a misuse keeps the session
  }
  return st;
}

wtet_status wtet_stage_end(wtet_ctx* c) {
  if (!c) return WTET_ERR_INVALID_ARGUMENT;
  end_stage(c);
  return WTET_OK;
}
