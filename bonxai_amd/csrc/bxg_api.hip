// bxg_api.hip -- host runtime of the MI355X sparse voxel map behind the C ABI
// declared in include/bonxai_gpu.h.
//
// A context owns one HIP stream (default priority), the scratch of the scan
// and bulk-op pipelines, pinned staging, events and the memory budget of its
// maps.  A map owns its device node pools (root directory + leaf pool) and a
// device header (MapDev) that every kernel reads, so pools can grow between
// calls without re-capturing graphs.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "bxg_kernels.hip"
#include "../../include/bonxai_gpu.h"

using namespace bxg;

namespace {

constexpr uint32_t kBlock = 256;
constexpr int kStageSlots = 8;
constexpr size_t kMaxPendingInserts = 256;  // asynchronous inserts a map holds before it is confirmed

struct PinnedSlot {
  void* host = nullptr;
  uint64_t cap = 0;
  hipEvent_t done = nullptr;
};

struct Prof {
  const char* name;
  hipEvent_t a, b;
};

// An event marking the completion of counter snapshots, shared by every map
// whose snapshot it covers (a scan wave's maps share one); the event goes
// back to the context's pool when the last map lets go of it.
struct SnapEvent {
  hipEvent_t e = nullptr;
  std::vector<hipEvent_t>* pool = nullptr;
  ~SnapEvent() {
    if (e) pool->push_back(e);
  }
};

}  // namespace

struct bxg_map;

struct bxg_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  uint64_t budget = 0;
  uint64_t used = 0;
  std::string err;
  std::recursive_mutex mu;
  int cu_count = 0;
  uint64_t hbm_bytes = 0;
  std::string name;

  // scan pipeline scratch
  uint64_t scan_cap = 0;
  ScanScratch sc{};
  uint32_t* d_counters = nullptr;
  ScanDesc* d_desc = nullptr;
  BatchHdr* d_hdr = nullptr;
  uint32_t desc_cap = 0;
  PinnedSlot stage[kStageSlots];
  int stage_next = 0;

  // bulk-op scratch
  uint64_t op_cap = 0;
  OpSlot* op_slots = nullptr;  // op_cap * 2 slots
  uint32_t* op_tpos = nullptr;
  uint32_t* op_leaf = nullptr;
  uint8_t* op_prebit = nullptr;

  // device staging of host-memory calls
  void* dstage[4] = {nullptr, nullptr, nullptr, nullptr};
  uint64_t dstage_cap[4] = {0, 0, 0, 0};

  uint64_t* d_small = nullptr;  // 8 words of results
  uint64_t* h_small = nullptr;  // pinned mirror

  hipEvent_t ev[64] = {};
  bool profiling = false;
  bool counting = false;
  // tuning (bxg_ctx_set_tuning): workgroups per CU of the work-stealing
  // update pass and leaf-allocation pass over the rays, the workgroup size
  // of both, the most rays a wavefront takes per cursor atomic (rays per
  // lane = ray_chunk / 64), the endpoint-table capacity per point and the
  // root-directory capacity per root (inverse load factors); the update
  // pass of a call with fewer points than its tuned grid has lanes runs on
  // small_ray_blocks_per_cu workgroups per CU
  uint32_t ray_blocks_per_cu = 4;
  uint32_t small_ray_blocks_per_cu = 2;
  uint32_t leaf_blocks_per_cu = 16;
  uint32_t ray_block = 256;
  uint32_t ray_chunk = 256;
  uint32_t ep_table_factor = 2;
  uint32_t root_slack = 2;
  uint32_t ray_window = 8;    // lockstep windows of the update pass (cells) and leaf pass (leaves)
  uint32_t leaf_window = 32;
  uint32_t fresh_pass = 1;    // maps empty at a scan's start: one create-and-mark walk
  uint32_t fresh_blocks_per_cu = 6;  // workgroups per CU of that walk
  uint32_t fresh_window = 16;        // its lockstep window (cells)
  uint32_t ray_order = 3;            // winner order: direction sector, length class, direction
  std::vector<Prof> prof_pending;
  std::vector<hipEvent_t> ev_free;
  std::vector<hipEvent_t> snap_pool;  // free snapshot events (SnapEvent)
  std::map<std::string, std::pair<uint64_t, double>> prof_totals;

  std::vector<bxg_map*> maps;
};

struct bxg_map {
  bxg_ctx* ctx = nullptr;
  MapDev h{};
  MapDev* d = nullptr;
  MapCounters* ctr = nullptr;
  MapCounters* h_ctr = nullptr;  // pinned mirror of the counters
  std::shared_ptr<SnapEvent> snap;  // completion of the last snapshot into h_ctr
  bool ctr_pending = false;
  uint32_t root_cap = 0;
  uint32_t leaf_cap = 0;
  uint64_t bytes = 0;
  uint64_t known_leaves = 0;
  uint64_t known_roots = 0;
  uint64_t known_scans = 0;
  uint64_t max_leaf_growth = 0;  // largest leaf-count increase seen between snapshots
  int4* d_extra = nullptr;  // rays due from addHit/MissPoint
  uint32_t extra_cap = 0;
  std::vector<int4> extra;
  // asynchronous inserts enqueued since the map was last confirmed, and the
  // device's scans_applied before the first of them (see settle)
  std::vector<bxg_scan_desc> pending;
  // the map's own copies of their sensor-to-world matrices (the caller's may
  // be gone by the time an overflowed insert runs again); a deque keeps the
  // copies in place as more are added
  std::deque<std::array<float, 16>> pending_tfs;
  uint64_t pending_base = 0;
  uint64_t replays = 0;  // pending inserts run again by settle after an overflow
};

struct bxg_graph {
  bxg_ctx* ctx = nullptr;
  uint32_t max_scans = 0;
  uint64_t max_points = 0;
  ScanScratch sc{};
  uint32_t* d_counters = nullptr;
  ScanDesc* d_desc = nullptr;
  BatchHdr* d_hdr = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<bxg_map*> maps;  // maps of the loaded frame
  std::vector<uint64_t> npoints;
  std::vector<double> ranges;
  // the loaded frame's scans (with their own copy of each transform), which
  // each launch leaves pending on its map until the map is confirmed
  std::vector<bxg_scan_desc> scans;
  std::vector<std::array<float, 16>> tfs;
  bool loaded = false;
};

namespace {

int fail(bxg_ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

int hip_fail(bxg_ctx* c, hipError_t e, const char* what) {
  return fail(c, BXG_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define BXG_HIP(c, expr)                                 \
  do {                                                   \
    hipError_t e_ = (expr);                              \
    if (e_ != hipSuccess) return hip_fail(c, e_, #expr); \
  } while (0)

#define BXG_TRY(expr)         \
  do {                        \
    int r_ = (expr);          \
    if (r_ != BXG_OK) return r_; \
  } while (0)

uint64_t pow2_at_least(uint64_t v) {
  uint64_t p = 1;
  while (p < v) p <<= 1;
  return p;
}

uint32_t nblocks(uint64_t n) { return (uint32_t)((n + kBlock - 1) / kBlock); }

hipEvent_t ev_take(bxg_ctx* c) {
  if (!c->ev_free.empty()) {
    hipEvent_t e = c->ev_free.back();
    c->ev_free.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  hipEventCreate(&e);
  return e;
}

// workgroups of k_lookup over n coordinates (kLookups a thread)
uint32_t lookup_blocks(uint64_t n) {
  return (uint32_t)((n + (uint64_t)kBlock * kLookups - 1) / ((uint64_t)kBlock * kLookups));
}

// launch on a stream of the context, bracketed by events when profiling
template <typename... KArgs, typename... Args>
int launch_on(bxg_ctx* c, hipStream_t stream, const char* name, void (*k)(KArgs...), uint32_t grid,
              uint32_t block, Args... args) {
  if (grid == 0) return BXG_OK;
  hipEvent_t a = nullptr;
  if (c->profiling) {
    a = ev_take(c);
    hipEventRecord(a, stream);
  }
  hipLaunchKernelGGL(k, dim3(grid), dim3(block), 0, stream, args...);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(c, e, name);
  if (c->profiling) {
    hipEvent_t b = ev_take(c);
    hipEventRecord(b, stream);
    c->prof_pending.push_back({name, a, b});
  }
  return BXG_OK;
}

// launch on the context stream
template <typename... KArgs, typename... Args>
int launch(bxg_ctx* c, const char* name, void (*k)(KArgs...), uint32_t grid, uint32_t block,
           Args... args) {
  return launch_on(c, c->stream, name, k, grid, block, args...);
}

int dmalloc(bxg_ctx* c, void** p, uint64_t bytes) {
  *p = nullptr;
  if (bytes == 0) bytes = 64;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return hip_fail(c, e, "hipMalloc");
  }
  return BXG_OK;
}

void dfree(void* p) {
  if (p) hipFree(p);
}

int fill32(bxg_ctx* c, uint32_t* p, uint64_t n, uint32_t v) {
  if (!n) return BXG_OK;
  const uint32_t g = (uint32_t)std::min<uint64_t>(nblocks(n), 4096);
  return launch(c, "k_fill_u32", k_fill_u32, g, kBlock, p, n, v);
}
int fill64(bxg_ctx* c, uint64_t* p, uint64_t n, uint64_t v) {
  if (!n) return BXG_OK;
  const uint32_t g = (uint32_t)std::min<uint64_t>(nblocks(n), 4096);
  return launch(c, "k_fill_u64", k_fill_u64, g, kBlock, p, n, v);
}

int ctx_sync(bxg_ctx* c) {
  BXG_HIP(c, hipStreamSynchronize(c->stream));
  return BXG_OK;
}

// device staging buffer `slot` of at least `bytes`
int dstage(bxg_ctx* c, int slot, uint64_t bytes, void** out) {
  if (c->dstage_cap[slot] < bytes) {
    BXG_TRY(ctx_sync(c));
    dfree(c->dstage[slot]);
    c->dstage[slot] = nullptr;
    c->dstage_cap[slot] = 0;
    const uint64_t cap = std::max<uint64_t>(pow2_at_least(bytes), 1 << 16);
    BXG_TRY(dmalloc(c, &c->dstage[slot], cap));
    c->dstage_cap[slot] = cap;
  }
  *out = c->dstage[slot];
  return BXG_OK;
}

// pinned host staging for stream-ordered uploads: the slot is reused only
// once its previous copy has completed
int pinned_slot(bxg_ctx* c, uint64_t bytes, PinnedSlot** out) {
  PinnedSlot& s = c->stage[c->stage_next];
  c->stage_next = (c->stage_next + 1) % kStageSlots;
  if (s.done) BXG_HIP(c, hipEventSynchronize(s.done));
  if (s.cap < bytes) {
    if (s.host) hipHostFree(s.host);
    s.host = nullptr;
    const uint64_t cap = std::max<uint64_t>(pow2_at_least(bytes), 4096);
    BXG_HIP(c, hipHostMalloc(&s.host, cap, hipHostMallocDefault));
    s.cap = cap;
  }
  if (!s.done) BXG_HIP(c, hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
  *out = &s;
  return BXG_OK;
}

// stream-ordered upload of a small host structure through pinned staging
int upload(bxg_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!bytes) return BXG_OK;
  PinnedSlot* s = nullptr;
  BXG_TRY(pinned_slot(c, bytes, &s));
  std::memcpy(s->host, src, bytes);
  BXG_HIP(c, hipMemcpyAsync(dst, s->host, bytes, hipMemcpyHostToDevice, c->stream));
  BXG_HIP(c, hipEventRecord(s->done, c->stream));
  return BXG_OK;
}

int h2d_sync(bxg_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!bytes) return BXG_OK;
  BXG_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
  return ctx_sync(c);
}
int d2h_sync(bxg_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!bytes) return BXG_OK;
  BXG_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
  return ctx_sync(c);
}

// ------------------------------------------------------------ map pools

// bytes of touched marks a leaf (probabilistic maps: one a cell)
uint64_t touched_bytes(const bxg_map* m) {
  return (m->h.flags & kMapProbabilistic) ? (uint64_t)m->h.leaf_size : 0;
}
uint64_t leaf_bytes(const bxg_map* m) {
  return (uint64_t)m->h.leaf_words * 8 + (uint64_t)m->h.leaf_size * m->h.value_bytes + sizeof(int4) +
         touched_bytes(m);
}
uint64_t root_bytes(const bxg_map* m) { return 8 + (uint64_t)m->h.inner_size * 4; }

int push_header(bxg_map* m) { return upload(m->ctx, m->d, &m->h, sizeof(MapDev)); }

int budget_check(bxg_ctx* c, uint64_t more) {
  if (c->budget && c->used + more > c->budget)
    return fail(c, BXG_ERR_BUDGET,
                "memory budget exceeded: " + std::to_string(c->used + more) + " > " +
                    std::to_string(c->budget));
  return BXG_OK;
}

// fold a counter snapshot into the host's view of the map
void take_counters(bxg_map* m) {
  const uint64_t leaves = std::min<uint64_t>(m->h_ctr->leaf_count, m->leaf_cap);
  if (leaves > m->known_leaves) m->max_leaf_growth = std::max(m->max_leaf_growth, leaves - m->known_leaves);
  m->known_leaves = leaves;
  m->known_roots = m->h_ctr->root_count;
  m->known_scans = m->h_ctr->scans_applied;
}

// read the device counters of a map (blocking)
int read_counters(bxg_map* m) {
  bxg_ctx* c = m->ctx;
  BXG_HIP(c, hipMemcpyAsync(m->h_ctr, m->ctr, sizeof(MapCounters), hipMemcpyDeviceToHost, c->stream));
  BXG_TRY(ctx_sync(c));
  m->ctr_pending = false;
  m->snap.reset();
  take_counters(m);
  return BXG_OK;
}

// Mark the snapshots enqueued so far for maps[0..n) with one event on the
// context stream (polled lazily by poll_counters).
int snapshot_event(bxg_ctx* c, bxg_map* const* maps, size_t n) {
  auto ev = std::make_shared<SnapEvent>();
  ev->pool = &c->snap_pool;
  if (!c->snap_pool.empty()) {
    ev->e = c->snap_pool.back();
    c->snap_pool.pop_back();
  } else {
    BXG_HIP(c, hipEventCreateWithFlags(&ev->e, hipEventDisableTiming));
  }
  BXG_HIP(c, hipEventRecord(ev->e, c->stream));
  for (size_t i = 0; i < n; ++i) {
    maps[i]->snap = ev;
    maps[i]->ctr_pending = true;
  }
  return BXG_OK;
}

// enqueue a counter snapshot after a mutating call (polled lazily)
int snapshot_counters(bxg_map* m) {
  bxg_ctx* c = m->ctx;
  BXG_HIP(c, hipMemcpyAsync(m->h_ctr, m->ctr, sizeof(MapCounters), hipMemcpyDeviceToHost, c->stream));
  return snapshot_event(c, &m, 1);
}

// take a finished snapshot into account without blocking
void poll_counters(bxg_map* m) {
  if (m->ctr_pending && hipEventQuery(m->snap->e) == hipSuccess) {
    m->ctr_pending = false;
    m->snap.reset();
    take_counters(m);
  }
}

int grow_leaves(bxg_map* m, uint64_t need) {
  bxg_ctx* c = m->ctx;
  if (need <= m->leaf_cap) return BXG_OK;
  uint64_t cap = std::max<uint64_t>(need, (uint64_t)m->leaf_cap * 2);
  cap = (cap + 255) & ~255ull;
  if (cap >= 0xFFFFFF00ull) return fail(c, BXG_ERR_OUT_OF_POOL, "leaf pool limit (2^32 leaves)");
  const uint64_t more = (cap - m->leaf_cap) * leaf_bytes(m);
  BXG_TRY(budget_check(c, more));
  BXG_TRY(read_counters(m));  // also synchronizes the stream
  const uint64_t used = std::min<uint64_t>(m->h_ctr->leaf_count, m->leaf_cap);
  const uint64_t lw = m->h.leaf_words, cb = (uint64_t)m->h.leaf_size * m->h.value_bytes;
  const uint64_t tb = touched_bytes(m);
  void *masks = nullptr, *cells = nullptr, *orig = nullptr, *touched = nullptr;
  BXG_TRY(dmalloc(c, &masks, cap * lw * 8));
  if (cb) BXG_TRY(dmalloc(c, &cells, cap * cb));
  BXG_TRY(dmalloc(c, &orig, cap * sizeof(int4)));
  if (tb) BXG_TRY(dmalloc(c, &touched, cap * tb));
  // the leaves in use move over; the rest of the pool stays uninitialised
  // (a leaf is written whole when it is created).  Touched marks start at
  // zero, and every scan leaves them so.
  if (tb) BXG_HIP(c, hipMemsetAsync(touched, 0, cap * tb, c->stream));
  if (used) {
    BXG_HIP(c, hipMemcpyAsync(masks, m->h.leaf_masks, used * lw * 8, hipMemcpyDeviceToDevice, c->stream));
    if (cb) BXG_HIP(c, hipMemcpyAsync(cells, m->h.cells, used * cb, hipMemcpyDeviceToDevice, c->stream));
    BXG_HIP(c, hipMemcpyAsync(orig, m->h.leaf_origin, used * sizeof(int4), hipMemcpyDeviceToDevice,
                              c->stream));
  }
  BXG_TRY(ctx_sync(c));
  dfree(m->h.leaf_masks);
  dfree(m->h.cells);
  dfree(m->h.leaf_origin);
  dfree(m->h.touched);
  m->h.leaf_masks = (uint64_t*)masks;
  m->h.cells = (uint8_t*)cells;
  m->h.leaf_origin = (int4*)orig;
  m->h.touched = (uint8_t*)touched;
  m->h.leaf_cap = (uint32_t)cap;
  c->used += more;
  m->bytes += more;
  m->leaf_cap = (uint32_t)cap;
  // allocations that failed for lack of room ran the bump counter past the
  // old capacity; every index below it is in use
  uint32_t lc = (uint32_t)used;
  BXG_HIP(c, hipMemcpyAsync(&m->ctr->leaf_count, &lc, 4, hipMemcpyHostToDevice, c->stream));
  BXG_TRY(ctx_sync(c));
  m->known_leaves = used;
  return push_header(m);
}

int grow_roots(bxg_map* m, uint64_t need_roots) {
  bxg_ctx* c = m->ctx;
  uint64_t cap = pow2_at_least(std::max<uint64_t>(need_roots * 2, 1024));
  if (cap <= m->root_cap) return BXG_OK;
  if (cap > (1ull << 24)) return fail(c, BXG_ERR_OUT_OF_POOL, "root directory limit (2^23 roots)");
  const uint64_t more = (cap - m->root_cap) * root_bytes(m);
  BXG_TRY(budget_check(c, more));
  BXG_TRY(ctx_sync(c));
  void *keys = nullptr, *children = nullptr;
  BXG_TRY(dmalloc(c, &keys, cap * 8));
  BXG_TRY(dmalloc(c, &children, cap * m->h.inner_size * 4));
  BXG_TRY(fill64(c, (uint64_t*)keys, cap, kEmptyKey));
  BXG_TRY(fill32(c, (uint32_t*)children, cap * m->h.inner_size, kEmptyChild));
  uint64_t* old_keys = m->h.root_keys;
  uint32_t* old_children = m->h.children;
  const uint32_t old_cap = m->root_cap;
  m->h.root_keys = (uint64_t*)keys;
  m->h.children = (uint32_t*)children;
  m->h.root_mask = (uint32_t)(cap - 1);
  m->h.max_probe = (uint32_t)std::min<uint64_t>(cap, 512);
  m->root_cap = (uint32_t)cap;
  BXG_TRY(push_header(m));
  uint32_t zero = 0;
  BXG_HIP(c, hipMemcpyAsync(&m->ctr->root_count, &zero, 4, hipMemcpyHostToDevice, c->stream));
  if (old_keys) BXG_TRY(launch(c, "k_rehash", k_rehash, nblocks(old_cap), kBlock, (const MapDev*)m->d,
                               (const uint64_t*)old_keys, (const uint32_t*)old_children, old_cap, 0u));
  BXG_TRY(ctx_sync(c));
  dfree(old_keys);
  dfree(old_children);
  c->used += more;
  m->bytes += more;
  return BXG_OK;
}

// Expected new leaves of an n-point scan with max range R.  Into a map
// without scans: at most eight leaves a point (long free-space rays of a
// sparse scan) and at most the leaves of the ball of radius R around the
// sensor; afterwards twice the largest growth seen so far.
uint64_t scan_leaf_estimate(const bxg_map* m, uint64_t n, double max_range) {
  double reach = max_range * m->h.inv_res / (double)(1u << m->h.leaf_bits) + 1.0;
  if (!(reach < 1e5)) reach = 1e5;
  const double ball = 4.19 * reach * reach * reach;
  const double est = std::min<double>((double)n * 8.0, ball) + 4096.0;
  const uint64_t first = (uint64_t)std::min<double>(est, (double)(1ull << 24));
  if (m->known_scans == 0) return first;
  return std::min<uint64_t>(first, 2 * m->max_leaf_growth + 4096);
}

// keep head-room ahead of a call expected to add up to `leaves` leaves and
// `roots` roots (estimates; the device reports an overflow if they are low).
// The leaf estimates are upper bounds already (the sensing ball, twice the
// largest growth seen, one leaf per cell written), so the pool grows only
// when it cannot hold the estimate, and then to a quarter beyond it (and at
// least to twice its size, grow_leaves).
int ensure_room(bxg_map* m, uint64_t leaves, uint64_t roots) {
  poll_counters(m);
  const uint64_t want_leaves = m->known_leaves + leaves;
  if (want_leaves > m->leaf_cap) BXG_TRY(grow_leaves(m, want_leaves + want_leaves / 4 + 1024));
  const uint64_t want_roots = m->known_roots + roots;
  const uint64_t slack = m->ctx->root_slack;
  if (want_roots * slack > m->root_cap) BXG_TRY(grow_roots(m, want_roots * slack));
  return BXG_OK;
}

// ----------------------------------------------------------- scratch

int ensure_scan_scratch(bxg_ctx* c, uint64_t points, uint32_t nscans) {
  if (points > c->scan_cap) {
    BXG_TRY(ctx_sync(c));
    dfree(c->sc.rec);
    dfree(c->sc.tpos);
    dfree(c->sc.winners);
    dfree(c->sc.wkey);
    dfree(c->sc.ep);
    uint32_t* hist = c->sc.hist;
    uint32_t* offs = c->sc.offs;
    c->sc = ScanScratch{};
    c->sc.hist = hist;
    c->sc.offs = offs;
    const uint64_t cap = std::max<uint64_t>(pow2_at_least(points), 1 << 12);
    if (cap > (1ull << 31)) return fail(c, BXG_ERR_INVALID_ARG, "too many points in one call");
    const uint64_t tcap = cap * c->ep_table_factor;
    void* p;
    BXG_TRY(dmalloc(c, &p, cap * sizeof(int4)));
    c->sc.rec = (int4*)p;
    BXG_TRY(dmalloc(c, &p, cap * 4));
    c->sc.tpos = (uint32_t*)p;
    BXG_TRY(dmalloc(c, &p, cap * 4));
    c->sc.winners = (uint32_t*)p;
    BXG_TRY(dmalloc(c, &p, cap * 4));
    c->sc.wkey = (uint32_t*)p;
    if (!c->sc.hist) {
      BXG_TRY(dmalloc(c, &p, (uint64_t)kMaxBuckets * 4));
      c->sc.hist = (uint32_t*)p;
      BXG_TRY(dmalloc(c, &p, (uint64_t)kMaxBuckets * 4));
      c->sc.offs = (uint32_t*)p;
      BXG_TRY(fill32(c, c->sc.hist, kMaxBuckets, 0));
    }
    BXG_TRY(dmalloc(c, &p, tcap * sizeof(EpSlot)));
    c->sc.ep = (EpSlot*)p;
    c->sc.ep_mask = (uint32_t)(tcap - 1);
    BXG_TRY(fill64(c, (uint64_t*)c->sc.ep, tcap * 2, ~0ull));  // empty: all ones
    c->scan_cap = cap;
  }
  c->sc.counters = c->d_counters;
  if (nscans > c->desc_cap) {
    BXG_TRY(ctx_sync(c));
    dfree(c->d_desc);
    c->d_desc = nullptr;
    const uint32_t cap = (uint32_t)pow2_at_least(std::max<uint32_t>(nscans, 64));
    void* p;
    BXG_TRY(dmalloc(c, &p, (uint64_t)cap * sizeof(ScanDesc)));
    c->d_desc = (ScanDesc*)p;
    c->desc_cap = cap;
  }
  return BXG_OK;
}

int ensure_op_scratch(bxg_ctx* c, uint64_t n) {
  if (n <= c->op_cap) return BXG_OK;
  BXG_TRY(ctx_sync(c));
  dfree(c->op_slots);
  dfree(c->op_tpos);
  dfree(c->op_leaf);
  dfree(c->op_prebit);
  c->op_slots = nullptr;
  c->op_tpos = c->op_leaf = nullptr;
  c->op_prebit = nullptr;
  c->op_cap = 0;
  const uint64_t cap = std::max<uint64_t>(pow2_at_least(n), 1 << 12);
  if (cap > (1ull << 31)) return fail(c, BXG_ERR_INVALID_ARG, "too many cells in one call");
  void* p;
  BXG_TRY(dmalloc(c, &p, cap * 2 * sizeof(OpSlot)));
  c->op_slots = (OpSlot*)p;
  BXG_TRY(dmalloc(c, &p, cap * 4));
  c->op_tpos = (uint32_t*)p;
  BXG_TRY(dmalloc(c, &p, cap * 4));
  c->op_leaf = (uint32_t*)p;
  BXG_TRY(dmalloc(c, &p, cap));
  c->op_prebit = (uint8_t*)p;
  c->op_cap = cap;
  return BXG_OK;
}

// The scratch of a bulk call of n inputs: a table of twice the power of two
// at or above n slots (load factor at most 1/2), emptied by the call itself.
OpScratch op_scratch(bxg_ctx* c, uint64_t n, uint64_t* slots) {
  OpScratch os;
  *slots = std::max<uint64_t>(pow2_at_least(n), 1 << 12) * 2;
  os.slots = c->op_slots;
  os.mask = (uint32_t)(*slots - 1);
  os.tpos = c->op_tpos;
  os.leaf = c->op_leaf;
  os.prebit = c->op_prebit;
  return os;
}

// ------------------------------------------------------------ scans

bool is_prob(const bxg_map* m) { return (m->h.flags & kMapProbabilistic) != 0; }

// posToCoord of the origin on the host, with the device's range rule
int origin_voxel(bxg_map* m, const double o[3], int32_t out[3]) {
  const double lim = (double)(1 << (20 + m->h.log2n));
  for (int a = 0; a < 3; ++a) {
    const double v = std::floor(o[a] * m->h.inv_res);
    if (!(v >= -lim && v < lim))
      return fail(m->ctx, BXG_ERR_COORD_RANGE, "origin outside the supported coordinate range");
    out[a] = (int32_t)v;
  }
  return BXG_OK;
}

// Fill the device descriptor of scan `in`, the tag-th of its wave, whose
// points start at flattened index off.
int build_desc(bxg_ctx* c, const bxg_scan_desc& in, uint32_t tag, uint64_t off, uint32_t extra_flags,
               ScanDesc& d) {
  bxg_map* m = in.map;
  if (!m || m->ctx != c) return fail(c, BXG_ERR_INVALID_ARG, "scan without a map of this context");
  if (!is_prob(m)) return fail(c, BXG_ERR_INVALID_ARG, "insertPointCloud needs a probabilistic map");
  if (in.n && !in.points_dev) return fail(c, BXG_ERR_INVALID_ARG, "null points");
  if (std::isnan(in.max_range)) return fail(c, BXG_ERR_INVALID_ARG, "max_range is NaN");
  if (in.n > 0xFFFFFFFFull) return fail(c, BXG_ERR_INVALID_ARG, "scan too large");
  std::memset(&d, 0, sizeof(d));
  d.map = m->d;
  d.points = in.points_dev;
  d.pt_off = off;
  d.n = (uint32_t)in.n;
  const bool f64 = (in.flags & BXG_POINTS_F64) != 0;
  d.flags = (f64 ? kScanF64 : 0u) | extra_flags;
  const uint32_t packed = f64 ? 24u : 12u;
  d.stride = in.stride ? in.stride : packed;
  if (d.stride < packed || d.stride % (f64 ? 8u : 4u))
    return fail(c, BXG_ERR_INVALID_ARG, "point stride smaller than a point or misaligned");
  double origin[3] = {in.origin[0], in.origin[1], in.origin[2]};
  if (in.sensor_to_world) {
    if (f64) return fail(c, BXG_ERR_INVALID_ARG, "the ingest transform takes float points");
    d.flags |= kScanTransform;
    std::memcpy(d.tf, in.sensor_to_world, sizeof(d.tf));
    // the sensor position: the matrix's translation (bonxai_server.cpp:172-175)
    for (int k = 0; k < 3; ++k) origin[k] = (double)in.sensor_to_world[4 * k + 3];
  }
  d.fx = origin[0];
  d.fy = origin[1];
  d.fz = origin[2];
  d.max_range = in.max_range;
  d.r2 = in.max_range * in.max_range;  // probabilistic_map.hpp:145
  const double reach = in.max_range * m->h.inv_res;
  d.reach = reach >= 1.0 ? (uint32_t)std::min(reach, 16777216.0) : 0u;
  int32_t o[3];
  BXG_TRY(origin_voxel(m, origin, o));
  d.ox = o[0];
  d.oy = o[1];
  d.oz = o[2];
  d.tag = tag;
  if (!(extra_flags & kScanNoRays) && !m->extra.empty()) {
    d.extra = m->d_extra;
    d.nextra = (uint32_t)m->extra.size();
  }
  return BXG_OK;
}

// Fill the device descriptors of one wave of scans (distinct maps).
int build_descs(bxg_ctx* c, const bxg_scan_desc* scans, uint32_t nscans, uint32_t extra_flags,
                std::vector<ScanDesc>& out, uint64_t* total) {
  out.resize(nscans);
  uint64_t off = 0;
  for (uint32_t s = 0; s < nscans; ++s) {
    BXG_TRY(build_desc(c, scans[s], s, off, extra_flags, out[s]));
    off += scans[s].n;
  }
  *total = off;
  if (off >= (1ull << 31)) return fail(c, BXG_ERR_INVALID_ARG, "too many points in one call");
  return BXG_OK;
}

// upload pending addHit/MissPoint rays of a map to the device
int upload_extra(bxg_map* m) {
  bxg_ctx* c = m->ctx;
  if (m->extra.empty()) return BXG_OK;
  if (m->extra.size() > m->extra_cap) {
    BXG_TRY(ctx_sync(c));
    dfree(m->d_extra);
    m->d_extra = nullptr;
    const uint32_t cap = (uint32_t)pow2_at_least(std::max<uint64_t>(m->extra.size(), 64));
    void* p;
    BXG_TRY(dmalloc(c, &p, (uint64_t)cap * sizeof(int4)));
    m->d_extra = (int4*)p;
    m->extra_cap = cap;
  }
  return upload(c, m->d_extra, m->extra.data(), m->extra.size() * sizeof(int4));
}

// the kernel sequence of one wave of scans on given scratch/descriptors
int enqueue_scan_kernels(bxg_ctx* c, const BatchHdr* hdr, const ScanDesc* desc, ScanScratch sc,
                         uint64_t grid_points, uint32_t nscans, bool any_extra) {
  const uint32_t gp = nblocks(grid_points);
  const uint32_t gw = std::max<uint32_t>(1, std::min<uint32_t>(gp, 16384));
  // work-stealing ray kernels: enough blocks to fill every CU, no more
  // than the rays could use (64 rays per wavefront, 4 wavefronts a block)
  const uint32_t cus = (uint32_t)std::max(c->cu_count, 1);
  // leaf pass: at most one lane per point (the winners are fewer); update
  // pass: up to 8, since a call with fewer rays than lanes walks each ray as
  // up to 8 segments (on the leaf pass the extra wavefronts cost more in
  // cursor traffic than the segments gain)
  const uint32_t rb = c->ray_block;
  const uint32_t most = (uint32_t)((grid_points + rb - 1) / rb);
  const uint32_t most8 = (uint32_t)std::min<uint64_t>((grid_points * 8 + rb - 1) / rb, 1u << 30);
  // the tuned grid serves large calls (batched frames, lockstep batches); a
  // call with fewer points than that grid has lanes walks ray segments on a
  // grid of small_ray_blocks_per_cu workgroups per CU (2 by default: more
  // lanes cut the rays into more, shorter segments, and each segment starts
  // with a seek and a root lookup of its own)
  const uint32_t gr_full = std::max<uint32_t>(1, cus * c->ray_blocks_per_cu);
  const uint32_t gr = grid_points >= (uint64_t)gr_full * rb
                          ? gr_full
                          : std::max<uint32_t>(1, std::min<uint32_t>(cus * c->small_ray_blocks_per_cu, most8));
  const uint32_t gl = std::max<uint32_t>(1, std::min<uint32_t>(cus * c->leaf_blocks_per_cu, most));
  // the fresh pass: its own workgroups per CU for large calls
  const uint32_t gf_full = std::max<uint32_t>(1, cus * c->fresh_blocks_per_cu);
  const uint32_t gf = grid_points >= (uint64_t)gf_full * rb
                          ? gf_full
                          : std::max<uint32_t>(1, std::min<uint32_t>(cus * c->small_ray_blocks_per_cu, most8));
  sc.chunk_max = c->ray_chunk;
  sc.win_cells = (int32_t)c->ray_window;
  sc.win_leaves = (int32_t)c->leaf_window;
  sc.fresh_pass = c->fresh_pass;
  sc.win_fresh = (int32_t)c->fresh_window;
  sc.ray_order = c->ray_order;
  // new-leaf pass: a wavefront per new leaf, grid-stride over them
  const uint32_t gm = std::max<uint32_t>(1, std::min<uint32_t>(cus * 8, gp));
  BXG_TRY(launch(c, "k_scan_begin", k_scan_begin, 1, kBlock, hdr, desc, sc));
  BXG_TRY(launch(c, "k_scan_classify", k_scan_classify, gp, kBlock, hdr, desc, sc));
  BXG_TRY(launch(c, "k_scan_winners", k_scan_winners, gp, kBlock, hdr, desc, sc));
  BXG_TRY(launch(c, "k_scan_bucket_reduce", k_scan_bucket_reduce, kTiles, 1024, hdr, sc));
  BXG_TRY(launch(c, "k_scan_bucket_scan", k_scan_bucket_scan, kTiles, 1024, hdr, sc));
  BXG_TRY(launch(c, "k_scan_scatter", k_scan_scatter, gp, kBlock, hdr, sc));
  BXG_TRY(launch(c, "k_scan_ray_leaves", k_scan_ray_leaves, gl, rb, desc, sc));
  if (any_extra)
    BXG_TRY(launch(c, "k_scan_extra_leaves", k_scan_extra_leaves, nscans, kBlock, hdr, desc));
  BXG_TRY(launch(c, "k_scan_endpoints", k_scan_endpoints, gw, kBlock, desc, sc));
  if (c->counting) {
    BXG_TRY(launch(c, "k_scan_fresh_rays", k_scan_fresh_rays<true>, gf, rb, desc, sc));
    BXG_TRY(launch(c, "k_scan_rays", k_scan_rays<true>, gr, rb, desc, sc));
    if (any_extra)
      BXG_TRY(launch(c, "k_scan_extra_rays", k_scan_extra_rays<true>, nscans, kBlock, hdr, desc));
    BXG_TRY(launch(c, "k_scan_new_endpoints", k_scan_new_endpoints<true>, gw, kBlock, desc, sc));
  } else {
    BXG_TRY(launch(c, "k_scan_fresh_rays", k_scan_fresh_rays<false>, gf, rb, desc, sc));
    BXG_TRY(launch(c, "k_scan_rays", k_scan_rays<false>, gr, rb, desc, sc));
    if (any_extra)
      BXG_TRY(launch(c, "k_scan_extra_rays", k_scan_extra_rays<false>, nscans, kBlock, hdr, desc));
    BXG_TRY(launch(c, "k_scan_new_endpoints", k_scan_new_endpoints<false>, gw, kBlock, desc, sc));
  }
  BXG_TRY(launch(c, "k_scan_new_leaves", k_scan_new_leaves, gm, kBlock, hdr, desc));
  BXG_TRY(launch(c, "k_scan_finalize", k_scan_finalize, 1, kBlock, hdr, desc, sc));
  return BXG_OK;
}

// reset the per-scan statistics of the maps of a counted run
int reset_scan_stats(bxg_map* m) {
  bxg_ctx* c = m->ctx;
  const uint64_t z[3] = {0, 0, 0};  // stat_rays, stat_ray_updates, stat_ray_visits
  return upload(c, &m->ctr->stat_rays, z, sizeof(z));
}

// bits of the scan index kept in a winner's bucket
uint32_t tag_bits_for(uint32_t nscans) {
  uint32_t b = 0;
  while (b < kMaxTagBits && (1u << b) < nscans) ++b;
  return b;
}

// one wave of scans, each into a different map
int run_wave(bxg_ctx* c, const bxg_scan_desc* scans, uint32_t nscans, uint32_t extra_flags) {
  std::vector<ScanDesc> descs;
  uint64_t total = 0;
  BXG_TRY(build_descs(c, scans, nscans, extra_flags, descs, &total));
  bool any_extra = false;
  for (uint32_t s = 0; s < nscans; ++s) {
    bxg_map* m = scans[s].map;
    // head-room for a scan: about one new leaf per point covers a first
    // LiDAR scan into an empty map (long free-space rays); an overflow is
    // still caught on the device and leaves the cells untouched
    BXG_TRY(ensure_room(m, scan_leaf_estimate(m, scans[s].n, scans[s].max_range), 1024));
    if (descs[s].nextra) {
      any_extra = true;
      BXG_TRY(upload_extra(m));
      descs[s].extra = m->d_extra;
    }
    if (c->counting) BXG_TRY(reset_scan_stats(m));
  }
  // the maps may have grown: descriptors point at device headers, which
  // stay put, so nothing to rebuild
  BXG_TRY(ensure_scan_scratch(c, std::max<uint64_t>(total, 1), nscans));
  BatchHdr hdr{nscans, tag_bits_for(nscans), total};
  BXG_TRY(upload(c, c->d_desc, descs.data(), descs.size() * sizeof(ScanDesc)));
  BXG_TRY(upload(c, c->d_hdr, &hdr, sizeof(hdr)));
  BXG_TRY(enqueue_scan_kernels(c, c->d_hdr, c->d_desc, c->sc, total, nscans, any_extra));
  // k_scan_finalize wrote every map's counters to its host mirror
  std::vector<bxg_map*> maps(nscans);
  for (uint32_t s = 0; s < nscans; ++s) maps[s] = scans[s].map;
  return snapshot_event(c, maps.data(), maps.size());
}

// split a batch into waves of distinct maps, keeping input order per map
int run_batch(bxg_ctx* c, const bxg_scan_desc* scans, uint32_t nscans, uint32_t extra_flags) {
  uint32_t start = 0;
  while (start < nscans) {
    uint32_t end = start;
    std::vector<bxg_map*> seen;
    while (end < nscans && std::find(seen.begin(), seen.end(), scans[end].map) == seen.end()) {
      seen.push_back(scans[end].map);
      ++end;
    }
    BXG_TRY(run_wave(c, scans + start, end - start, extra_flags));
    start = end;
  }
  return BXG_OK;
}

int status_code(uint32_t st) {
  if (!st) return BXG_OK;
  if (st & kStCoordRange) return BXG_ERR_COORD_RANGE;
  if (st & (kStOutOfPool | kStRootFull)) return BXG_ERR_OUT_OF_POOL;
  return BXG_ERR_INVALID_ARG;
}

int map_status(bxg_map* m) {
  BXG_TRY(read_counters(m));
  const uint32_t st = m->h_ctr->status;
  if (st) {
    const int code = status_code(st);
    return fail(m->ctx, code, std::string("device error on map: ") + bxg_status_string(code));
  }
  return BXG_OK;
}

int clear_status(bxg_map* m) {
  const uint32_t z = 0;
  BXG_HIP(m->ctx, hipMemcpyAsync(&m->ctr->status, &z, 4, hipMemcpyHostToDevice, m->ctx->stream));
  BXG_TRY(ctx_sync(m->ctx));
  // any snapshot enqueued before has landed; the mirror follows the device
  m->ctr_pending = false;
  m->snap.reset();
  m->h_ctr->status = 0;
  return BXG_OK;
}

// grow after a device overflow, from the counters read at the failure
int grow_after_overflow(bxg_map* m) {
  BXG_TRY(read_counters(m));
  const uint32_t st = m->h_ctr->status;
  BXG_TRY(clear_status(m));
  if (st & kStOutOfPool) BXG_TRY(grow_leaves(m, (uint64_t)m->leaf_cap * 2));
  if (st & kStRootFull) BXG_TRY(grow_roots(m, (uint64_t)m->root_cap));
  return BXG_OK;
}

// ---------------------------------------------------------- bulk ops

// Remember an enqueued insert until the map is confirmed, with a copy of
// its transform owned by the map.
void pending_push(bxg_map* m, const bxg_scan_desc& s) {
  m->pending.push_back(s);
  if (s.sensor_to_world) {
    m->pending_tfs.emplace_back();
    std::memcpy(m->pending_tfs.back().data(), s.sensor_to_world, sizeof(float) * 16);
    m->pending.back().sensor_to_world = m->pending_tfs.back().data();
  }
}

// Start a pending window on m if none is open: the device's scans_applied
// before the scans about to be enqueued, from the counter mirror when no
// snapshot is in flight (else read, blocking).  A map holding a device error
// reports it instead.
int pending_begin(bxg_map* m) {
  if (!m->pending.empty()) return BXG_OK;
  poll_counters(m);
  if (m->ctr_pending) BXG_TRY(read_counters(m));
  if (m->h_ctr->status) return map_status(m);
  m->pending_base = m->h_ctr->scans_applied;
  return BXG_OK;
}

// Confirm the asynchronous inserts enqueued on m since it was last confirmed
// (blocking).  A pool overflow among them is repaired here: the device ran
// them in order, the first to overflow and every later one on the map wrote
// no cell, and scans_applied tells how many completed; the pools grow and
// the rest run again in order, each re-run checked.  Any other call on the
// map settles first, so it sees every earlier insert applied.  Errors other
// than an overflow stay on the map for map_status to report.
int settle(bxg_map* m) {
  if (m->pending.empty()) return BXG_OK;
  bxg_ctx* c = m->ctx;
  std::vector<bxg_scan_desc> todo;
  todo.swap(m->pending);
  std::deque<std::array<float, 16>> tfs;  // what todo's transforms point at
  tfs.swap(m->pending_tfs);
  BXG_TRY(read_counters(m));
  const uint32_t overflow = kStOutOfPool | kStRootFull;
  if ((m->h_ctr->status & ~overflow) || !(m->h_ctr->status & overflow)) return BXG_OK;
  const uint64_t applied = m->h_ctr->scans_applied - m->pending_base;
  for (uint64_t i = applied; i < todo.size(); ++i) {
    for (int attempt = 0;; ++attempt) {
      if (m->h_ctr->status & overflow) {
        if (attempt > 8) return fail(c, BXG_ERR_OUT_OF_POOL, "could not grow the node pool");
        BXG_TRY(grow_after_overflow(m));
      }
      ++m->replays;
      BXG_TRY(run_batch(c, &todo[i], 1, 0));
      BXG_TRY(read_counters(m));
      if (m->h_ctr->status & ~overflow) return BXG_OK;
      if (!(m->h_ctr->status & overflow)) break;
    }
  }
  return BXG_OK;
}

int check_coords_args(bxg_map* m, const void* coords, uint64_t n) {
  if (!m) return BXG_ERR_INVALID_ARG;
  if (n && !coords) return fail(m->ctx, BXG_ERR_INVALID_ARG, "null coordinates");
  if (n >= (1ull << 31)) return fail(m->ctx, BXG_ERR_INVALID_ARG, "too many coordinates in one call");
  return settle(m);
}

int map_failed(bxg_map* m) {
  BXG_TRY(settle(m));
  poll_counters(m);
  if (!m->ctr_pending && m->h_ctr->status)
    return fail(m->ctx, BXG_ERR_MAP_FAILED, "the map holds a device error; call bxg_map_clear_error");
  return BXG_OK;
}

int cell_op(bxg_map* m, const int32_t* coords, const void* vals, uint64_t n, uint8_t* was_on,
            uint32_t op) {
  bxg_ctx* c = m->ctx;
  BXG_TRY(check_coords_args(m, coords, n));
  if (!n) return BXG_OK;
  if (op != kOpSetOff)
    BXG_TRY(ensure_room(m, std::min<uint64_t>(n, 1u << 22), std::min<uint64_t>(n, 1u << 16)));
  BXG_TRY(ensure_op_scratch(c, n));
  uint64_t slots = 0;
  const OpScratch os = op_scratch(c, n, &slots);
  BXG_TRY(launch(c, "k_op_clear", k_op_clear, (uint32_t)std::min<uint64_t>(nblocks(slots), 4096), kBlock,
                 os.slots, slots));
  BXG_TRY(launch(c, "k_op_begin", k_op_begin, 1, 64, (const MapDev*)m->d));
  if (op == kOpSetValue && !was_on)
    BXG_TRY(launch(c, "k_op_resolve", k_op_resolve<true, false>, nblocks(n), kBlock, (const MapDev*)m->d,
                   coords, n, os));
  else if (op != kOpSetOff)
    BXG_TRY(launch(c, "k_op_resolve", k_op_resolve<true, true>, nblocks(n), kBlock, (const MapDev*)m->d,
                   coords, n, os));
  else
    BXG_TRY(launch(c, "k_op_resolve", k_op_resolve<false, true>, nblocks(n), kBlock, (const MapDev*)m->d,
                   coords, n, os));
  if (op != kOpSetOff) {
    // the leaves the call created: at most one per input
    const uint64_t words = std::min<uint64_t>(n, m->leaf_cap) *
                           ((uint64_t)m->h.leaf_words + (uint64_t)m->h.leaf_size * m->h.value_bytes / 8);
    const uint32_t gz = (uint32_t)std::max<uint64_t>(
        1, std::min<uint64_t>(nblocks(words), (uint64_t)std::max(c->cu_count, 1) * 8));
    BXG_TRY(launch(c, "k_zero_new_leaves", k_zero_new_leaves, gz, kBlock, (const MapDev*)m->d));
  }
  BXG_TRY(launch(c, "k_op_apply", k_op_apply, nblocks(n), kBlock, (const MapDev*)m->d, coords,
                 (const uint8_t*)vals, n, op, was_on, os));
  return snapshot_counters(m);
}

int known_leaf_count(bxg_map* m, uint32_t* out) {
  BXG_TRY(read_counters(m));
  *out = (uint32_t)m->known_leaves;
  return BXG_OK;
}

}  // namespace

// ------------------------------------------------ allocation-free scans
//
// The scratch of bxg_insert_point_cloud_noalloc: the scan pipeline's per-call
// buffers for at most BXG_NOALLOC_MAX_POINTS points (endpoint records and
// table, winner lists, bucket counts and cursors, batch header and
// descriptor), in a file-scope device array.  Every GPU the library runs on
// holds its own copy from the library's load to the end of the process.  Host
// code never uses the array's name (a host-side stand-in, not a device
// address): the device address comes from hipGetSymbolAddress, taken once per
// device with that device current, at the first bxg_ctx_create on it.
namespace {

constexpr uint64_t kNaPoints = BXG_NOALLOC_MAX_POINTS;
constexpr uint64_t kNaEpSlots = kNaPoints * 2;  // endpoint table at load factor <= 1/2
constexpr uint64_t kNaOffHdr = 0;
constexpr uint64_t kNaOffDesc = 256;
constexpr uint64_t kNaOffCounters = 1024;
constexpr uint64_t kNaOffHist = 4096;
constexpr uint64_t kNaOffOffs = kNaOffHist + (uint64_t)kMaxBuckets * 4;
constexpr uint64_t kNaOffRec = kNaOffOffs + (uint64_t)kMaxBuckets * 4;
constexpr uint64_t kNaOffTpos = kNaOffRec + kNaPoints * sizeof(int4);
constexpr uint64_t kNaOffWinners = kNaOffTpos + kNaPoints * 4;
constexpr uint64_t kNaOffWkey = kNaOffWinners + kNaPoints * 4;
constexpr uint64_t kNaOffEp = kNaOffWkey + kNaPoints * 4;
constexpr uint64_t kNaBytes = kNaOffEp + kNaEpSlots * sizeof(EpSlot);
static_assert(sizeof(BatchHdr) <= kNaOffDesc - kNaOffHdr, "batch header slot");
static_assert(sizeof(ScanDesc) <= kNaOffCounters - kNaOffDesc, "descriptor slot");
static_assert(kCounterWords * 4 <= kNaOffHist - kNaOffCounters, "counter slot");
static_assert(kNaOffRec % 256 == 0 && kNaOffEp % 256 == 0, "aligned buffers");

constexpr int kNaMaxDevices = 64;
std::mutex g_na_mu;                      // guards g_na_ptr
uint8_t* g_na_ptr[kNaMaxDevices] = {};   // device address of the array, per device
std::mutex g_na_run[kNaMaxDevices];      // one call at a time per device

}  // namespace

static __device__ __align__(256) uint8_t scratch[kNaBytes];

namespace {

// Take the array's device address on `device` (current) once.
int na_scratch_init(int device) {
  if (device < 0 || device >= kNaMaxDevices) return BXG_ERR_NOT_SUPPORTED;
  std::lock_guard<std::mutex> lk(g_na_mu);
  if (g_na_ptr[device]) return BXG_OK;
  void* ptr = nullptr;
  if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return BXG_ERR_HIP;
  g_na_ptr[device] = (uint8_t*)ptr;
  return BXG_OK;
}

uint8_t* na_scratch(const bxg_ctx* c) {
  if (c->device < 0 || c->device >= kNaMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lk(g_na_mu);
  return g_na_ptr[c->device];
}

// the scan pipeline's buffers inside the array, for a call of n points
ScanScratch na_layout(uint8_t* base, uint64_t n, uint64_t* ep_slots) {
  ScanScratch sc{};
  sc.rec = (int4*)(base + kNaOffRec);
  sc.tpos = (uint32_t*)(base + kNaOffTpos);
  sc.winners = (uint32_t*)(base + kNaOffWinners);
  sc.wkey = (uint32_t*)(base + kNaOffWkey);
  sc.hist = (uint32_t*)(base + kNaOffHist);
  sc.offs = (uint32_t*)(base + kNaOffOffs);
  sc.counters = (uint32_t*)(base + kNaOffCounters);
  sc.ep = (EpSlot*)(base + kNaOffEp);
  const uint64_t slots = std::max<uint64_t>(pow2_at_least(n), 1 << 12) * 2;  // <= kNaEpSlots
  sc.ep_mask = (uint32_t)(slots - 1);
  *ep_slots = slots;
  return sc;
}

// One scan of n device points into m on the array's buffers, waited for.
// The map's pools are used as they are: a scan they cannot hold writes no
// cell and returns BXG_ERR_OUT_OF_POOL.
int insert_noalloc(bxg_map* m, const bxg_scan_desc& in) {
  bxg_ctx* c = m->ctx;
  uint8_t* base = na_scratch(c);
  if (!base) return fail(c, BXG_ERR_HIP, "the allocation-free scratch has no device address");
  std::lock_guard<std::mutex> run(g_na_run[c->device]);
  const uint32_t overflow = kStOutOfPool | kStRootFull;
  // earlier asynchronous inserts of the map: confirmed here when none of them
  // overflowed (repairing an overflow grows the pools, which this call never
  // does)
  if (!m->pending.empty()) {
    BXG_TRY(read_counters(m));
    if (m->h_ctr->status & overflow)
      return fail(c, BXG_ERR_MAP_FAILED,
                  "an earlier asynchronous insert of the map ran out of nodes: bxg_map_sync repairs it "
                  "(and grows the pools)");
    BXG_TRY(settle(m));
  }
  poll_counters(m);
  if (m->ctr_pending) BXG_TRY(read_counters(m));
  if (m->h_ctr->status)
    return fail(c, BXG_ERR_MAP_FAILED, "the map holds a device error; call bxg_map_clear_error");
  if (!m->extra.empty())
    return fail(c, BXG_ERR_NOT_SUPPORTED,
                "addHitPoint/addMissPoint rays are due: bxg_insert_point_cloud traces them");
  ScanDesc d;
  BXG_TRY(build_desc(c, in, 0, 0, 0u, d));
  const BatchHdr hdr{1u, tag_bits_for(1), in.n};
  uint64_t slots = 0;
  const ScanScratch sc = na_layout(base, in.n, &slots);
  BatchHdr* d_hdr = (BatchHdr*)(base + kNaOffHdr);
  ScanDesc* d_desc = (ScanDesc*)(base + kNaOffDesc);
  // the array keeps the last call's contents: clear what starts empty
  BXG_TRY(fill64(c, (uint64_t*)sc.ep, slots * 2, ~0ull));
  BXG_TRY(fill32(c, sc.hist, kMaxBuckets, 0));
  BXG_TRY(fill32(c, sc.counters, kCounterWords, 0));
  BXG_TRY(launch(c, "k_na_setup", k_na_setup, 1, 64, hdr, d, d_hdr, d_desc, m->ctr,
                 c->counting ? 1u : 0u));
  BXG_TRY(enqueue_scan_kernels(c, d_hdr, d_desc, sc, in.n, 1, false));
  BXG_TRY(read_counters(m));  // waits for the scan
  const uint32_t st = m->h_ctr->status;
  if (st & ~overflow) return map_status(m);
  if (st) {
    BXG_TRY(launch(c, "k_na_recover", k_na_recover, 1, 64, m->ctr, m->leaf_cap));
    BXG_TRY(read_counters(m));
    return fail(c, BXG_ERR_OUT_OF_POOL,
                "the map's pools cannot hold the scan (no cell changed): reserve more and call again");
  }
  return BXG_OK;
}

}  // namespace

// ====================================================================
//                                C ABI
// ====================================================================

extern "C" {

const char* bxg_status_string(int status) {
  switch (status) {
    case BXG_OK: return "ok";
    case BXG_ERR_INVALID_ARG: return "invalid argument";
    case BXG_ERR_OUT_OF_POOL: return "node pool or root directory exhausted";
    case BXG_ERR_HIP: return "HIP runtime error";
    case BXG_ERR_COORD_RANGE: return "coordinate outside the supported range";
    case BXG_ERR_BUDGET: return "memory budget exceeded";
    case BXG_ERR_NOT_SUPPORTED: return "not supported";
    case BXG_ERR_CAPACITY: return "output buffer too small";
    case BXG_ERR_MAP_FAILED: return "map holds an unacknowledged device error";
    default: return "unknown status";
  }
}

int bxg_ctx_create(int device, uint64_t budget_bytes, bxg_ctx** out) {
  if (!out) return BXG_ERR_INVALID_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return BXG_ERR_HIP;
  if (device < 0 || device >= ndev) return BXG_ERR_INVALID_ARG;
  if (hipSetDevice(device) != hipSuccess) return BXG_ERR_HIP;
  bxg_ctx* c = new bxg_ctx();
  c->device = device;
  c->budget = budget_bytes;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    c->cu_count = prop.multiProcessorCount;
    c->hbm_bytes = prop.totalGlobalMem;
    c->name = prop.gcnArchName;
  }
  int rc = BXG_OK;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) rc = BXG_ERR_HIP;
  void* p = nullptr;
  if (rc == BXG_OK) rc = dmalloc(c, &p, kCounterWords * 4);
  c->d_counters = (uint32_t*)p;
  if (rc == BXG_OK) rc = dmalloc(c, &p, 64);
  c->d_hdr = (BatchHdr*)p;
  if (rc == BXG_OK) rc = dmalloc(c, &p, 64);
  c->d_small = (uint64_t*)p;
  if (rc == BXG_OK && hipHostMalloc((void**)&c->h_small, 64, hipHostMallocDefault) != hipSuccess)
    rc = BXG_ERR_HIP;
  for (int i = 0; rc == BXG_OK && i < 64; ++i)
    if (hipEventCreate(&c->ev[i]) != hipSuccess) rc = BXG_ERR_HIP;
  if (rc == BXG_OK) rc = fill32(c, c->d_counters, kCounterWords, 0);
  if (rc == BXG_OK) rc = ctx_sync(c);
  // the allocation-free scratch's address on this device (an error here
  // leaves the other entry points working; the allocation-free one reports it)
  if (rc == BXG_OK) na_scratch_init(device);
  if (rc != BXG_OK) {
    bxg_ctx_destroy(c);
    return rc;
  }
  *out = c;
  return BXG_OK;
}

int bxg_ctx_destroy(bxg_ctx* c) {
  if (!c) return BXG_ERR_INVALID_ARG;
  if (c->stream) hipStreamSynchronize(c->stream);
  while (!c->maps.empty()) bxg_map_destroy(c->maps.back());
  dfree(c->sc.rec);
  dfree(c->sc.tpos);
  dfree(c->sc.winners);
  dfree(c->sc.wkey);
  dfree(c->sc.hist);
  dfree(c->sc.offs);
  dfree(c->sc.ep);
  dfree(c->d_counters);
  dfree(c->d_desc);
  dfree(c->d_hdr);
  dfree(c->op_slots);
  dfree(c->op_tpos);
  dfree(c->op_leaf);
  dfree(c->op_prebit);
  for (int i = 0; i < 4; ++i) dfree(c->dstage[i]);
  dfree(c->d_small);
  if (c->h_small) hipHostFree(c->h_small);
  for (auto& s : c->stage) {
    if (s.host) hipHostFree(s.host);
    if (s.done) hipEventDestroy(s.done);
  }
  for (auto& e : c->ev)
    if (e) hipEventDestroy(e);
  for (auto& p : c->prof_pending) {
    hipEventDestroy(p.a);
    hipEventDestroy(p.b);
  }
  for (auto e : c->ev_free) hipEventDestroy(e);
  for (auto e : c->snap_pool) hipEventDestroy(e);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return BXG_OK;
}

int bxg_ctx_sync(bxg_ctx* c) {
  if (!c) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  return ctx_sync(c);
}

const char* bxg_ctx_last_error(bxg_ctx* c) { return c ? c->err.c_str() : "no context"; }

int bxg_ctx_stream(bxg_ctx* c, void** s) {
  if (!c || !s) return BXG_ERR_INVALID_ARG;
  *s = (void*)c->stream;
  return BXG_OK;
}

int bxg_ctx_memory(bxg_ctx* c, uint64_t* used, uint64_t* budget) {
  if (!c) return BXG_ERR_INVALID_ARG;
  if (used) *used = c->used;
  if (budget) *budget = c->budget;
  return BXG_OK;
}

int bxg_ctx_device_info(bxg_ctx* c, char* name, uint64_t cap, int* cu, uint64_t* hbm) {
  if (!c) return BXG_ERR_INVALID_ARG;
  if (name && cap) {
    std::strncpy(name, c->name.c_str(), cap - 1);
    name[cap - 1] = 0;
  }
  if (cu) *cu = c->cu_count;
  if (hbm) *hbm = c->hbm_bytes;
  return BXG_OK;
}

int bxg_malloc(bxg_ctx* c, uint64_t bytes, void** out) {
  if (!c || !out) return BXG_ERR_INVALID_ARG;
  return dmalloc(c, out, bytes);
}
int bxg_free(bxg_ctx* c, void* p) {
  if (!c) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  BXG_TRY(ctx_sync(c));
  dfree(p);
  return BXG_OK;
}
int bxg_memcpy_h2d(bxg_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  return h2d_sync(c, dst, src, bytes);
}
int bxg_memcpy_d2h(bxg_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  return d2h_sync(c, dst, src, bytes);
}
int bxg_memcpy_d2d(bxg_ctx* c, void* dst, const void* src, uint64_t bytes) {
  if (!c) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  if (!bytes) return BXG_OK;
  BXG_HIP(c, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, c->stream));
  return BXG_OK;
}

int bxg_event_record(bxg_ctx* c, uint32_t slot) {
  if (!c || slot >= 64) return BXG_ERR_INVALID_ARG;
  BXG_HIP(c, hipEventRecord(c->ev[slot], c->stream));
  return BXG_OK;
}
int bxg_event_elapsed(bxg_ctx* c, uint32_t a, uint32_t b, float* ms) {
  if (!c || a >= 64 || b >= 64 || !ms) return BXG_ERR_INVALID_ARG;
  BXG_HIP(c, hipEventSynchronize(c->ev[b]));
  BXG_HIP(c, hipEventElapsedTime(ms, c->ev[a], c->ev[b]));
  return BXG_OK;
}

static int prof_fold(bxg_ctx* c) {
  if (c->prof_pending.empty()) return BXG_OK;
  BXG_TRY(ctx_sync(c));
  for (auto& p : c->prof_pending) {
    float ms = 0.f;
    hipEventElapsedTime(&ms, p.a, p.b);
    auto& t = c->prof_totals[p.name];
    t.first += 1;
    t.second += ms;
    c->ev_free.push_back(p.a);
    c->ev_free.push_back(p.b);
  }
  c->prof_pending.clear();
  return BXG_OK;
}

int bxg_profile_enable(bxg_ctx* c, int on) {
  if (!c) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  BXG_TRY(prof_fold(c));
  c->profiling = on != 0;
  return BXG_OK;
}
int bxg_profile_read(bxg_ctx* c, const char* kernel, uint64_t* launches, double* total_ms) {
  if (!c || !kernel) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  BXG_TRY(prof_fold(c));
  auto it = c->prof_totals.find(kernel);
  if (launches) *launches = it == c->prof_totals.end() ? 0 : it->second.first;
  if (total_ms) *total_ms = it == c->prof_totals.end() ? 0.0 : it->second.second;
  return BXG_OK;
}
int bxg_profile_reset(bxg_ctx* c) {
  if (!c) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  BXG_TRY(prof_fold(c));
  c->prof_totals.clear();
  return BXG_OK;
}
}  // extern "C"

namespace {

// the tuning keys: name, field, smallest and largest value, required divisor
struct TuningKey {
  const char* name;
  uint32_t bxg_ctx::*field;
  int64_t lo, hi, step;
};
const TuningKey kTuning[] = {
    {"ray_blocks_per_cu", &bxg_ctx::ray_blocks_per_cu, 1, 256, 1},
    {"leaf_blocks_per_cu", &bxg_ctx::leaf_blocks_per_cu, 1, 256, 1},
    {"small_ray_blocks_per_cu", &bxg_ctx::small_ray_blocks_per_cu, 1, 64, 1},
    {"ray_block", &bxg_ctx::ray_block, 64, 256, 64},
    {"ray_chunk", &bxg_ctx::ray_chunk, 64, 1024, 64},
    {"ep_table_factor", &bxg_ctx::ep_table_factor, 2, 8, 1},
    {"root_slack", &bxg_ctx::root_slack, 2, 8, 1},
    {"ray_window", &bxg_ctx::ray_window, 1, 1 << 20, 1},
    {"leaf_window", &bxg_ctx::leaf_window, 1, 1 << 20, 1},
    {"fresh_pass", &bxg_ctx::fresh_pass, 0, 1, 1},
    {"fresh_blocks_per_cu", &bxg_ctx::fresh_blocks_per_cu, 1, 256, 1},
    {"fresh_window", &bxg_ctx::fresh_window, 1, 1 << 20, 1},
    {"ray_order", &bxg_ctx::ray_order, 0, 5, 1},
};

const TuningKey* tuning_key(const char* key) {
  for (const TuningKey& k : kTuning)
    if (!std::strcmp(key, k.name)) return &k;
  return nullptr;
}

}  // namespace

extern "C" {

int bxg_ctx_set_tuning(bxg_ctx* c, const char* key, int64_t value) {
  if (!c || !key) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  const TuningKey* k = tuning_key(key);
  if (!k) return fail(c, BXG_ERR_INVALID_ARG, std::string("unknown tuning key ") + key);
  if (value < k->lo || value > k->hi || value % k->step)
    return fail(c, BXG_ERR_INVALID_ARG, std::string(key) + ": value out of range");
  if (k->field == &bxg_ctx::ep_table_factor && (value & (value - 1)))
    return fail(c, BXG_ERR_INVALID_ARG, "ep_table_factor must be a power of two");
  if (k->field == &bxg_ctx::ep_table_factor && (uint32_t)value != c->ep_table_factor) {
    BXG_TRY(ctx_sync(c));
    c->scan_cap = 0;  // the endpoint table is rebuilt at the next scan
  }
  c->*(k->field) = (uint32_t)value;
  return BXG_OK;
}

int bxg_ctx_get_tuning(bxg_ctx* c, const char* key, int64_t* value) {
  if (!c || !key || !value) return BXG_ERR_INVALID_ARG;
  const TuningKey* k = tuning_key(key);
  if (!k) return fail(c, BXG_ERR_INVALID_ARG, std::string("unknown tuning key ") + key);
  *value = c->*(k->field);
  return BXG_OK;
}

int bxg_set_counting(bxg_ctx* c, int on) {
  if (!c) return BXG_ERR_INVALID_ARG;
  c->counting = on != 0;
  return BXG_OK;
}

// -------------------------------------------------------------- maps

int bxg_map_create(bxg_ctx* c, double voxel_size, uint32_t inner_bits, uint32_t leaf_bits,
                   uint32_t value_bytes, uint32_t flags, bxg_map** out) {
  if (!c || !out) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  *out = nullptr;
  // VoxelGrid's constructor rejects bits < 1 (bonxai.hpp:399-401)
  if (inner_bits < 1 || leaf_bits < 1)
    return fail(c, BXG_ERR_INVALID_ARG, "The minimum value of the inner_bits and leaf_bits should be 1");
  if (inner_bits > 4 || leaf_bits > 4)
    return fail(c, BXG_ERR_NOT_SUPPORTED, "inner_bits and leaf_bits above 4 are not supported");
  if (!(voxel_size > 0.0) || !std::isfinite(voxel_size))
    return fail(c, BXG_ERR_INVALID_ARG, "voxel_size must be positive and finite");
  if (value_bytes > 256) return fail(c, BXG_ERR_NOT_SUPPORTED, "cell values above 256 bytes");
  if ((flags & BXG_MAP_PROBABILISTIC) && value_bytes != 4)
    return fail(c, BXG_ERR_INVALID_ARG, "probabilistic cells are 4 bytes");
  bxg_map* m = new bxg_map();
  m->ctx = c;
  MapDev& h = m->h;
  h.res = voxel_size;
  h.inv_res = 1.0 / voxel_size;  // bonxai.hpp:395
  h.inner_bits = inner_bits;
  h.leaf_bits = leaf_bits;
  h.log2n = inner_bits + leaf_bits;
  h.inner_size = 1u << (3 * inner_bits);
  h.leaf_size = 1u << (3 * leaf_bits);
  h.leaf_words = h.leaf_size >= 64 ? h.leaf_size / 64 : 1;
  h.value_bytes = value_bytes;
  h.flags = (flags & BXG_MAP_PROBABILISTIC) ? kMapProbabilistic : 0u;
  // Options defaults (probabilistic_map.hpp:56-64): logods() of 0.4f, 0.7f,
  // 0.12f, 0.97f and 0.5
  h.opt_miss = -405465;
  h.opt_hit = 847297;
  h.opt_cmin = -1992430;
  h.opt_cmax = 3476099;
  h.opt_thr = 0;
  int rc = BXG_OK;
  void* p = nullptr;
  rc = dmalloc(c, &p, sizeof(MapDev));
  m->d = (MapDev*)p;
  if (rc == BXG_OK) rc = dmalloc(c, &p, sizeof(MapCounters));
  m->ctr = (MapCounters*)p;
  if (rc == BXG_OK && hipHostMalloc((void**)&m->h_ctr, sizeof(MapCounters), hipHostMallocDefault) != hipSuccess)
    rc = BXG_ERR_HIP;
  if (rc == BXG_OK && hipHostGetDevicePointer((void**)&h.ctr_host, m->h_ctr, 0) != hipSuccess)
    rc = BXG_ERR_HIP;
  if (rc == BXG_OK) {
    std::memset(m->h_ctr, 0, sizeof(MapCounters));
    m->h_ctr->cur = 1;  // _update_count starts at 1 (probabilistic_map.hpp:129)
    h.ctr = m->ctr;
    rc = h2d_sync(c, m->ctr, m->h_ctr, sizeof(MapCounters));
  }
  c->maps.push_back(m);
  if (rc == BXG_OK) rc = grow_roots(m, 2048);
  if (rc == BXG_OK) rc = grow_leaves(m, 4096);
  if (rc == BXG_OK) rc = push_header(m);
  if (rc == BXG_OK) rc = ctx_sync(c);
  if (rc != BXG_OK) {
    bxg_map_destroy(m);
    return rc;
  }
  *out = m;
  return BXG_OK;
}

int bxg_map_destroy(bxg_map* m) {
  if (!m) return BXG_ERR_INVALID_ARG;
  bxg_ctx* c = m->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  hipStreamSynchronize(c->stream);
  dfree(m->h.root_keys);
  dfree(m->h.children);
  dfree(m->h.leaf_masks);
  dfree(m->h.cells);
  dfree(m->h.leaf_origin);
  dfree(m->h.touched);
  dfree(m->d);
  dfree(m->ctr);
  dfree(m->d_extra);
  m->snap.reset();
  if (m->h_ctr) hipHostFree(m->h_ctr);
  c->used -= std::min(c->used, m->bytes);
  c->maps.erase(std::remove(c->maps.begin(), c->maps.end(), m), c->maps.end());
  delete m;
  return BXG_OK;
}

int bxg_map_reserve(bxg_map* m, uint64_t leaves, uint64_t roots) {
  if (!m) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  BXG_TRY(settle(m));
  if (leaves > m->leaf_cap) BXG_TRY(grow_leaves(m, leaves));
  if (roots * 2 > m->root_cap) BXG_TRY(grow_roots(m, roots));
  return BXG_OK;
}

int bxg_map_clear(bxg_map* m, int option) {
  if (!m) return BXG_ERR_INVALID_ARG;
  bxg_ctx* c = m->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  BXG_TRY(settle(m));
  uint32_t nl = 0;
  BXG_TRY(known_leaf_count(m, &nl));
  if (option == BXG_SET_ALL_CELLS_OFF) {
    const uint64_t work = std::max<uint64_t>((uint64_t)nl * m->h.leaf_words,
                                             (uint64_t)nl * m->h.leaf_size * m->h.value_bytes / 4);
    BXG_TRY(launch(c, "k_cells_off", k_cells_off, (uint32_t)std::min<uint64_t>(nblocks(work), 8192),
                   kBlock, (const MapDev*)m->d, nl));
    return ctx_sync(c);
  }
  if (option != BXG_CLEAR_MEMORY) return fail(c, BXG_ERR_INVALID_ARG, "unknown clear option");
  // CLEAR_MEMORY: empty directory and pool (the pools keep their capacity;
  // the old leaves' bytes stay: a leaf is written whole when it is created)
  BXG_TRY(fill64(c, m->h.root_keys, m->root_cap, kEmptyKey));
  BXG_TRY(fill32(c, m->h.children, (uint64_t)m->root_cap * m->h.inner_size, kEmptyChild));
  static const uint32_t z[2] = {0, 0};  // leaf_count, root_count; then status
  BXG_HIP(c, hipMemcpyAsync(&m->ctr->leaf_count, z, 8, hipMemcpyHostToDevice, c->stream));
  BXG_HIP(c, hipMemcpyAsync(&m->ctr->status, z, 4, hipMemcpyHostToDevice, c->stream));
  m->extra.clear();
  BXG_TRY(ctx_sync(c));
  m->known_leaves = m->known_roots = 0;
  return BXG_OK;
}

int bxg_map_sync(bxg_map* m) {
  if (!m) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  BXG_TRY(settle(m));
  return map_status(m);
}

int bxg_map_clear_error(bxg_map* m) {
  if (!m) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  BXG_TRY(settle(m));
  BXG_TRY(read_counters(m));
  const uint32_t st = m->h_ctr->status;
  BXG_TRY(clear_status(m));
  // make room for what failed
  if (st & kStOutOfPool) BXG_TRY(grow_leaves(m, (uint64_t)m->leaf_cap * 2));
  if (st & kStRootFull) BXG_TRY(grow_roots(m, (uint64_t)m->root_cap));
  return BXG_OK;
}

int bxg_map_get_stats(bxg_map* m, bxg_map_stats* out) {
  if (!m || !out) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  BXG_TRY(settle(m));
  BXG_TRY(read_counters(m));
  const MapCounters& k = *m->h_ctr;
  out->leaf_count = std::min<uint64_t>(k.leaf_count, m->leaf_cap);
  out->leaf_capacity = m->leaf_cap;
  out->root_count = k.root_count;
  out->root_capacity = m->root_cap;
  out->scans_applied = k.scans_applied;
  out->device_bytes = m->bytes;
  out->last_rays = k.stat_rays;
  out->last_ray_updates = k.stat_ray_updates;
  out->last_ray_visits = k.stat_ray_visits;
  out->replays = m->replays;
  out->update_count = (uint32_t)k.cur;
  out->status = k.status;
  return BXG_OK;
}

int bxg_map_active_cells(bxg_map* m, uint64_t* out) {
  if (!m || !out) return BXG_ERR_INVALID_ARG;
  bxg_ctx* c = m->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  BXG_TRY(settle(m));
  uint32_t nl = 0;
  BXG_TRY(known_leaf_count(m, &nl));
  BXG_HIP(c, hipMemsetAsync(c->d_small, 0, 8, c->stream));
  BXG_TRY(launch(c, "k_count_on", k_count_on, nblocks((uint64_t)nl * m->h.leaf_words), kBlock,
                 (const MapDev*)m->d, nl, c->d_small));
  BXG_TRY(d2h_sync(c, c->h_small, c->d_small, 8));
  *out = c->h_small[0];
  return BXG_OK;
}

int bxg_map_mem_usage(bxg_map* m, uint64_t* out) {
  if (!m || !out) return BXG_ERR_INVALID_ARG;
  *out = m->bytes + sizeof(MapDev) + sizeof(MapCounters);
  return BXG_OK;
}

double bxg_map_voxel_size(bxg_map* m) { return m ? m->h.res : 0.0; }

int bxg_map_digest(bxg_map* m, uint64_t out[4]) {
  if (!m || !out) return BXG_ERR_INVALID_ARG;
  bxg_ctx* c = m->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  BXG_TRY(settle(m));
  uint32_t nl = 0;
  BXG_TRY(known_leaf_count(m, &nl));
  BXG_HIP(c, hipMemsetAsync(c->d_small, 0, 32, c->stream));
  BXG_TRY(launch(c, "k_digest", k_digest, nblocks((uint64_t)nl * m->h.leaf_size), kBlock,
                 (const MapDev*)m->d, nl, c->d_small));
  BXG_TRY(d2h_sync(c, c->h_small, c->d_small, 32));
  std::memcpy(out, c->h_small, 32);
  return BXG_OK;
}

int bxg_map_release_unused(bxg_map* m) {
  if (!m) return BXG_ERR_INVALID_ARG;
  bxg_ctx* c = m->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  BXG_TRY(map_failed(m));
  uint32_t nl = 0;
  BXG_TRY(known_leaf_count(m, &nl));
  // 1. live leaves and their order-preserving new indices
  const uint32_t ntiles = (uint32_t)((nl + kScanTile - 1) / kScanTile);
  void *dlive = nullptr, *didx = nullptr, *dtiles = nullptr;
  BXG_TRY(dstage(c, 0, (uint64_t)nl * 4 + 64, &dlive));
  BXG_TRY(dstage(c, 1, (uint64_t)nl * 4 + 64, &didx));
  BXG_TRY(dstage(c, 2, (uint64_t)ntiles * 4 + 64, &dtiles));
  uint32_t live_leaves = 0;
  if (nl) {
    BXG_TRY(launch(c, "k_leaf_live", k_leaf_live, nblocks(nl), kBlock, (const MapDev*)m->d, nl,
                   (uint32_t*)dlive));
    BXG_TRY(launch(c, "k_tile_reduce", k_tile_reduce, ntiles, 1024, (const uint32_t*)dlive,
                   (uint64_t)nl, (uint32_t*)dtiles));
    BXG_TRY(launch(c, "k_tile_sums_scan", k_tile_sums_scan, 1, 1024, (uint32_t*)dtiles, ntiles,
                   (uint32_t*)c->d_small));
    BXG_TRY(launch(c, "k_tile_scan", k_tile_scan, ntiles, 1024, (const uint32_t*)dlive, (uint64_t)nl,
                   (const uint32_t*)dtiles, (uint32_t*)didx));
    BXG_TRY(d2h_sync(c, c->h_small, c->d_small, 4));
    std::memcpy(&live_leaves, c->h_small, 4);
  }
  // 2. a right-sized leaf pool holding the live leaves at the front
  const uint64_t new_cap = ((uint64_t)live_leaves * 5 / 4 + 1024 + 255) & ~255ull;
  const uint64_t lw = m->h.leaf_words, cb = (uint64_t)m->h.leaf_size * m->h.value_bytes;
  const uint64_t tb = touched_bytes(m);
  void *masks = nullptr, *cells = nullptr, *orig = nullptr, *touched = nullptr;
  BXG_TRY(dmalloc(c, &masks, new_cap * lw * 8));
  if (cb) BXG_TRY(dmalloc(c, &cells, new_cap * cb));
  BXG_TRY(dmalloc(c, &orig, new_cap * sizeof(int4)));
  if (tb) BXG_TRY(dmalloc(c, &touched, new_cap * tb));
  if (tb) BXG_HIP(c, hipMemsetAsync(touched, 0, new_cap * tb, c->stream));
  if (nl)
    BXG_TRY(launch(c, "k_compact_leaves", k_compact_leaves, nl, kBlock, (const MapDev*)m->d, nl,
                   (const uint32_t*)dlive, (const uint32_t*)didx, (uint64_t*)masks,
                   (uint8_t*)cells, (int4*)orig));
  // 3. a directory without the roots left empty, sized for what remains
  BXG_TRY(ctx_sync(c));
  uint64_t live_roots = m->known_roots;
  const uint64_t rcap = pow2_at_least(std::max<uint64_t>(live_roots * 2, 1024));
  void *keys = nullptr, *children = nullptr;
  BXG_TRY(dmalloc(c, &keys, rcap * 8));
  BXG_TRY(dmalloc(c, &children, rcap * m->h.inner_size * 4));
  BXG_TRY(fill64(c, (uint64_t*)keys, rcap, kEmptyKey));
  BXG_TRY(fill32(c, (uint32_t*)children, rcap * m->h.inner_size, kEmptyChild));
  uint64_t* old_keys = m->h.root_keys;
  uint32_t* old_children = m->h.children;
  const uint32_t old_rcap = m->root_cap;
  uint64_t* old_masks = m->h.leaf_masks;
  uint8_t* old_cells = m->h.cells;
  int4* old_orig = m->h.leaf_origin;
  uint8_t* old_touched = m->h.touched;
  const uint64_t old_bytes = (uint64_t)m->leaf_cap * leaf_bytes(m) + (uint64_t)old_rcap * root_bytes(m);
  m->h.leaf_masks = (uint64_t*)masks;
  m->h.cells = (uint8_t*)cells;
  m->h.leaf_origin = (int4*)orig;
  m->h.touched = (uint8_t*)touched;
  m->h.leaf_cap = (uint32_t)new_cap;
  m->leaf_cap = (uint32_t)new_cap;
  m->h.root_keys = (uint64_t*)keys;
  m->h.children = (uint32_t*)children;
  m->h.root_mask = (uint32_t)(rcap - 1);
  m->h.max_probe = (uint32_t)std::min<uint64_t>(rcap, 512);
  m->root_cap = (uint32_t)rcap;
  BXG_TRY(push_header(m));
  const uint32_t z2[2] = {live_leaves, 0};  // leaf_count, root_count
  BXG_HIP(c, hipMemcpyAsync(&m->ctr->leaf_count, z2, 8, hipMemcpyHostToDevice, c->stream));
  BXG_TRY(launch(c, "k_rehash", k_rehash, nblocks(old_rcap), kBlock, (const MapDev*)m->d,
                 (const uint64_t*)old_keys, (const uint32_t*)old_children, old_rcap, 1u));
  BXG_TRY(ctx_sync(c));
  dfree(old_keys);
  dfree(old_children);
  dfree(old_masks);
  dfree(old_cells);
  dfree(old_orig);
  dfree(old_touched);
  const uint64_t new_bytes = new_cap * leaf_bytes(m) + rcap * root_bytes(m);
  c->used = c->used - old_bytes + new_bytes;
  m->bytes = m->bytes - old_bytes + new_bytes;
  return read_counters(m);
}

// ------------------------------------------------------ bulk VoxelGrid

int bxg_set_values(bxg_map* m, const int32_t* coords, const void* vals, uint64_t n, uint8_t* was_on) {
  if (!m) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  if (m->h.value_bytes && n && !vals) return fail(m->ctx, BXG_ERR_INVALID_ARG, "null values");
  return cell_op(m, coords, vals, n, was_on, kOpSetValue);
}
int bxg_set_cells_on(bxg_map* m, const int32_t* coords, const void* defaults, uint64_t n,
                     uint8_t* was_on) {
  if (!m) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  return cell_op(m, coords, defaults, n, was_on, kOpSetOn);
}
int bxg_set_cells_off(bxg_map* m, const int32_t* coords, uint64_t n, uint8_t* was_on) {
  if (!m) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  return cell_op(m, coords, nullptr, n, was_on, kOpSetOff);
}
int bxg_lookup(bxg_map* m, const int32_t* coords, uint64_t n, void* vals, uint8_t* found) {
  if (!m) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  BXG_TRY(check_coords_args(m, coords, n));
  if (!n) return BXG_OK;
  return launch(m->ctx, "k_lookup", k_lookup, lookup_blocks(n), kBlock, (const MapDev*)m->d, coords, n,
                (uint8_t*)vals, found);
}
int bxg_classify(bxg_map* m, const int32_t* coords, uint64_t n, int8_t* state) {
  if (!m) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  BXG_TRY(check_coords_args(m, coords, n));
  if (!is_prob(m)) return fail(m->ctx, BXG_ERR_INVALID_ARG, "classify needs a probabilistic map");
  if (n && !state) return fail(m->ctx, BXG_ERR_INVALID_ARG, "null state buffer");
  if (!n) return BXG_OK;
  return launch(m->ctx, "k_classify_cells", k_classify_cells, nblocks(n), kBlock,
                (const MapDev*)m->d, coords, n, state);
}

int bxg_export(bxg_map* m, int filter, int32_t* coords, void* vals, uint64_t cap, uint64_t* count) {
  if (!m || !count) return BXG_ERR_INVALID_ARG;
  bxg_ctx* c = m->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  BXG_TRY(settle(m));
  if (filter < 0 || filter > 2) return fail(c, BXG_ERR_INVALID_ARG, "unknown export filter");
  if (filter && !is_prob(m)) return fail(c, BXG_ERR_INVALID_ARG, "filtered export needs a probabilistic map");
  uint32_t nl = 0;
  BXG_TRY(known_leaf_count(m, &nl));
  BXG_HIP(c, hipMemsetAsync(c->d_small, 0, 8, c->stream));
  BXG_TRY(launch(c, "k_export", k_export, nblocks((uint64_t)nl * m->h.leaf_size), kBlock,
                 (const MapDev*)m->d, nl, (uint32_t)filter, coords, (uint8_t*)vals, cap,
                 c->d_small));
  BXG_TRY(d2h_sync(c, c->h_small, c->d_small, 8));
  *count = c->h_small[0];
  return *count > cap ? fail(c, BXG_ERR_CAPACITY, "export buffer too small") : BXG_OK;
}

// host forms ---------------------------------------------------------

static int stage_in(bxg_ctx* c, int slot, const void* src, uint64_t bytes, void** dev) {
  BXG_TRY(dstage(c, slot, std::max<uint64_t>(bytes, 64), dev));
  if (bytes && src) BXG_HIP(c, hipMemcpyAsync(*dev, src, bytes, hipMemcpyHostToDevice, c->stream));
  return BXG_OK;
}

static int host_cell_op(bxg_map* m, const int32_t* coords, const void* vals, uint64_t n,
                        uint8_t* was_on, uint32_t op) {
  bxg_ctx* c = m->ctx;
  BXG_TRY(check_coords_args(m, coords, n));
  if (!n) return BXG_OK;
  BXG_TRY(map_failed(m));
  const uint32_t vb = m->h.value_bytes;
  void *dc, *dv = nullptr, *dw = nullptr;
  BXG_TRY(stage_in(c, 0, coords, n * 12, &dc));
  if (vals && vb) BXG_TRY(stage_in(c, 1, vals, n * vb, &dv));
  if (was_on) BXG_TRY(stage_in(c, 2, nullptr, n, &dw));
  for (int attempt = 0;; ++attempt) {
    BXG_TRY(cell_op(m, (const int32_t*)dc, dv, n, (uint8_t*)dw, op));
    BXG_TRY(read_counters(m));
    if (!m->h_ctr->status) break;
    if (m->h_ctr->status & kStCoordRange) {
      BXG_TRY(clear_status(m));
      return fail(c, BXG_ERR_COORD_RANGE, "coordinate outside the supported range");
    }
    if (attempt > 8) return fail(c, BXG_ERR_OUT_OF_POOL, "could not grow the node pool");
    BXG_TRY(grow_after_overflow(m));
  }
  if (was_on) BXG_TRY(d2h_sync(c, was_on, dw, n));
  return BXG_OK;
}

int bxg_set_values_host(bxg_map* m, const int32_t* coords, const void* vals, uint64_t n,
                        uint8_t* was_on) {
  if (!m) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  if (m->h.value_bytes && n && !vals) return fail(m->ctx, BXG_ERR_INVALID_ARG, "null values");
  return host_cell_op(m, coords, vals, n, was_on, kOpSetValue);
}
int bxg_set_cells_on_host(bxg_map* m, const int32_t* coords, const void* defaults, uint64_t n,
                          uint8_t* was_on) {
  if (!m) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  return host_cell_op(m, coords, defaults, n, was_on, kOpSetOn);
}
int bxg_set_cells_off_host(bxg_map* m, const int32_t* coords, uint64_t n, uint8_t* was_on) {
  if (!m) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  return host_cell_op(m, coords, nullptr, n, was_on, kOpSetOff);
}
int bxg_lookup_host(bxg_map* m, const int32_t* coords, uint64_t n, void* vals, uint8_t* found) {
  if (!m) return BXG_ERR_INVALID_ARG;
  bxg_ctx* c = m->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  BXG_TRY(check_coords_args(m, coords, n));
  if (!n) return BXG_OK;
  const uint32_t vb = m->h.value_bytes;
  void *dc, *dv = nullptr, *df = nullptr;
  BXG_TRY(stage_in(c, 0, coords, n * 12, &dc));
  if (vals && vb) BXG_TRY(stage_in(c, 1, nullptr, n * vb, &dv));
  if (found) BXG_TRY(stage_in(c, 2, nullptr, n, &df));
  BXG_TRY(launch(c, "k_lookup", k_lookup, lookup_blocks(n), kBlock, (const MapDev*)m->d,
                 (const int32_t*)dc, n, (uint8_t*)dv, (uint8_t*)df));
  if (dv) BXG_HIP(c, hipMemcpyAsync(vals, dv, n * vb, hipMemcpyDeviceToHost, c->stream));
  if (df) BXG_HIP(c, hipMemcpyAsync(found, df, n, hipMemcpyDeviceToHost, c->stream));
  return ctx_sync(c);
}
int bxg_classify_host(bxg_map* m, const int32_t* coords, uint64_t n, int8_t* state) {
  if (!m) return BXG_ERR_INVALID_ARG;
  bxg_ctx* c = m->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  BXG_TRY(check_coords_args(m, coords, n));
  if (!n) return BXG_OK;
  void *dc, *ds;
  BXG_TRY(stage_in(c, 0, coords, n * 12, &dc));
  BXG_TRY(stage_in(c, 2, nullptr, n, &ds));
  BXG_TRY(bxg_classify(m, (const int32_t*)dc, n, (int8_t*)ds));
  return d2h_sync(c, state, ds, n);
}
int bxg_export_host(bxg_map* m, int filter, int32_t* coords, void* vals, uint64_t cap,
                    uint64_t* count) {
  if (!m || !count) return BXG_ERR_INVALID_ARG;
  bxg_ctx* c = m->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  BXG_TRY(settle(m));
  const uint32_t vb = m->h.value_bytes;
  void *dc = nullptr, *dv = nullptr;
  if (coords && cap) BXG_TRY(dstage(c, 0, cap * 12, &dc));
  if (vals && vb && cap) BXG_TRY(dstage(c, 1, cap * vb, &dv));
  const int rc = bxg_export(m, filter, (int32_t*)dc, dv, cap, count);
  if (rc != BXG_OK && rc != BXG_ERR_CAPACITY) return rc;
  const uint64_t k = std::min(*count, cap);
  if (dc && k) BXG_HIP(c, hipMemcpyAsync(coords, dc, k * 12, hipMemcpyDeviceToHost, c->stream));
  if (dv && k) BXG_HIP(c, hipMemcpyAsync(vals, dv, k * vb, hipMemcpyDeviceToHost, c->stream));
  BXG_TRY(ctx_sync(c));
  return rc;
}

int bxg_read_leaves_host(bxg_map* m, const int32_t* origins, uint64_t n, uint64_t* masks,
                         void* cells, uint8_t* exists) {
  if (!m) return BXG_ERR_INVALID_ARG;
  bxg_ctx* c = m->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  BXG_TRY(settle(m));
  if (!n) return BXG_OK;
  if (!origins || !masks || !exists || (m->h.value_bytes && !cells) || n >= (1u << 31))
    return fail(c, BXG_ERR_INVALID_ARG, "bad leaf read arguments");
  const uint64_t mb = (uint64_t)m->h.leaf_words * 8, cb = (uint64_t)m->h.leaf_size * m->h.value_bytes;
  void *dorig, *dm, *dc, *de;
  BXG_TRY(stage_in(c, 0, origins, n * 12, &dorig));
  BXG_TRY(stage_in(c, 1, nullptr, n * mb, &dm));
  BXG_TRY(stage_in(c, 2, nullptr, std::max<uint64_t>(n * cb, 64), &dc));
  BXG_TRY(stage_in(c, 3, nullptr, n, &de));
  BXG_TRY(launch(c, "k_read_leaves", k_read_leaves, (uint32_t)n, kBlock, (const MapDev*)m->d,
                 (const int32_t*)dorig, (uint32_t)n, (uint64_t*)dm, (uint8_t*)dc, (uint8_t*)de));
  BXG_HIP(c, hipMemcpyAsync(masks, dm, n * mb, hipMemcpyDeviceToHost, c->stream));
  if (cb) BXG_HIP(c, hipMemcpyAsync(cells, dc, n * cb, hipMemcpyDeviceToHost, c->stream));
  BXG_HIP(c, hipMemcpyAsync(exists, de, n, hipMemcpyDeviceToHost, c->stream));
  return ctx_sync(c);
}

int bxg_write_leaves_host(bxg_map* m, const int32_t* origins, uint64_t n, const uint64_t* masks,
                          const void* cells) {
  if (!m) return BXG_ERR_INVALID_ARG;
  bxg_ctx* c = m->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  if (!n) return BXG_OK;
  if (!origins || !masks || (m->h.value_bytes && !cells) || n >= (1u << 31))
    return fail(c, BXG_ERR_INVALID_ARG, "bad leaf write arguments");
  BXG_TRY(map_failed(m));
  const uint64_t mb = (uint64_t)m->h.leaf_words * 8, cb = (uint64_t)m->h.leaf_size * m->h.value_bytes;
  void *dorig, *dm, *dc = nullptr;
  BXG_TRY(stage_in(c, 0, origins, n * 12, &dorig));
  BXG_TRY(stage_in(c, 1, masks, n * mb, &dm));
  if (cb) BXG_TRY(stage_in(c, 2, cells, n * cb, &dc));
  for (int attempt = 0;; ++attempt) {
    BXG_TRY(ensure_room(m, n, std::min<uint64_t>(n, 1u << 16)));
    BXG_TRY(launch(c, "k_touch", k_touch, nblocks(n), kBlock, (const MapDev*)m->d,
                   (const int32_t*)dorig, n));
    BXG_TRY(read_counters(m));
    if (!m->h_ctr->status) break;
    if (m->h_ctr->status & kStCoordRange) {
      BXG_TRY(clear_status(m));
      return fail(c, BXG_ERR_COORD_RANGE, "coordinate outside the supported range");
    }
    if (attempt > 8) return fail(c, BXG_ERR_OUT_OF_POOL, "could not grow the node pool");
    BXG_TRY(grow_after_overflow(m));
  }
  BXG_TRY(launch(c, "k_write_leaves", k_write_leaves, (uint32_t)n, kBlock, (const MapDev*)m->d,
                 (const int32_t*)dorig, (uint32_t)n, (const uint64_t*)dm, (const uint8_t*)dc));
  BXG_TRY(snapshot_counters(m));
  return ctx_sync(c);
}

int bxg_leaf_origins_host(bxg_map* m, int32_t* origins, uint64_t cap, uint64_t* count) {
  if (!m || !count) return BXG_ERR_INVALID_ARG;
  bxg_ctx* c = m->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  BXG_TRY(settle(m));
  uint32_t nl = 0;
  BXG_TRY(known_leaf_count(m, &nl));
  *count = nl;
  if (!origins || !cap || !nl) return nl > cap ? fail(c, BXG_ERR_CAPACITY, "origin buffer too small") : BXG_OK;
  void* d;
  BXG_TRY(dstage(c, 0, (uint64_t)nl * 12, &d));
  BXG_TRY(launch(c, "k_leaf_origins", k_leaf_origins, nblocks(nl), kBlock, (const MapDev*)m->d, nl,
                 (int32_t*)d));
  BXG_TRY(d2h_sync(c, origins, d, std::min<uint64_t>(nl, cap) * 12));
  return nl > cap ? fail(c, BXG_ERR_CAPACITY, "origin buffer too small") : BXG_OK;
}

// ------------------------------------------------------ probabilistic

int bxg_set_options(bxg_map* m, const bxg_prob_options* o) {
  if (!m || !o) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  BXG_TRY(settle(m));
  m->h.opt_miss = o->prob_miss_log;
  m->h.opt_hit = o->prob_hit_log;
  m->h.opt_cmin = o->clamp_min_log;
  m->h.opt_cmax = o->clamp_max_log;
  m->h.opt_thr = o->occupancy_threshold_log;
  return push_header(m);
}
int bxg_get_options(bxg_map* m, bxg_prob_options* o) {
  if (!m || !o) return BXG_ERR_INVALID_ARG;
  o->prob_miss_log = m->h.opt_miss;
  o->prob_hit_log = m->h.opt_hit;
  o->clamp_min_log = m->h.opt_cmin;
  o->clamp_max_log = m->h.opt_cmax;
  o->occupancy_threshold_log = m->h.opt_thr;
  return BXG_OK;
}
int bxg_get_update_count(bxg_map* m, uint32_t* out) {
  if (!m || !out) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  BXG_TRY(settle(m));
  BXG_TRY(read_counters(m));
  *out = (uint32_t)m->h_ctr->cur;
  return BXG_OK;
}
int bxg_set_update_count(bxg_map* m, uint32_t v) {
  if (!m || v < 1 || v > 3) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  BXG_TRY(settle(m));
  const int32_t cur = (int32_t)v;
  BXG_HIP(m->ctx, hipMemcpyAsync(&m->ctr->cur, &cur, 4, hipMemcpyHostToDevice, m->ctx->stream));
  return ctx_sync(m->ctx);
}

}  // extern "C"

namespace {

// one scan, enqueued (the device-pointer insert entry points)
// An insert of device points: enqueued and remembered until the map is next
// confirmed (settle), which re-runs it should a pool overflow have skipped
// it; the caller keeps the points valid until then.  An insert that also
// traces rays due from addHit/MissPoint is checked at once (those rays are
// consumed by the call).
int insert_async(bxg_map* m, const bxg_scan_desc& s) {
  bxg_ctx* c = m->ctx;
  if (!m->extra.empty()) {
    BXG_TRY(settle(m));
    for (int attempt = 0;; ++attempt) {
      BXG_TRY(run_batch(c, &s, 1, 0));
      BXG_TRY(read_counters(m));
      const uint32_t st = m->h_ctr->status;
      if (!(st & (kStOutOfPool | kStRootFull))) break;
      if (attempt > 8) return fail(c, BXG_ERR_OUT_OF_POOL, "could not grow the node pool");
      BXG_TRY(grow_after_overflow(m));
    }
    BXG_TRY(map_status(m));
    m->extra.clear();
    return BXG_OK;
  }
  BXG_TRY(pending_begin(m));
  BXG_TRY(run_batch(c, &s, 1, 0));
  pending_push(m, s);
  if (m->pending.size() >= kMaxPendingInserts) BXG_TRY(settle(m));
  return BXG_OK;
}

// one scan of host points: staged, run, re-run after growing the pools if
// the map ran out of nodes (a failed run wrote no cell)
int insert_host(bxg_map* m, bxg_scan_desc s, const void* points) {
  bxg_ctx* c = m->ctx;
  BXG_TRY(map_failed(m));
  const uint64_t packed = (s.flags & BXG_POINTS_F64) ? 24 : 12;
  const uint64_t bytes = s.n ? (s.n - 1) * (s.stride ? s.stride : packed) + packed : 0;
  void* dp;
  BXG_TRY(stage_in(c, 3, points, bytes, &dp));
  s.points_dev = dp;
  for (int attempt = 0;; ++attempt) {
    BXG_TRY(run_batch(c, &s, 1, 0));
    BXG_TRY(read_counters(m));
    const uint32_t st = m->h_ctr->status;
    if (!st) break;
    if (st & (kStCoordRange | kStScratchFull)) {
      BXG_TRY(clear_status(m));
      return fail(c, status_code(st), "insertPointCloud failed: coordinate outside the supported range");
    }
    if (attempt > 8) return fail(c, BXG_ERR_OUT_OF_POOL, "could not grow the node pool");
    BXG_TRY(grow_after_overflow(m));
  }
  m->extra.clear();
  return BXG_OK;
}

bxg_scan_desc scan_of(bxg_map* m, const void* points, uint64_t n, uint32_t flags, uint32_t stride,
                      const double* origin, const float* tf, double max_range) {
  bxg_scan_desc s;
  std::memset(&s, 0, sizeof(s));
  s.map = m;
  s.points_dev = points;
  s.n = n;
  s.flags = flags;
  s.stride = stride;
  if (origin) std::memcpy(s.origin, origin, sizeof(s.origin));
  s.sensor_to_world = tf;
  s.max_range = max_range;
  return s;
}

}  // namespace

extern "C" {

int bxg_insert_point_cloud(bxg_map* m, const void* points, uint64_t n, uint32_t flags,
                           const double origin[3], double max_range) {
  if (!m || !origin) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  return insert_async(m, scan_of(m, points, n, flags, 0, origin, nullptr, max_range));
}

int bxg_insert_point_cloud_tf(bxg_map* m, const float* points, uint64_t n, uint32_t stride,
                              const float sensor_to_world[16], double max_range) {
  if (!m || !sensor_to_world) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  return insert_async(m, scan_of(m, points, n, 0, stride, nullptr, sensor_to_world, max_range));
}

int bxg_insert_point_cloud_noalloc(bxg_map* m, const void* points, uint64_t n, uint32_t flags,
                                   const double origin[3], double max_range) {
  if (!m || !origin) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  if (n > kNaPoints)
    return fail(m->ctx, BXG_ERR_INVALID_ARG,
                "more points than the allocation-free scratch holds (BXG_NOALLOC_MAX_POINTS = " +
                    std::to_string(kNaPoints) + ")");
  return insert_noalloc(m, scan_of(m, points, n, flags, 0, origin, nullptr, max_range));
}

// Run one wave, then re-run (alone, after growing its pools) every scan
// whose map ran out of nodes; a failed scan left its map's cells untouched.
static int run_wave_checked(bxg_ctx* c, const bxg_scan_desc* scans, uint32_t nscans) {
  BXG_TRY(run_wave(c, scans, nscans, 0));
  BXG_TRY(ctx_sync(c));
  for (uint32_t s = 0; s < nscans; ++s) {
    bxg_map* m = scans[s].map;
    for (int attempt = 0;; ++attempt) {
      BXG_TRY(read_counters(m));
      const uint32_t st = m->h_ctr->status;
      if (!st) break;
      if (st & (kStCoordRange | kStScratchFull)) {
        BXG_TRY(clear_status(m));
        return fail(c, status_code(st), "insert failed: coordinate outside the supported range");
      }
      if (attempt > 8) return fail(c, BXG_ERR_OUT_OF_POOL, "could not grow the node pool");
      BXG_TRY(grow_after_overflow(m));
      BXG_TRY(run_wave(c, scans + s, 1, 0));
    }
    m->extra.clear();
  }
  return BXG_OK;
}

int bxg_insert_batch(bxg_ctx* c, const bxg_scan_desc* scans, uint32_t nscans) {
  if (!c || (nscans && !scans)) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  for (uint32_t s = 0; s < nscans; ++s) {
    if (!scans[s].map) return fail(c, BXG_ERR_INVALID_ARG, "scan without a map");
    BXG_TRY(map_failed(scans[s].map));
  }
  uint32_t start = 0;
  while (start < nscans) {
    uint32_t end = start;
    std::vector<bxg_map*> seen;
    while (end < nscans && std::find(seen.begin(), seen.end(), scans[end].map) == seen.end()) {
      seen.push_back(scans[end].map);
      ++end;
    }
    BXG_TRY(run_wave_checked(c, scans + start, end - start));
    start = end;
  }
  return BXG_OK;
}

int bxg_insert_point_cloud_host(bxg_map* m, const void* points, uint64_t n, uint32_t flags,
                                const double origin[3], double max_range) {
  if (!m || !origin || (n && !points)) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  return insert_host(m, scan_of(m, nullptr, n, flags, 0, origin, nullptr, max_range), points);
}

int bxg_insert_point_cloud_tf_host(bxg_map* m, const float* points, uint64_t n, uint32_t stride,
                                   const float sensor_to_world[16], double max_range) {
  if (!m || !sensor_to_world || (n && !points)) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(m->ctx->mu);
  return insert_host(m, scan_of(m, nullptr, n, 0, stride, nullptr, sensor_to_world, max_range), points);
}

int bxg_add_point_host(bxg_map* m, const double xyz[3], int is_hit) {
  if (!m || !xyz) return BXG_ERR_INVALID_ARG;
  bxg_ctx* c = m->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  if (!is_prob(m)) return fail(c, BXG_ERR_INVALID_ARG, "addHit/MissPoint needs a probabilistic map");
  BXG_TRY(map_failed(m));
  int32_t e[3];
  BXG_TRY(origin_voxel(m, xyz, e));  // posToCoord with the range rule
  void* dp;
  BXG_TRY(stage_in(c, 3, xyz, 24, &dp));
  bxg_scan_desc s;
  std::memset(&s, 0, sizeof(s));
  s.map = m;
  s.points_dev = dp;
  s.n = 1;
  s.flags = BXG_POINTS_F64;
  std::memcpy(s.origin, xyz, sizeof(s.origin));
  s.max_range = 0.0;
  for (int attempt = 0;; ++attempt) {
    BXG_TRY(read_counters(m));
    const uint64_t before = m->h_ctr->stat_endpoint_updates;
    BXG_TRY(run_batch(c, &s, 1, kScanNoRays | (is_hit ? kScanForceHit : kScanForceMiss)));
    BXG_TRY(read_counters(m));
    const uint32_t st = m->h_ctr->status;
    if (!st) {
      if (m->h_ctr->stat_endpoint_updates != before) m->extra.push_back(make_int4(e[0], e[1], e[2], 0));
      break;
    }
    if (st & (kStCoordRange | kStScratchFull)) {
      BXG_TRY(clear_status(m));
      return fail(c, status_code(st), "addHit/MissPoint failed");
    }
    if (attempt > 8) return fail(c, BXG_ERR_OUT_OF_POOL, "could not grow the node pool");
    BXG_TRY(grow_after_overflow(m));
  }
  return BXG_OK;
}

// ---------------------------------------------------------- graphs

int bxg_graph_create(bxg_ctx* c, uint32_t max_scans, uint64_t max_points, bxg_graph** out) {
  if (!c || !out || !max_scans || !max_points) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  *out = nullptr;
  if (max_points >= (1ull << 31) || max_scans >= 65535)
    return fail(c, BXG_ERR_INVALID_ARG, "graph capacity too large");
  bxg_graph* g = new bxg_graph();
  g->ctx = c;
  g->max_scans = max_scans;
  g->max_points = max_points;
  const uint64_t cap = pow2_at_least(std::max<uint64_t>(max_points, 4096));
  const uint64_t tcap = cap * c->ep_table_factor;
  void* p = nullptr;
  int rc = dmalloc(c, &p, cap * sizeof(int4));
  g->sc.rec = (int4*)p;
  if (rc == BXG_OK) rc = dmalloc(c, &p, cap * 4), g->sc.tpos = (uint32_t*)p;
  if (rc == BXG_OK) rc = dmalloc(c, &p, cap * 4), g->sc.winners = (uint32_t*)p;
  if (rc == BXG_OK) rc = dmalloc(c, &p, cap * 4), g->sc.wkey = (uint32_t*)p;
  if (rc == BXG_OK) rc = dmalloc(c, &p, (uint64_t)kMaxBuckets * 4), g->sc.hist = (uint32_t*)p;
  if (rc == BXG_OK) rc = dmalloc(c, &p, (uint64_t)kMaxBuckets * 4), g->sc.offs = (uint32_t*)p;
  if (rc == BXG_OK) rc = fill32(c, g->sc.hist, kMaxBuckets, 0);
  if (rc == BXG_OK) rc = dmalloc(c, &p, tcap * sizeof(EpSlot)), g->sc.ep = (EpSlot*)p;
  if (rc == BXG_OK) rc = dmalloc(c, &p, kCounterWords * 4), g->d_counters = (uint32_t*)p;
  if (rc == BXG_OK) rc = dmalloc(c, &p, (uint64_t)max_scans * sizeof(ScanDesc)), g->d_desc = (ScanDesc*)p;
  if (rc == BXG_OK) rc = dmalloc(c, &p, sizeof(BatchHdr)), g->d_hdr = (BatchHdr*)p;
  g->sc.ep_mask = (uint32_t)(tcap - 1);
  g->sc.counters = g->d_counters;
  if (rc == BXG_OK) rc = fill64(c, (uint64_t*)g->sc.ep, tcap * 2, ~0ull);  // empty: all ones
  if (rc == BXG_OK) rc = fill32(c, g->d_counters, kCounterWords, 0);
  if (rc == BXG_OK) {
    BatchHdr hdr{0, 0, 0};
    rc = h2d_sync(c, g->d_hdr, &hdr, sizeof(hdr));
  }
  if (rc == BXG_OK) {
    // capture the pipeline once; the kernels read the frame size from d_hdr
    const bool prof = c->profiling;
    c->profiling = false;
    hipError_t e = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) {
      rc = hip_fail(c, e, "hipStreamBeginCapture");
    } else {
      rc = enqueue_scan_kernels(c, g->d_hdr, g->d_desc, g->sc, max_points, max_scans, false);
      hipGraph_t graph = nullptr;
      e = hipStreamEndCapture(c->stream, &graph);
      if (rc == BXG_OK && e != hipSuccess) rc = hip_fail(c, e, "hipStreamEndCapture");
      g->graph = graph;
      if (rc == BXG_OK) {
        e = hipGraphInstantiate(&g->exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = hip_fail(c, e, "hipGraphInstantiate");
      }
    }
    c->profiling = prof;
  }
  if (rc != BXG_OK) {
    bxg_graph_destroy(g);
    return rc;
  }
  *out = g;
  return BXG_OK;
}

int bxg_graph_set_scans(bxg_graph* g, const bxg_scan_desc* scans, uint32_t nscans) {
  if (!g || (nscans && !scans)) return BXG_ERR_INVALID_ARG;
  bxg_ctx* c = g->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  if (nscans > g->max_scans) return fail(c, BXG_ERR_INVALID_ARG, "more scans than the graph holds");
  // the previous frame's scans may still be pending on their maps; they
  // hold their own copies of their transforms, so the frame can change
  // without waiting for them (the descriptor upload is stream-ordered)
  std::vector<ScanDesc> descs;
  uint64_t total = 0;
  BXG_TRY(build_descs(c, scans, nscans, 0u, descs, &total));
  if (total > g->max_points) return fail(c, BXG_ERR_INVALID_ARG, "more points than the graph holds");
  std::vector<bxg_map*> maps;
  for (uint32_t s = 0; s < nscans; ++s) {
    bxg_map* m = scans[s].map;
    if (std::find(maps.begin(), maps.end(), m) != maps.end())
      return fail(c, BXG_ERR_INVALID_ARG, "a graph frame holds one scan per map");
    if (!m->extra.empty())
      return fail(c, BXG_ERR_NOT_SUPPORTED, "graph frames do not trace addHit/MissPoint rays");
    descs[s].extra = nullptr;
    descs[s].nextra = 0;
    maps.push_back(m);
  }
  BatchHdr hdr{nscans, tag_bits_for(nscans), total};
  BXG_TRY(upload(c, g->d_desc, descs.data(), descs.size() * sizeof(ScanDesc)));
  BXG_TRY(upload(c, g->d_hdr, &hdr, sizeof(hdr)));
  g->maps = maps;
  g->npoints.clear();
  g->ranges.clear();
  g->scans.assign(scans, scans + nscans);
  g->tfs.assign(nscans, std::array<float, 16>{});
  for (uint32_t s = 0; s < nscans; ++s) {
    g->npoints.push_back(scans[s].n);
    g->ranges.push_back(scans[s].max_range);
    if (scans[s].sensor_to_world) {
      std::memcpy(g->tfs[s].data(), scans[s].sensor_to_world, sizeof(float) * 16);
      g->scans[s].sensor_to_world = g->tfs[s].data();
    }
  }
  g->loaded = true;
  return BXG_OK;
}

int bxg_graph_launch(bxg_graph* g) {
  if (!g) return BXG_ERR_INVALID_ARG;
  bxg_ctx* c = g->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  if (!g->loaded) return fail(c, BXG_ERR_INVALID_ARG, "no frame loaded");
  for (bxg_map* m : g->maps) BXG_TRY(pending_begin(m));
  for (size_t s = 0; s < g->maps.size(); ++s)
    BXG_TRY(ensure_room(g->maps[s], scan_leaf_estimate(g->maps[s], g->npoints[s], g->ranges[s]), 1024));
  BXG_HIP(c, hipGraphLaunch(g->exec, c->stream));
  // the frame's k_scan_finalize wrote each map's counters to its mirror
  BXG_TRY(snapshot_event(c, g->maps.data(), g->maps.size()));
  // each map keeps its scan of the frame until it is confirmed (settle),
  // which re-runs it should the frame have overflowed the map's pools
  for (size_t s = 0; s < g->maps.size(); ++s) {
    bxg_map* m = g->maps[s];
    pending_push(m, g->scans[s]);
    if (m->pending.size() >= kMaxPendingInserts) BXG_TRY(settle(m));
  }
  return BXG_OK;
}

int bxg_graph_destroy(bxg_graph* g) {
  if (!g) return BXG_ERR_INVALID_ARG;
  bxg_ctx* c = g->ctx;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  hipStreamSynchronize(c->stream);  // the scratch below may still be in use
  if (g->exec) hipGraphExecDestroy(g->exec);
  if (g->graph) hipGraphDestroy(g->graph);
  dfree(g->sc.rec);
  dfree(g->sc.tpos);
  dfree(g->sc.winners);
  dfree(g->sc.wkey);
  dfree(g->sc.hist);
  dfree(g->sc.offs);
  dfree(g->sc.ep);
  dfree(g->d_counters);
  dfree(g->d_desc);
  dfree(g->d_hdr);
  delete g;
  return BXG_OK;
}

}  // extern "C"

#include "bxg_server.hip"
