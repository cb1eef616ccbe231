// bxg_server.hip -- the serving engine: many maps resident in HBM, insert /
// lookup / export requests admitted continuously while others run.
// Included by bxg_api.hip (it uses the runtime's internal helpers).
//
// Scheduling:
//   - every map has a FIFO of requests and at most one request on the GPU, so
//     each request sees exactly the requests submitted before it on its map
//     (the result of a request is the same alone and in any mix);
//   - requests of different maps overlap: the inserts admitted together form
//     one scan wave on the context stream (one pass over their flattened
//     point queue), lookups and exports run on the engine's streams beside it;
//   - admission: a request needs a free in-flight slot; an insert also needs
//     its map's pools grown to the scan's head-room estimate within the
//     context's memory budget, else it waits until memory frees up (it fails
//     with BXG_ERR_BUDGET when nothing is left in flight to wait for);
//   - a scan that still runs out of nodes on the device changed nothing: its
//     map's pools grow and the request is re-queued at the head of its FIFO.
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>
#include <unordered_map>

namespace {

uint64_t now_ns() {
  return (uint64_t)std::chrono::duration_cast<std::chrono::nanoseconds>(
             std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

enum : uint32_t { kReqInsert = 0, kReqLookup = 1, kReqExport = 2 };

struct SrvReq {
  uint64_t id = 0;
  uint32_t kind = kReqInsert;
  bxg_map* map = nullptr;
  bxg_scan_desc scan{};  // insert
  bool had_extra = false;
  const int32_t* coords = nullptr;  // lookup
  void* vals = nullptr;
  uint8_t* found = nullptr;
  uint64_t n = 0;
  int filter = 0;  // export
  int32_t* out_coords = nullptr;
  uint64_t cap = 0;
  int status = -1;  // -1 while pending
  uint64_t count = 0;
  uint64_t t_submit = 0, t_admit = 0, t_done = 0;
  hipEvent_t ev = nullptr;
  int slot = -1;
  int attempts = 0;
};

}  // namespace

struct bxg_server {
  bxg_ctx* c = nullptr;
  bxg_server_options opt{};
  std::thread th;
  std::mutex mu;  // guards everything below up to `ninflight`
  std::condition_variable cv_work, cv_done;
  bool stop = false;
  uint64_t next_id = 1;
  std::unordered_map<uint64_t, std::unique_ptr<SrvReq>> reqs;
  std::unordered_map<bxg_map*, std::deque<SrvReq*>> queues;
  std::unordered_map<bxg_map*, bool> busy;
  std::vector<bxg_map*> order;  // maps in first-seen order, served round robin
  size_t rr = 0;
  size_t queued = 0;
  size_t ninflight = 0;
  std::vector<int> free_slots;  // export count slots
  bxg_server_stats stats{};
  // engine thread only
  std::vector<SrvReq*> inflight;
  std::vector<hipStream_t> streams;
  uint32_t next_stream = 0;
  std::vector<hipEvent_t> events;
  uint64_t* d_counts = nullptr;
  uint64_t* h_counts = nullptr;
};

namespace {

hipEvent_t srv_event(bxg_server* s) {
  if (!s->events.empty()) {
    hipEvent_t e = s->events.back();
    s->events.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
  return e;
}

// record a request's completion event on a stream and mark it in flight
int srv_track(bxg_server* s, SrvReq* r, hipStream_t st) {
  r->ev = srv_event(s);
  if (!r->ev) return fail(s->c, BXG_ERR_HIP, "hipEventCreate");
  BXG_HIP(s->c, hipEventRecord(r->ev, st));
  s->inflight.push_back(r);
  return BXG_OK;
}

// finish a request that never reached the GPU (or failed to launch)
void srv_finish(bxg_server* s, SrvReq* r, int status) {
  std::lock_guard<std::mutex> lk(s->mu);
  s->busy[r->map] = false;
  r->status = status;
  r->t_done = now_ns();
  ++s->stats.completed;
  if (status != BXG_OK) ++s->stats.failed;
  s->cv_done.notify_all();
}

// a lookup or an export on one of the engine's streams
void srv_launch_read(bxg_server* s, SrvReq* r) {
  bxg_ctx* c = s->c;
  bxg_map* m = r->map;
  int rc = BXG_OK;
  {
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    hipStream_t st = s->streams[s->next_stream++ % s->streams.size()];
    r->t_admit = now_ns();
    // inserts the map took outside the engine are confirmed first
    rc = settle(m);
    if (rc != BXG_OK) {
      // reported below
    } else if (r->kind == kReqLookup) {
      rc = launch_on(c, st, "k_lookup", k_lookup, lookup_blocks(r->n), kBlock, (const MapDev*)m->d, r->coords,
                     r->n, (uint8_t*)r->vals, r->found);
    } else {
      // the map is idle: its last counter snapshot has landed or is about to
      if (m->ctr_pending) hipEventSynchronize(m->snap->e);
      poll_counters(m);
      const uint32_t nl = (uint32_t)m->known_leaves;
      uint64_t* dc = s->d_counts + r->slot;
      rc = hipMemsetAsync(dc, 0, 8, st) == hipSuccess ? BXG_OK : fail(c, BXG_ERR_HIP, "hipMemsetAsync");
      if (rc == BXG_OK)
        rc = launch_on(c, st, "k_export", k_export, nblocks((uint64_t)nl * m->h.leaf_size), kBlock,
                       (const MapDev*)m->d, nl, (uint32_t)r->filter, r->out_coords, (uint8_t*)r->vals,
                       r->cap, dc);
      if (rc == BXG_OK && hipMemcpyAsync(s->h_counts + r->slot, dc, 8, hipMemcpyDeviceToHost, st) != hipSuccess)
        rc = fail(c, BXG_ERR_HIP, "hipMemcpyAsync");
    }
    if (rc == BXG_OK) rc = srv_track(s, r, st);
  }
  if (rc != BXG_OK) {
    if (r->slot >= 0) {
      std::lock_guard<std::mutex> lk(s->mu);
      s->free_slots.push_back(r->slot);
      r->slot = -1;
    }
    srv_finish(s, r, rc);
  }
}

// the admitted inserts as one scan wave; inserts whose pools cannot grow
// within the budget go back to the head of their queue (their count is
// returned)
size_t srv_launch_inserts(bxg_server* s, std::vector<SrvReq*>& ins) {
  bxg_ctx* c = s->c;
  std::vector<SrvReq*> go, back;
  std::vector<bxg_scan_desc> scans;
  int rc = BXG_OK;
  {
    std::lock_guard<std::recursive_mutex> lk(c->mu);
    for (SrvReq* r : ins) {
      bxg_map* m = r->map;
      // the map is idle, so its pools may move (inserts it took outside the
      // engine are confirmed first)
      int g = settle(m);
      if (g == BXG_OK) g = ensure_room(m, scan_leaf_estimate(m, r->scan.n, r->scan.max_range), 1024);
      if (g == BXG_ERR_BUDGET) {
        back.push_back(r);
      } else if (g != BXG_OK) {
        srv_finish(s, r, g);
      } else {
        r->had_extra = !m->extra.empty();
        r->t_admit = now_ns();
        go.push_back(r);
        scans.push_back(r->scan);
      }
    }
    if (!go.empty()) {
      rc = run_wave(c, scans.data(), (uint32_t)scans.size(), 0);
      for (size_t i = 0; i < go.size() && rc == BXG_OK; ++i) rc = srv_track(s, go[i], c->stream);
    }
    std::lock_guard<std::mutex> lk2(s->mu);
    if (!go.empty() && rc == BXG_OK) {
      ++s->stats.waves;
      s->stats.wave_scans += go.size();
    }
  }
  if (rc != BXG_OK) {
    // nothing of a wave that failed to launch is tracked: finish them all
    for (SrvReq* r : go) {
      auto it = std::find(s->inflight.begin(), s->inflight.end(), r);
      if (it != s->inflight.end()) continue;
      srv_finish(s, r, rc);
    }
  }
  if (!back.empty()) {
    std::lock_guard<std::mutex> lk(s->mu);
    const bool waiting_helps = !s->inflight.empty();
    for (auto it = back.rbegin(); it != back.rend(); ++it) {
      SrvReq* r = *it;
      s->busy[r->map] = false;
      if (waiting_helps) {
        s->queues[r->map].push_front(r);
        ++s->queued;
        ++s->stats.deferred;
      } else {
        r->status = BXG_ERR_BUDGET;
        r->t_done = now_ns();
        ++s->stats.completed;
        ++s->stats.failed;
      }
    }
    if (!waiting_helps) s->cv_done.notify_all();
    return waiting_helps ? back.size() : 0;
  }
  return 0;
}

// collect finished requests; true if any finished
bool srv_complete(bxg_server* s) {
  bool any = false;
  for (size_t i = 0; i < s->inflight.size();) {
    SrvReq* r = s->inflight[i];
    const hipError_t q = hipEventQuery(r->ev);
    if (q == hipErrorNotReady) {
      ++i;
      continue;
    }
    any = true;
    s->inflight[i] = s->inflight.back();
    s->inflight.pop_back();
    int st = q == hipSuccess ? BXG_OK : BXG_ERR_HIP;
    bool requeue = false;
    {
      std::lock_guard<std::recursive_mutex> lk(s->c->mu);
      if (st == BXG_OK && r->kind == kReqInsert) {
        bxg_map* m = r->map;
        poll_counters(m);  // the snapshot was enqueued before the event
        const uint32_t dst = m->h_ctr->status;
        if (dst && !(dst & kStCoordRange) && r->attempts < 8) {
          st = grow_after_overflow(m);
          requeue = st == BXG_OK;
        } else if (dst) {
          st = status_code(dst);
          clear_status(m);
        } else if (r->had_extra) {
          m->extra.clear();  // the addHit/MissPoint rays went with this scan
        }
      } else if (r->kind == kReqExport) {
        r->count = s->h_counts[r->slot];
        if (st == BXG_OK && r->count > r->cap) st = BXG_ERR_CAPACITY;
      }
    }
    s->events.push_back(r->ev);
    r->ev = nullptr;
    std::lock_guard<std::mutex> lk(s->mu);
    --s->ninflight;
    s->busy[r->map] = false;
    if (r->slot >= 0) {
      s->free_slots.push_back(r->slot);
      r->slot = -1;
    }
    if (requeue) {
      ++r->attempts;
      ++s->stats.retried;
      s->queues[r->map].push_front(r);
      ++s->queued;
    } else {
      r->status = st;
      r->t_done = now_ns();
      ++s->stats.completed;
      if (st != BXG_OK) ++s->stats.failed;
    }
  }
  if (any) s->cv_done.notify_all();
  return any;
}

// admit the head request of every idle map while slots last; true if a
// request reached the GPU or finished
bool srv_admit(bxg_server* s) {
  std::vector<SrvReq*> ins, reads;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    size_t slots = s->opt.max_inflight > s->ninflight ? s->opt.max_inflight - s->ninflight : 0;
    const size_t nm = s->order.size();
    for (size_t k = 0; k < nm && slots; ++k) {
      bxg_map* m = s->order[(s->rr + k) % nm];
      auto& q = s->queues[m];
      if (q.empty() || s->busy[m]) continue;
      SrvReq* r = q.front();
      if (r->kind == kReqInsert) {
        if (ins.size() >= s->opt.max_batch_scans) continue;
        ins.push_back(r);
      } else {
        if (r->kind == kReqExport) {
          if (s->free_slots.empty()) continue;
          r->slot = s->free_slots.back();
          s->free_slots.pop_back();
        }
        reads.push_back(r);
      }
      q.pop_front();
      --s->queued;
      s->busy[m] = true;
      ++s->ninflight;
      --slots;
    }
    if (nm) s->rr = (s->rr + 1) % nm;
  }
  if (ins.empty() && reads.empty()) return false;
  // requests that do not reach the GPU leave ninflight again
  const size_t before = s->inflight.size();
  for (SrvReq* r : reads) srv_launch_read(s, r);
  const size_t deferred = ins.empty() ? 0 : srv_launch_inserts(s, ins);
  const size_t tracked = s->inflight.size() - before;
  const size_t admitted = ins.size() + reads.size();
  if (tracked < admitted) {
    std::lock_guard<std::mutex> lk(s->mu);
    s->ninflight -= admitted - tracked;
  }
  return tracked + deferred < admitted || tracked > 0;
}

void srv_loop(bxg_server* s) {
  hipSetDevice(s->c->device);
  for (;;) {
    bool progress = srv_complete(s);
    progress = srv_admit(s) || progress;
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->ninflight == 0 && s->queued == 0) {
      if (s->stop) break;
      s->cv_work.wait(lk, [&] { return s->stop || s->queued > 0; });
      continue;
    }
    if (!progress) {
      lk.unlock();
      std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  }
}

int srv_submit(bxg_server* s, std::unique_ptr<SrvReq> r, uint64_t* id) {
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop) return BXG_ERR_INVALID_ARG;
  r->id = s->next_id++;
  r->t_submit = now_ns();
  bxg_map* m = r->map;
  if (!s->queues.count(m)) s->order.push_back(m);
  s->queues[m].push_back(r.get());
  ++s->queued;
  ++s->stats.submitted;
  if (id) *id = r->id;
  s->reqs[r->id] = std::move(r);
  s->cv_work.notify_one();
  return BXG_OK;
}

void srv_info(const SrvReq& r, bxg_request_info* out) {
  if (!out) return;
  out->status = r.status;
  out->kind = r.kind;
  out->count = r.count;
  out->submit_ns = r.t_submit;
  out->admit_ns = r.t_admit;
  out->done_ns = r.t_done;
}

}  // namespace

extern "C" {

int bxg_server_create(bxg_ctx* c, const bxg_server_options* o, bxg_server** out) {
  if (!c || !out) return BXG_ERR_INVALID_ARG;
  *out = nullptr;
  std::lock_guard<std::recursive_mutex> lk(c->mu);
  auto s = std::make_unique<bxg_server>();
  s->c = c;
  if (o) s->opt = *o;
  if (!s->opt.max_inflight) s->opt.max_inflight = 32;
  if (!s->opt.max_batch_scans) s->opt.max_batch_scans = 16;
  s->opt.max_batch_scans = std::min<uint32_t>(s->opt.max_batch_scans, 1u << kMaxTagBits);
  if (!s->opt.streams) s->opt.streams = 4;
  if (s->opt.max_inflight > 4096 || s->opt.streams > 16)
    return fail(c, BXG_ERR_INVALID_ARG, "server options out of range");
  for (uint32_t i = 0; i < s->opt.streams; ++i) {
    hipStream_t st = nullptr;
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
      for (hipStream_t x : s->streams) hipStreamDestroy(x);
      return fail(c, BXG_ERR_HIP, "hipStreamCreate");
    }
    s->streams.push_back(st);
  }
  void* p = nullptr;
  int rc = dmalloc(c, &p, (uint64_t)s->opt.max_inflight * 8);
  s->d_counts = (uint64_t*)p;
  if (rc == BXG_OK && hipHostMalloc((void**)&s->h_counts, (uint64_t)s->opt.max_inflight * 8,
                                    hipHostMallocDefault) != hipSuccess)
    rc = fail(c, BXG_ERR_HIP, "hipHostMalloc");
  if (rc != BXG_OK) {
    for (hipStream_t x : s->streams) hipStreamDestroy(x);
    dfree(s->d_counts);
    return rc;
  }
  for (uint32_t i = 0; i < s->opt.max_inflight; ++i) s->free_slots.push_back((int)i);
  bxg_server* raw = s.release();
  raw->th = std::thread(srv_loop, raw);
  *out = raw;
  return BXG_OK;
}

int bxg_server_drain(bxg_server* s) {
  if (!s) return BXG_ERR_INVALID_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->queued == 0 && s->ninflight == 0; });
  return BXG_OK;
}

int bxg_server_destroy(bxg_server* s) {
  if (!s) return BXG_ERR_INVALID_ARG;
  bxg_server_drain(s);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->cv_work.notify_all();
  }
  s->th.join();
  for (hipStream_t st : s->streams) {
    hipStreamSynchronize(st);
    hipStreamDestroy(st);
  }
  for (hipEvent_t e : s->events) hipEventDestroy(e);
  dfree(s->d_counts);
  if (s->h_counts) hipHostFree(s->h_counts);
  delete s;
  return BXG_OK;
}

int bxg_server_insert(bxg_server* s, bxg_map* m, const void* points_dev, uint64_t n, uint32_t flags,
                      const double origin[3], double max_range, uint64_t* id) {
  if (!s || !m || !origin || (n && !points_dev) || m->ctx != s->c) return BXG_ERR_INVALID_ARG;
  if (!is_prob(m)) return fail(s->c, BXG_ERR_INVALID_ARG, "insert needs a probabilistic map");
  if (n >= (1ull << 31)) return fail(s->c, BXG_ERR_INVALID_ARG, "too many points in one request");
  auto r = std::make_unique<SrvReq>();
  r->kind = kReqInsert;
  r->map = m;
  r->scan.map = m;
  r->scan.points_dev = points_dev;
  r->scan.n = n;
  r->scan.flags = flags;
  std::memcpy(r->scan.origin, origin, sizeof(r->scan.origin));
  r->scan.max_range = max_range;
  return srv_submit(s, std::move(r), id);
}

int bxg_server_lookup(bxg_server* s, bxg_map* m, const int32_t* coords_dev, uint64_t n, void* vals_dev,
                      uint8_t* found_dev, uint64_t* id) {
  if (!s || !m || (n && !coords_dev) || m->ctx != s->c || n >= (1ull << 31)) return BXG_ERR_INVALID_ARG;
  auto r = std::make_unique<SrvReq>();
  r->kind = kReqLookup;
  r->map = m;
  r->coords = coords_dev;
  r->n = n;
  r->vals = vals_dev;
  r->found = found_dev;
  return srv_submit(s, std::move(r), id);
}

int bxg_server_export(bxg_server* s, bxg_map* m, int filter, int32_t* coords_dev, void* vals_dev,
                      uint64_t cap, uint64_t* id) {
  if (!s || !m || m->ctx != s->c || filter < 0 || filter > 2) return BXG_ERR_INVALID_ARG;
  if (filter && !is_prob(m)) return fail(s->c, BXG_ERR_INVALID_ARG, "filtered export needs a probabilistic map");
  auto r = std::make_unique<SrvReq>();
  r->kind = kReqExport;
  r->map = m;
  r->filter = filter;
  r->out_coords = coords_dev;
  r->vals = vals_dev;
  r->cap = cap;
  return srv_submit(s, std::move(r), id);
}

int bxg_server_query(bxg_server* s, uint64_t id, bxg_request_info* info) {
  if (!s) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return BXG_ERR_INVALID_ARG;
  srv_info(*it->second, info);
  return BXG_OK;
}

int bxg_server_wait(bxg_server* s, uint64_t id, bxg_request_info* info) {
  if (!s) return BXG_ERR_INVALID_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->reqs.find(id);
  if (it == s->reqs.end()) return BXG_ERR_INVALID_ARG;
  SrvReq* r = it->second.get();
  s->cv_done.wait(lk, [&] { return r->status != -1; });
  srv_info(*r, info);
  const int st = r->status;
  s->reqs.erase(it);
  return st;
}

int bxg_server_get_stats(bxg_server* s, bxg_server_stats* out) {
  if (!s || !out) return BXG_ERR_INVALID_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  out->queued = s->queued;
  out->inflight = s->ninflight;
  return BXG_OK;
}

}  // extern "C"
