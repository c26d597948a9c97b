// Serving loop: continuous batching of matmul requests onto the grouped
// persistent GEMM kernel.
//
// One worker thread per server runs
//     reap finished launches -> admit queued requests -> launch
// in a loop.  Admission is in arrival order and bounded by the HBM arena
// (each request's footprint is its C, plus A and B when they come from host
// memory) and by max_resident.  A lane is a stream plus a grouped plan; in
// GROUPED mode an idle lane takes up to max_batch admitted requests into one
// launch, so with several lanes new batches start while earlier ones run.
// Completion is detected by polling each launch's event, and a request's
// latency runs from its arrival to the moment the worker sees its launch
// finish.
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "common.hpp"

namespace fg {

using Clock = std::chrono::steady_clock;

// First-fit allocator over one device allocation (256-byte granules).
class Arena {
 public:
  int init(size_t bytes) {
    cap_ = (bytes + 255) & ~(size_t)255;
    FG_CHECK(hipMalloc(reinterpret_cast<void**>(&base_), cap_));
    free_[0] = cap_;
    return FERRAY_OK;
  }
  void destroy() {
    if (base_) (void)hipFree(base_);
    base_ = nullptr;
  }
  bool alloc(size_t bytes, size_t* off) {
    bytes = round(bytes);
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second < bytes) continue;
      *off = it->first;
      const size_t rest = it->second - bytes, start = it->first + bytes;
      free_.erase(it);
      if (rest) free_[start] = rest;
      used_ += bytes;
      peak_ = std::max(peak_, used_);
      return true;
    }
    return false;
  }
  void release(size_t off, size_t bytes) {
    bytes = round(bytes);
    used_ -= bytes;
    auto it = free_.emplace(off, bytes).first;
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
      }
    }
  }
  static size_t round(size_t b) { return (std::max<size_t>(b, 1) + 255) & ~(size_t)255; }
  uint8_t* base() const { return base_; }
  size_t capacity() const { return cap_; }
  size_t peak() const { return peak_; }

 private:
  uint8_t* base_ = nullptr;
  size_t cap_ = 0, used_ = 0, peak_ = 0;
  std::map<size_t, size_t> free_;  // offset -> size
};

enum { ST_QUEUED = 0, ST_ADMITTED = 1, ST_RUNNING = 2, ST_DONE = 3 };

struct Req {
  ferray_request r;
  int state = ST_QUEUED;
  int status = FERRAY_OK;
  bool holds_c = false, holds_ab = false;
  int readers = 0;  // ferray_server_result copying C out with the lock dropped
  size_t c_off = 0, c_bytes = 0, a_off = 0, a_bytes = 0, b_off = 0, b_bytes = 0;
  Clock::time_point t_arrive, t_done;
};

struct Flight {  // PER_REQUEST mode: one launch per request
  uint64_t ticket;
  hipEvent_t ev;
};

struct Lane {
  hipStream_t stream = nullptr;
  ferray_grouped_plan* plan = nullptr;
  hipEvent_t ev = nullptr;
  std::vector<uint64_t> batch;  // GROUPED: the launch in flight
  bool busy = false;
  std::deque<Flight> flights;   // PER_REQUEST: launches in stream order
};

static size_t elem_bytes(int dtype) {
  switch (dtype) {
    case FERRAY_DTYPE_F64: return 8;
    case FERRAY_DTYPE_F32: return 4;
    case FERRAY_DTYPE_C64: return 16;
    case FERRAY_DTYPE_C32: return 8;
    default: return 1;
  }
}
static size_t c_elem_bytes(int dtype) { return dtype == FERRAY_DTYPE_I8 ? 4 : elem_bytes(dtype); }

}  // namespace fg

struct ferray_server {
  ferray_server_config cfg{};
  fg::Arena arena;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::unordered_map<uint64_t, fg::Req> reqs;
  std::deque<uint64_t> pending, ready;
  uint64_t next_ticket = 1;
  size_t resident = 0;
  std::vector<fg::Lane> lanes;
  size_t next_lane = 0;
  std::vector<hipEvent_t> event_pool;
  std::thread worker;
  bool stop = false;
  ferray_server_stats stats{};
};

namespace fg {

static void free_req_memory(ferray_server* s, Req& q, bool keep_c) {
  if (q.holds_ab) {
    s->arena.release(q.a_off, q.a_bytes);
    s->arena.release(q.b_off, q.b_bytes);
    q.holds_ab = false;
  }
  if (q.holds_c && !keep_c) {
    s->arena.release(q.c_off, q.c_bytes);
    q.holds_c = false;
  }
}

// Marks a request done (lock held).
static void complete(ferray_server* s, uint64_t t, int status, Clock::time_point now) {
  Req& q = s->reqs[t];
  q.state = ST_DONE;
  q.status = status;
  q.t_done = now;
  const bool keep = s->cfg.keep_results && status == FERRAY_OK && !s->stop;
  free_req_memory(s, q, keep);
  if (!q.holds_c) --s->resident;
  ++s->stats.completed;
}

static void reap(ferray_server* s) {
  const auto now = Clock::now();
  bool any = false;
  for (Lane& l : s->lanes) {
    if (!l.busy) continue;
    const hipError_t e = hipEventQuery(l.ev);
    if (e == hipErrorNotReady) continue;
    const int st = e == hipSuccess ? FERRAY_OK : map_hip_error(e);
    for (uint64_t t : l.batch) complete(s, t, st, now);
    l.batch.clear();
    l.busy = false;
    any = true;
  }
  // A stream completes its launches in order: poll each lane's oldest.
  for (Lane& l : s->lanes) {
    while (!l.flights.empty()) {
      const Flight f = l.flights.front();
      const hipError_t e = hipEventQuery(f.ev);
      if (e == hipErrorNotReady) break;
      complete(s, f.ticket, e == hipSuccess ? FERRAY_OK : map_hip_error(e), now);
      s->event_pool.push_back(f.ev);
      l.flights.pop_front();
      any = true;
    }
  }
  if (any) s->cv_done.notify_all();
}

static void admit(ferray_server* s) {
  while (!s->pending.empty() && s->resident < s->cfg.max_resident) {
    const uint64_t t = s->pending.front();
    Req& q = s->reqs[t];
    size_t c_off, a_off = 0, b_off = 0;
    if (!s->arena.alloc(q.c_bytes, &c_off)) return;
    if (q.r.operands_on_host) {
      if (!s->arena.alloc(q.a_bytes, &a_off)) {
        s->arena.release(c_off, q.c_bytes);
        return;
      }
      if (!s->arena.alloc(q.b_bytes, &b_off)) {
        s->arena.release(c_off, q.c_bytes);
        s->arena.release(a_off, q.a_bytes);
        return;
      }
      q.holds_ab = true;
      q.a_off = a_off;
      q.b_off = b_off;
    }
    q.holds_c = true;
    q.c_off = c_off;
    q.state = ST_ADMITTED;
    ++s->resident;
    s->pending.pop_front();
    s->ready.push_back(t);
  }
}

// Device operands of an admitted request (host operands are copied into the
// arena on `st` first).
static int stage_operands(ferray_server* s, const Req& q, hipStream_t st, const void** a,
                          const void** b) {
  if (!q.r.operands_on_host) {
    *a = q.r.a;
    *b = q.r.b;
    return FERRAY_OK;
  }
  uint8_t* base = s->arena.base();
  if (q.a_bytes) FG_CHECK(hipMemcpyAsync(base + q.a_off, q.r.a, q.a_bytes, hipMemcpyHostToDevice, st));
  if (q.b_bytes) FG_CHECK(hipMemcpyAsync(base + q.b_off, q.r.b, q.b_bytes, hipMemcpyHostToDevice, st));
  *a = base + q.a_off;
  *b = base + q.b_off;
  return FERRAY_OK;
}

static int launch_single(const ferray_request& r, const void* a, const void* b, void* c,
                         hipStream_t st) {
  switch (r.dtype) {
    case FERRAY_DTYPE_F64:
      return ferray_gemm_f64_dev(r.m, r.n, r.k, r.alpha_re, (const double*)a, (const double*)b, 0.0,
                                 (double*)c, st);
    case FERRAY_DTYPE_F32:
      return ferray_gemm_f32_dev(r.m, r.n, r.k, (float)r.alpha_re, (const float*)a,
                                 (const float*)b, 0.0f, (float*)c, st);
    case FERRAY_DTYPE_C64:
      return ferray_gemm_c64_dev(r.m, r.n, r.k, r.alpha_re, r.alpha_im, (const double*)a,
                                 (const double*)b, 0.0, 0.0, (double*)c, st);
    case FERRAY_DTYPE_C32:
      return ferray_gemm_c32_dev(r.m, r.n, r.k, (float)r.alpha_re, (float)r.alpha_im,
                                 (const float*)a, (const float*)b, 0.0f, 0.0f, (float*)c, st);
    default:
      return ferray_gemm_i8_signed_dev(r.m, r.n, r.k, (const int8_t*)a, (const int8_t*)b,
                                       (int32_t*)c, false, st);
  }
}

// Launches what can be launched; called and returns with the lock held, but
// drops it around HIP calls.  Returns whether anything was launched.
static bool launch(ferray_server* s, std::unique_lock<std::mutex>& lk) {
  bool launched = false;
  if (s->cfg.mode == FERRAY_SERVE_PER_REQUEST) {
    while (!s->ready.empty()) {
      const uint64_t t = s->ready.front();
      s->ready.pop_front();
      Req& q = s->reqs[t];
      q.state = ST_RUNNING;
      const Req snap = q;
      Lane& l = s->lanes[s->next_lane++ % s->lanes.size()];
      hipEvent_t ev = nullptr;
      if (!s->event_pool.empty()) {
        ev = s->event_pool.back();
        s->event_pool.pop_back();
      }
      lk.unlock();
      int st = FERRAY_OK;
      if (!ev && hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) {
        ev = nullptr;
        st = FERRAY_ERR_HIP;
      }
      const void *a = nullptr, *b = nullptr;
      if (!st) st = stage_operands(s, snap, l.stream, &a, &b);
      if (!st) st = launch_single(snap.r, a, b, s->arena.base() + snap.c_off, l.stream);
      if (!st && hipEventRecord(ev, l.stream) != hipSuccess) st = FERRAY_ERR_HIP;
      lk.lock();
      ++s->stats.launches;
      ++s->stats.launched_requests;
      if (st) {
        if (ev) s->event_pool.push_back(ev);
        complete(s, t, st, Clock::now());
        s->cv_done.notify_all();
      } else {
        l.flights.push_back(Flight{t, ev});
      }
      launched = true;
    }
    return launched;
  }
  while (!s->ready.empty()) {
    Lane* lane = nullptr;
    for (size_t i = 0; i < s->lanes.size(); ++i) {
      Lane& l = s->lanes[(s->next_lane + i) % s->lanes.size()];
      if (!l.busy) {
        lane = &l;
        s->next_lane = (s->next_lane + i + 1) % s->lanes.size();
        break;
      }
    }
    if (!lane) break;
    std::vector<uint64_t> batch;
    std::vector<Req> snaps;
    while (!s->ready.empty() && batch.size() < s->cfg.max_batch) {
      const uint64_t t = s->ready.front();
      s->ready.pop_front();
      Req& q = s->reqs[t];
      q.state = ST_RUNNING;
      batch.push_back(t);
      snaps.push_back(q);
    }
    lane->busy = true;
    lk.unlock();
    int st = FERRAY_OK;
    std::vector<ferray_gemm_problem> probs(snaps.size());
    for (size_t i = 0; i < snaps.size() && !st; ++i) {
      const Req& q = snaps[i];
      const void *a = nullptr, *b = nullptr;
      st = stage_operands(s, q, lane->stream, &a, &b);
      ferray_gemm_problem& p = probs[i];
      p.dtype = q.r.dtype;
      p.accumulate = 0;
      p.m = q.r.m;
      p.n = q.r.n;
      p.k = q.r.k;
      p.a = a;
      p.b = b;
      p.c = s->arena.base() + q.c_off;
      p.alpha_re = q.r.alpha_re;
      p.alpha_im = q.r.alpha_im;
      p.beta_re = 0.0;
      p.beta_im = 0.0;
    }
    if (!st) st = ferray_grouped_plan_launch(lane->plan, probs.data(), probs.size(), lane->stream);
    if (!st && hipEventRecord(lane->ev, lane->stream) != hipSuccess) st = FERRAY_ERR_HIP;
    lk.lock();
    ++s->stats.launches;
    s->stats.launched_requests += batch.size();
    if (st) {
      lane->busy = false;
      const auto now = Clock::now();
      for (uint64_t t : batch) complete(s, t, st, now);
      s->cv_done.notify_all();
    } else {
      lane->batch = std::move(batch);
    }
    launched = true;
  }
  return launched;
}

static bool in_flight(const ferray_server* s) {
  for (const Lane& l : s->lanes)
    if (l.busy || !l.flights.empty()) return true;
  return false;
}

static void worker_main(ferray_server* s) {
  (void)ensure_device();
  std::unique_lock<std::mutex> lk(s->mu);
  for (;;) {
    reap(s);
    admit(s);
    const bool launched = launch(s, lk);
    const bool busy = in_flight(s);
    if (s->stop && s->pending.empty() && s->ready.empty() && !busy) break;
    if (launched) continue;
    if (busy) {
      // Poll completions; submissions and releases take the lock meanwhile.
      lk.unlock();
      std::this_thread::yield();
      lk.lock();
    } else {
      s->cv_work.wait(lk);
    }
  }
}

static int check_request(const ferray_server* s, const ferray_request& r, size_t* a_bytes,
                         size_t* b_bytes, size_t* c_bytes) {
  if (r.dtype < FERRAY_DTYPE_F64 || r.dtype > FERRAY_DTYPE_I8) return FERRAY_ERR_INVALID;
  if (r.m == 0 || r.n == 0) return FERRAY_ERR_INVALID;
  if (r.m > (1u << 29) || r.n > (1u << 28) || r.k > (1u << 28)) return FERRAY_ERR_INVALID;
  if (r.dtype == FERRAY_DTYPE_I8 && r.k > (1u << 16)) return FERRAY_ERR_UNSUPPORTED;
  if (r.k && (!r.a || !r.b)) return FERRAY_ERR_INVALID;
  const size_t e = elem_bytes(r.dtype);
  *a_bytes = r.m * r.k * e;
  *b_bytes = r.k * r.n * e;
  *c_bytes = r.m * r.n * c_elem_bytes(r.dtype);
  size_t need = Arena::round(*c_bytes);
  if (r.operands_on_host) need += Arena::round(*a_bytes) + Arena::round(*b_bytes);
  if (need > s->arena.capacity()) return FERRAY_ERR_OOM;
  return FERRAY_OK;
}

static int submit_at(ferray_server* s, const ferray_request& r, Clock::time_point arrive,
                     uint64_t* ticket) {
  Req q;
  q.r = r;
  int st = check_request(s, r, &q.a_bytes, &q.b_bytes, &q.c_bytes);
  if (st) return st;
  q.t_arrive = arrive;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return FERRAY_ERR_INVALID;
    const uint64_t t = s->next_ticket++;
    s->reqs.emplace(t, q);
    s->pending.push_back(t);
    ++s->stats.submitted;
    *ticket = t;
  }
  s->cv_work.notify_one();
  return FERRAY_OK;
}

// Waits, with the lock held by `lk`, until `t` is done and no result copy of
// it is in flight; returns the request, or nullptr if the ticket is unknown
// or was released meanwhile.  The ticket is looked up again on every wake-up:
// a submission may rehash the map while the lock is dropped, and another
// thread may release the ticket.
static Req* wait_done(ferray_server* s, uint64_t t, std::unique_lock<std::mutex>& lk) {
  Req* q = nullptr;
  s->cv_done.wait(lk, [&] {
    auto it = s->reqs.find(t);
    q = it == s->reqs.end() ? nullptr : &it->second;
    return !q || (q->state == ST_DONE && q->readers == 0);
  });
  return q;
}

static void release_locked(ferray_server* s, uint64_t t) {
  auto it = s->reqs.find(t);
  if (it == s->reqs.end()) return;
  Req& q = it->second;
  if (q.holds_c) {
    s->arena.release(q.c_off, q.c_bytes);
    q.holds_c = false;
    --s->resident;
  }
  s->reqs.erase(it);
  s->cv_work.notify_one();
}

}  // namespace fg

using namespace fg;

extern "C" {

int ferray_server_create(const ferray_server_config* cfg, ferray_server** out) {
  int st = ensure_device();
  if (st) return st;
  if (!cfg || !out || !cfg->hbm_bytes || !cfg->max_resident || !cfg->max_batch ||
      cfg->lanes < 1 || cfg->lanes > 16 ||
      (cfg->mode != FERRAY_SERVE_GROUPED && cfg->mode != FERRAY_SERVE_PER_REQUEST))
    return FERRAY_ERR_INVALID;
  *out = nullptr;
  auto* s = new (std::nothrow) ferray_server();
  if (!s) return FERRAY_ERR_OOM;
  s->cfg = *cfg;
  st = s->arena.init(cfg->hbm_bytes);
  s->lanes.resize(cfg->lanes);
  for (Lane& l : s->lanes) {
    if (st) break;
    if (hipStreamCreateWithFlags(&l.stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&l.ev, hipEventDisableTiming) != hipSuccess) {
      st = FERRAY_ERR_HIP;
      break;
    }
    if (cfg->mode == FERRAY_SERVE_GROUPED)
      st = ferray_grouped_plan_create(cfg->max_batch, (size_t)1 << 30, &l.plan);
  }
  if (st) {
    for (Lane& l : s->lanes) {
      if (l.plan) ferray_grouped_plan_destroy(l.plan);
      if (l.ev) (void)hipEventDestroy(l.ev);
      if (l.stream) (void)hipStreamDestroy(l.stream);
    }
    s->arena.destroy();
    delete s;
    return st;
  }
  s->worker = std::thread(worker_main, s);
  *out = s;
  return FERRAY_OK;
}

int ferray_server_destroy(ferray_server* s) {
  if (!s) return FERRAY_OK;
  {
    // Kept results die with the server; dropping them first lets queued
    // requests be admitted while the worker drains.
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    for (auto& kv : s->reqs) {
      Req& q = kv.second;
      if (q.state == ST_DONE && q.holds_c) {
        s->arena.release(q.c_off, q.c_bytes);
        q.holds_c = false;
        --s->resident;
      }
    }
  }
  s->cv_work.notify_all();
  if (s->worker.joinable()) s->worker.join();
  for (Lane& l : s->lanes) {
    if (l.stream) (void)hipStreamSynchronize(l.stream);
    if (l.plan) ferray_grouped_plan_destroy(l.plan);
    if (l.ev) (void)hipEventDestroy(l.ev);
    if (l.stream) {
      release_workspace(l.stream);
      (void)hipStreamDestroy(l.stream);
    }
  }
  for (hipEvent_t e : s->event_pool) (void)hipEventDestroy(e);
  s->arena.destroy();
  delete s;
  return FERRAY_OK;
}

int ferray_server_submit(ferray_server* s, const ferray_request* r, uint64_t* ticket) {
  if (int st0 = ensure_device()) return st0;
  if (!s || !r || !ticket) return FERRAY_ERR_INVALID;
  return submit_at(s, *r, Clock::now(), ticket);
}

int ferray_server_poll(ferray_server* s, uint64_t t, int* state) {
  if (!s || !state) return FERRAY_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(t);
  if (it == s->reqs.end()) return FERRAY_ERR_INVALID;
  *state = it->second.state;
  return FERRAY_OK;
}

int ferray_server_wait(ferray_server* s, uint64_t t, double* latency_ms) {
  if (!s) return FERRAY_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  Req* q = wait_done(s, t, lk);
  if (!q) return FERRAY_ERR_INVALID;
  if (latency_ms)
    *latency_ms = std::chrono::duration<double, std::milli>(q->t_done - q->t_arrive).count();
  return q->status;
}

int ferray_server_result_dev(ferray_server* s, uint64_t t, const void** c_dev) {
  if (!s || !c_dev) return FERRAY_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  Req* q = wait_done(s, t, lk);
  if (!q) return FERRAY_ERR_INVALID;
  if (q->status) return q->status;
  if (!q->holds_c) return FERRAY_ERR_INVALID;
  *c_dev = s->arena.base() + q->c_off;
  return FERRAY_OK;
}

int ferray_server_result(ferray_server* s, uint64_t t, void* c_host, size_t bytes) {
  if (int st0 = ensure_device()) return st0;
  if (!s || !c_host) return FERRAY_ERR_INVALID;
  const void* src = nullptr;
  {
    std::unique_lock<std::mutex> lk(s->mu);
    Req* q = wait_done(s, t, lk);
    if (!q) return FERRAY_ERR_INVALID;
    if (q->status) return q->status;
    if (!q->holds_c || bytes != q->c_bytes) return FERRAY_ERR_INVALID;
    src = s->arena.base() + q->c_off;
    q->readers = 1;  // a concurrent release or result waits for the copy
  }
  const hipError_t e = hipMemcpy(c_host, src, bytes, hipMemcpyDeviceToHost);
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->reqs.find(t);
  if (it != s->reqs.end()) {
    it->second.readers = 0;
    if (e == hipSuccess) release_locked(s, t);
  }
  s->cv_done.notify_all();
  return e == hipSuccess ? FERRAY_OK : map_hip_error(e);
}

int ferray_server_release(ferray_server* s, uint64_t t) {
  if (!s) return FERRAY_ERR_INVALID;
  std::unique_lock<std::mutex> lk(s->mu);
  Req* q = wait_done(s, t, lk);
  if (!q) return FERRAY_ERR_INVALID;
  release_locked(s, t);
  return FERRAY_OK;
}

int ferray_server_stats_get(ferray_server* s, ferray_server_stats* out) {
  if (!s || !out) return FERRAY_ERR_INVALID;
  std::lock_guard<std::mutex> lk(s->mu);
  *out = s->stats;
  out->arena_bytes = s->arena.capacity();
  out->arena_peak = s->arena.peak();
  return FERRAY_OK;
}

int ferray_serve_trace(const ferray_server_config* cfg, const ferray_request* reqs,
                       const double* arrival_s, size_t count, double* latency_ms, double* wall_s,
                       ferray_server_stats* stats) {
  if (!cfg || (count && (!reqs || !arrival_s || !latency_ms))) return FERRAY_ERR_INVALID;
  for (size_t i = 1; i < count; ++i)
    if (arrival_s[i] < arrival_s[i - 1]) return FERRAY_ERR_INVALID;
  ferray_server_config c = *cfg;
  c.keep_results = 0;
  ferray_server* s = nullptr;
  int st = ferray_server_create(&c, &s);
  if (st) return st;
  std::vector<uint64_t> tickets(count, 0);
  std::vector<int> sub(count, FERRAY_OK);
  const auto t0 = Clock::now() + std::chrono::milliseconds(2);
  for (size_t i = 0; i < count; ++i) {
    const auto target =
        t0 + std::chrono::duration_cast<Clock::duration>(std::chrono::duration<double>(arrival_s[i]));
    for (;;) {
      const auto now = Clock::now();
      if (now >= target) break;
      if (target - now > std::chrono::microseconds(300))
        std::this_thread::sleep_for(std::chrono::microseconds(100));
    }
    sub[i] = submit_at(s, reqs[i], target, &tickets[i]);
  }
  Clock::time_point last = t0;
  int first_err = FERRAY_OK;
  for (size_t i = 0; i < count; ++i) {
    if (sub[i]) {
      latency_ms[i] = -1.0;
      if (!first_err) first_err = sub[i];
      continue;
    }
    std::unique_lock<std::mutex> lk(s->mu);
    Req* q = wait_done(s, tickets[i], lk);
    latency_ms[i] = std::chrono::duration<double, std::milli>(q->t_done - q->t_arrive).count();
    last = std::max(last, q->t_done);
    if (q->status && !first_err) first_err = q->status;
    s->reqs.erase(tickets[i]);
  }
  if (wall_s) *wall_s = std::chrono::duration<double>(last - t0).count();
  if (stats) ferray_server_stats_get(s, stats);
  ferray_server_destroy(s);
  return first_err;
}

}  // extern "C"
