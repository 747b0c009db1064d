// Serving: a request loop with continuous batching.
//
// Requests (one pattern each) are prepared on the submitting thread, queued FIFO, and admitted
// onto one of `slots` HIP streams as soon as a slot is free and their device footprint fits the
// HBM budget. A slot takes the front request together with the requests queued behind it
// (same options, FIFO order, up to max_group of them, within the budget) as one group: one
// device set for all of them, i.e. one launch sequence, as a batch. Device state is allocated
// and freed stream-ordered on the slot (hipMallocAsync / hipFreeAsync: no device-wide
// synchronisation). Groups on different slots run concurrently: their workgroups fill the CUs
// others leave free, or, with cu_share set, each group's persistent grids are sized for that
// many CUs. A completion thread polls the slots' events, reads the results of a finished group
// on its stream and releases the slot. Latency is submit-to-result on the host clock.
// No counterpart in the reference (it runs one matrix per process invocation).
#include <hip/hip_runtime_api.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <thread>

#include "engine.h"

using namespace gsofa;

namespace {

using Clock = std::chrono::steady_clock;

enum class State { Pending, Running, Done, Failed, Cancelled, Rejected };

struct Request {
  int64_t id = 0;
  int32_t n = 0;
  Prepared prep;  // released once the request can no longer run (finished, cancelled, rejected)
  Options opt;
  int64_t footprint = 0;
  State state = State::Pending;
  bool cancel_requested = false;
  int error = GSOFA_OK;
  std::string error_text;
  Clock::time_point t_submit, t_admit, t_done;
  MatResult result;
  double device_ms = 0.0;  // of the request's group
  uint64_t nzu_digest = 0, sup_digest = 0;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t start = nullptr, done = nullptr;
  std::vector<Request*> group;  // empty: the slot is free
  std::unique_ptr<DeviceSet> set;
  int64_t footprint = 0;        // charged against the budget while the group runs
  int error = GSOFA_OK;
  std::string error_text;
  std::atomic<bool> launched{false};  // set after the slot's done event was recorded
};

bool same_options(const Options& a, const Options& b) {
  return a.chunk == b.chunk && a.n_processes == b.n_processes && a.structure == b.structure &&
         a.device == b.device && a.heavy_threads == b.heavy_threads && a.heavy_wgs_per_cu == b.heavy_wgs_per_cu &&
         a.light_wgs_per_cu == b.light_wgs_per_cu && a.qcap == b.qcap && a.chain_block == b.chain_block &&
         a.pair_directions == b.pair_directions && a.bitmap_in_hbm == b.bitmap_in_hbm && a.cu_share == b.cu_share &&
         a.split_blocks == b.split_blocks && a.symmetric == b.symmetric &&
         a.early_top == b.early_top && a.pre_escalate == b.pre_escalate;
}

// Host bytes a request keeps alive (its prepared pattern).
int64_t prep_bytes(const Prepared& p) {
  return (int64_t)(p.rowptr.capacity() + p.adj.capacity() + p.dup_u.capacity() + p.dup_all.capacity()) * 4;
}

void release_prep(Request* r) { r->prep = Prepared(); }

uint64_t fnv1a(const void* data, size_t bytes) {
  const unsigned char* p = static_cast<const unsigned char*>(data);
  uint64_t h = 1469598103934665603ull;
  for (size_t i = 0; i < bytes; ++i) {
    h ^= p[i];
    h *= 1099511628211ull;
  }
  return h;
}

}  // namespace

struct gsofa_server {
  int device = 0;
  int cu_share = 0;
  int max_group = 16;
  int64_t budget = 0;
  std::atomic<int64_t> in_use{0};
  std::vector<std::unique_ptr<Slot>> slots;
  std::mutex mu;
  std::condition_variable cv_admit, cv_done;
  std::deque<Request*> pending;
  std::map<int64_t, std::unique_ptr<Request>> requests;
  int64_t next_id = 1;
  bool stopping = false;
  int running = 0, peak_running = 0;
  int64_t completed = 0;
  std::thread admit_thread, done_thread;

  void admit_loop();
  void done_loop();
  void finish(Slot& s);
  Slot* free_slot() {
    for (auto& s : slots)
      if (s->group.empty()) return s.get();
    return nullptr;
  }
};

void gsofa_server::admit_loop() {
  (void)hipSetDevice(device);
  std::unique_lock<std::mutex> lk(mu);
  while (true) {
    cv_admit.wait(lk, [&] {
      if (stopping) return true;
      if (pending.empty()) return false;
      Request* r = pending.front();
      if (r->cancel_requested) return true;
      return free_slot() != nullptr && in_use.load() + r->footprint <= budget;
    });
    if (stopping) return;
    Request* r = pending.front();
    if (r->cancel_requested) {
      pending.pop_front();
      r->state = State::Cancelled;
      release_prep(r);
      cv_done.notify_all();
      continue;
    }
    // the group: the front request and the ones queued behind it, in order, while options
    // match and the group's own footprint fits. That footprint is computed for the group as a
    // whole: one paired member sizes every workgroup region of the set for paired items, so it
    // can exceed the sum of the members' own footprints.
    std::vector<Request*> grp{r};
    std::vector<const Prepared*> gmats{&r->prep};
    int64_t fp = r->footprint;
    pending.pop_front();
    while (!pending.empty() && (int)grp.size() < max_group) {
      Request* q = pending.front();
      if (q->cancel_requested) {
        pending.pop_front();
        q->state = State::Cancelled;
        release_prep(q);
        cv_done.notify_all();
        continue;
      }
      if (!same_options(q->opt, r->opt)) break;
      gmats.push_back(&q->prep);
      int64_t fq = 0;
      if (DeviceSet::footprint(gmats, r->opt, &fq) != GSOFA_OK || in_use.load() + fq > budget) {
        gmats.pop_back();
        break;
      }
      grp.push_back(q);
      fp = fq;
      pending.pop_front();
    }
    Slot& s = *free_slot();
    s.group = grp;
    s.footprint = fp;
    s.error = GSOFA_OK;
    s.error_text.clear();
    s.launched.store(false, std::memory_order_relaxed);
    const Clock::time_point now = Clock::now();
    for (Request* q : grp) {
      q->state = State::Running;
      q->t_admit = now;
    }
    in_use += fp;
    running += (int)grp.size();
    peak_running = std::max(peak_running, running);
    lk.unlock();
    // allocation, upload and launch outside the lock; all on the slot's stream
    std::vector<const Prepared*> mats;
    mats.reserve(grp.size());
    for (Request* q : grp) mats.push_back(&q->prep);
    int rc = GSOFA_OK;
    s.set.reset(new (std::nothrow) DeviceSet);
    if (!s.set) rc = fail(GSOFA_ERR_OUT_OF_MEMORY, "device set");
    if (rc == GSOFA_OK) rc = s.set->create(mats, r->opt, s.stream);
    if (rc == GSOFA_OK && hipEventRecord(s.start, s.stream) != hipSuccess) rc = fail(GSOFA_ERR_HIP, "hipEventRecord");
    if (rc == GSOFA_OK) rc = s.set->enqueue(s.stream);
    if (rc != GSOFA_OK) {
      s.error = rc;
      s.error_text = gsofa_last_error();
    }
    if (hipEventRecord(s.done, s.stream) != hipSuccess && rc == GSOFA_OK) {
      s.error = GSOFA_ERR_HIP;
      s.error_text = "hipEventRecord";
    }
    s.launched.store(true, std::memory_order_release);
    lk.lock();
  }
}

void gsofa_server::finish(Slot& s) {
  std::vector<MatResult> res;
  float ms = 0.f;
  if (s.error == GSOFA_OK) {
    (void)hipEventElapsedTime(&ms, s.start, s.done);
    int rc = s.set->read_results(res);
    if (rc != GSOFA_OK) {
      s.error = rc;
      s.error_text = gsofa_last_error();
    }
  }
  for (size_t i = 0; i < s.group.size(); ++i) {
    Request* r = s.group[i];
    r->error = s.error;
    r->error_text = s.error_text;
    if (r->error != GSOFA_OK) continue;
    r->device_ms = ms;
    r->result = res[i];
    const size_t n = (size_t)r->prep.n;
    std::vector<int32_t> nzu(n);
    std::vector<uint8_t> flags(n);
    const int rc = s.set->read_counts((int)i, nullptr, nullptr, flags.data(), nzu.data());
    if (rc == GSOFA_OK) {
      r->nzu_digest = fnv1a(nzu.data(), n * 4);
      r->sup_digest = fnv1a(flags.data(), n);
    } else {
      r->error = rc;
      r->error_text = gsofa_last_error();
    }
  }
  s.set.reset();  // stream-ordered frees on the slot
  for (Request* r : s.group) release_prep(r);  // only n and the results are kept
  const Clock::time_point now = Clock::now();
  std::lock_guard<std::mutex> lk(mu);
  in_use -= s.footprint;
  running -= (int)s.group.size();
  completed += (int64_t)s.group.size();
  for (Request* r : s.group) {
    r->t_done = now;
    if (r->cancel_requested) r->state = State::Cancelled;
    else r->state = r->error == GSOFA_OK ? State::Done : State::Failed;
  }
  s.group.clear();
  s.footprint = 0;
  cv_admit.notify_all();
  cv_done.notify_all();
}

void gsofa_server::done_loop() {
  (void)hipSetDevice(device);
  while (true) {
    bool any = false;
    {
      std::lock_guard<std::mutex> lk(mu);
      if (stopping && running == 0) return;
    }
    for (auto& sp : slots) {
      Slot& s = *sp;
      bool busy;
      {
        std::lock_guard<std::mutex> lk(mu);
        busy = !s.group.empty();
      }
      // the admission thread records s.done after enqueueing; an unrecorded event reads as done
      if (!busy || !s.launched.load(std::memory_order_acquire)) continue;
      hipError_t q = hipEventQuery(s.done);
      if (q == hipErrorNotReady) continue;
      any = true;
      finish(s);
    }
    if (!any) std::this_thread::sleep_for(std::chrono::microseconds(50));
  }
}

extern "C" {

void gsofa_server_options_default(gsofa_server_options* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->slots = 4;
  o->hbm_budget_bytes = 0;  // 0 = 80% of free device memory at start
  o->max_group = 16;
}

int gsofa_server_start(const gsofa_server_options* opt, gsofa_server** out) {
  if (!out) return fail(GSOFA_ERR_INVALID_ARG, "out is null");
  *out = nullptr;
  gsofa_server_options def;
  gsofa_server_options_default(&def);
  if (!opt) opt = &def;
  if (opt->slots <= 0 || opt->slots > 64) return fail(GSOFA_ERR_INVALID_ARG, "slots must be in [1, 64]");
  if (opt->max_group < 0 || opt->max_group > 1024) return fail(GSOFA_ERR_INVALID_ARG, "max_group must be in [0, 1024]");
  DeviceInfo info;
  int rc = device_info(opt->device, info);
  if (rc != GSOFA_OK) return rc;
  if (hipSetDevice(opt->device) != hipSuccess) return fail(GSOFA_ERR_HIP, "hipSetDevice");
  std::unique_ptr<gsofa_server> srv(new (std::nothrow) gsofa_server);
  if (!srv) return fail(GSOFA_ERR_OUT_OF_MEMORY, "server");
  srv->device = opt->device;
  srv->cu_share = opt->cu_share < 0 ? 0 : opt->cu_share;
  srv->max_group = opt->max_group > 0 ? opt->max_group : 16;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return fail(GSOFA_ERR_HIP, "hipMemGetInfo");
  srv->budget = opt->hbm_budget_bytes > 0 ? opt->hbm_budget_bytes : (int64_t)(0.8 * (double)free_b);
  for (int k = 0; k < opt->slots; ++k) {
    srv->slots.emplace_back(new (std::nothrow) Slot);
    Slot* s = srv->slots.back().get();
    if (!s) return fail(GSOFA_ERR_OUT_OF_MEMORY, "slot");
    if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&s->start) != hipSuccess || hipEventCreate(&s->done) != hipSuccess)
      return fail(GSOFA_ERR_HIP, "slot stream/event creation");
  }
  srv->admit_thread = std::thread([p = srv.get()] { p->admit_loop(); });
  srv->done_thread = std::thread([p = srv.get()] { p->done_loop(); });
  *out = srv.release();
  return GSOFA_OK;
}

int gsofa_server_submit(gsofa_server* srv, const gsofa_csr* a, const gsofa_options* opt, int64_t* id) {
  if (!srv || !id) return fail(GSOFA_ERR_INVALID_ARG, "null argument");
  std::unique_ptr<Request> r(new (std::nothrow) Request);
  if (!r) return fail(GSOFA_ERR_OUT_OF_MEMORY, "request");
  gsofa_options def;
  gsofa_options_default(&def);
  const gsofa_options* o = opt ? opt : &def;
  // Served requests report totals, nzU and supernode digests; the explicit structure and the
  // autotuner (which times candidates on the device) are plan and batch features.
  if (o->structure) return fail(GSOFA_ERR_INVALID_ARG, "the server does not produce the explicit structure");
  if (o->autotune)
    return fail(GSOFA_ERR_INVALID_ARG,
                "the server does not autotune: tune a plan or batch and pass heavy_threads, heavy_wgs_per_cu and "
                "chain_block");
  r->opt.chunk = o->chunk_size;
  r->opt.n_processes = o->n_processes;
  r->opt.device = srv->device;
  r->opt.heavy_threads = o->heavy_threads;
  r->opt.heavy_wgs_per_cu = o->heavy_wgs_per_cu;
  r->opt.light_wgs_per_cu = o->light_wgs_per_cu;
  r->opt.qcap = o->queue_capacity;
  r->opt.chain_block = o->chain_block > 0 ? o->chain_block : 0;
  r->opt.pair_directions = o->pair_directions < 0 ? -1 : (o->pair_directions ? 1 : 0);
  r->opt.bitmap_in_hbm = o->bitmap_in_hbm ? 1 : 0;
  // The cuts and pre-escalation need the elimination tree (as much host time as the preparation
  // itself): not on the admission path, whose host work bounds the served rate.
  r->opt.split_blocks = -1;
  r->opt.pre_escalate = -1;
  r->opt.structure = 0;
  DeviceInfo info;
  int rc = device_info(srv->device, info);
  if (rc != GSOFA_OK) return rc;
  r->opt.cu_share = srv->cu_share;  // 0: grids sized for the whole device
  rc = prepare(a, r->prep);
  if (rc != GSOFA_OK) return rc;
  r->n = r->prep.n;
  rc = (int)DeviceSet::footprint({&r->prep}, r->opt, &r->footprint);
  if (rc != GSOFA_OK) return rc;
  r->t_submit = Clock::now();
  std::lock_guard<std::mutex> lk(srv->mu);
  if (srv->stopping) return fail(GSOFA_ERR_STATE, "server is stopping");
  r->id = srv->next_id++;
  *id = r->id;
  if (r->footprint > srv->budget) {
    r->state = State::Rejected;
    r->error = GSOFA_ERR_REJECTED;
    r->error_text = "request footprint " + std::to_string(r->footprint) + " B exceeds the HBM budget " +
                    std::to_string(srv->budget) + " B";
    release_prep(r.get());
    srv->requests[r->id] = std::move(r);
    return GSOFA_OK;
  }
  srv->pending.push_back(r.get());
  srv->requests[r->id] = std::move(r);
  srv->cv_admit.notify_all();
  return GSOFA_OK;
}

static int report(Request* r, int n_processes_dummy, gsofa_result* out) {
  (void)n_processes_dummy;
  switch (r->state) {
    case State::Pending:
    case State::Running: return fail(GSOFA_ERR_NOT_READY, "request not finished");
    case State::Cancelled: return fail(GSOFA_ERR_CANCELLED, "request cancelled");
    case State::Rejected: return fail(GSOFA_ERR_REJECTED, r->error_text);
    case State::Failed: return fail(r->error, r->error_text);
    case State::Done: break;
  }
  if (out) {
    const MatResult& m = r->result;
    out->fill_count = m.fill;
    out->nnz_L = m.nnz_l;
    out->nnz_U = m.nnz_u;
    out->nnz_F = m.nnz_l + m.nnz_u + (uint64_t)r->n;
    out->edge_checks = m.W;
    out->reach = m.R;
    out->supernodes_counted = m.nsup;
    out->supernodes = m.nsup + r->opt.n_processes;
    out->heavy_sources = m.heavy;
    out->time_ms = r->device_ms;
  }
  return GSOFA_OK;
}

int gsofa_server_poll(gsofa_server* srv, int64_t id, gsofa_result* out) {
  if (!srv) return fail(GSOFA_ERR_INVALID_ARG, "server is null");
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->requests.find(id);
  if (it == srv->requests.end()) return fail(GSOFA_ERR_INVALID_ARG, "unknown request id");
  return report(it->second.get(), 0, out);
}

int gsofa_server_wait(gsofa_server* srv, int64_t id, double timeout_s, gsofa_result* out) {
  if (!srv) return fail(GSOFA_ERR_INVALID_ARG, "server is null");
  std::unique_lock<std::mutex> lk(srv->mu);
  auto it = srv->requests.find(id);
  if (it == srv->requests.end()) return fail(GSOFA_ERR_INVALID_ARG, "unknown request id");
  Request* r = it->second.get();
  auto finished = [&] { return r->state != State::Pending && r->state != State::Running; };
  if (timeout_s < 0) srv->cv_done.wait(lk, finished);
  else if (!srv->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), finished))
    return fail(GSOFA_ERR_NOT_READY, "timed out");
  return report(r, 0, out);
}

int gsofa_server_cancel(gsofa_server* srv, int64_t id) {
  if (!srv) return fail(GSOFA_ERR_INVALID_ARG, "server is null");
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->requests.find(id);
  if (it == srv->requests.end()) return fail(GSOFA_ERR_INVALID_ARG, "unknown request id");
  Request* r = it->second.get();
  r->cancel_requested = true;
  if (r->state == State::Pending) {
    for (auto p = srv->pending.begin(); p != srv->pending.end(); ++p)
      if (*p == r) {
        srv->pending.erase(p);
        break;
      }
    r->state = State::Cancelled;
    release_prep(r);
    srv->cv_done.notify_all();
    srv->cv_admit.notify_all();
  }
  return GSOFA_OK;
}

int gsofa_server_release(gsofa_server* srv, int64_t id) {
  if (!srv) return fail(GSOFA_ERR_INVALID_ARG, "server is null");
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->requests.find(id);
  if (it == srv->requests.end()) return fail(GSOFA_ERR_INVALID_ARG, "unknown request id");
  const State st = it->second->state;
  if (st == State::Pending || st == State::Running)
    return fail(GSOFA_ERR_STATE, "request is still queued or running (cancel it first)");
  srv->requests.erase(it);
  return GSOFA_OK;
}

int gsofa_server_retained(gsofa_server* srv, int64_t* requests, int64_t* host_bytes) {
  if (!srv) return fail(GSOFA_ERR_INVALID_ARG, "server is null");
  std::lock_guard<std::mutex> lk(srv->mu);
  int64_t b = 0;
  for (const auto& kv : srv->requests) b += (int64_t)sizeof(Request) + prep_bytes(kv.second->prep);
  if (requests) *requests = (int64_t)srv->requests.size();
  if (host_bytes) *host_bytes = b;
  return GSOFA_OK;
}

int gsofa_server_request_info(gsofa_server* srv, int64_t id, double* latency_ms, double* queue_ms,
                              uint64_t* nzu_digest, uint64_t* supernode_digest) {
  if (!srv) return fail(GSOFA_ERR_INVALID_ARG, "server is null");
  std::lock_guard<std::mutex> lk(srv->mu);
  auto it = srv->requests.find(id);
  if (it == srv->requests.end()) return fail(GSOFA_ERR_INVALID_ARG, "unknown request id");
  Request* r = it->second.get();
  int rc = report(r, 0, nullptr);
  if (rc != GSOFA_OK) return rc;
  using ms = std::chrono::duration<double, std::milli>;
  if (latency_ms) *latency_ms = ms(r->t_done - r->t_submit).count();
  if (queue_ms) *queue_ms = ms(r->t_admit - r->t_submit).count();
  if (nzu_digest) *nzu_digest = r->nzu_digest;
  if (supernode_digest) *supernode_digest = r->sup_digest;
  return GSOFA_OK;
}

int gsofa_server_stats(gsofa_server* srv, int64_t* completed, int32_t* running, int32_t* peak_running,
                       int64_t* pending, int64_t* hbm_in_use, int64_t* hbm_budget) {
  if (!srv) return fail(GSOFA_ERR_INVALID_ARG, "server is null");
  std::lock_guard<std::mutex> lk(srv->mu);
  if (completed) *completed = srv->completed;
  if (running) *running = srv->running;
  if (peak_running) *peak_running = srv->peak_running;
  if (pending) *pending = (int64_t)srv->pending.size();
  if (hbm_in_use) *hbm_in_use = srv->in_use.load();
  if (hbm_budget) *hbm_budget = srv->budget;
  return GSOFA_OK;
}

int gsofa_server_stop(gsofa_server* srv) {
  if (!srv) return GSOFA_OK;
  {
    std::lock_guard<std::mutex> lk(srv->mu);
    srv->stopping = true;
    for (Request* r : srv->pending) {
      r->state = State::Cancelled;
      release_prep(r);
    }
    srv->pending.clear();
    srv->cv_admit.notify_all();
    srv->cv_done.notify_all();
  }
  if (srv->admit_thread.joinable()) srv->admit_thread.join();
  if (srv->done_thread.joinable()) srv->done_thread.join();  // drains running requests
  for (auto& sp : srv->slots) {
    Slot& s = *sp;
    if (s.stream) (void)hipStreamSynchronize(s.stream);
    if (s.start) (void)hipEventDestroy(s.start);
    if (s.done) (void)hipEventDestroy(s.done);
    if (s.stream) (void)hipStreamDestroy(s.stream);
  }
  delete srv;
  return GSOFA_OK;
}

}  // extern "C"
