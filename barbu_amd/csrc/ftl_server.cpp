// Serving loop: resident creatures, admission by HBM budget, continuous
// batching of asynchronous step requests.  No counterpart in the reference
// (one groom stepped per frame from the render thread, src/fx/hair.cc:89-125).
//
// One dispatcher thread owns the device work.  It loops over
//   1. completions: a slot whose batch event has fired hands its requests
//      back (latency = now - submit time) and frees their creatures;
//   2. launches: while a slot is free and requests of idle resident
//      creatures are pending, it takes them in arrival order (one per
//      creature, up to the per-batch traffic cap), applies each request's
//      inputs to its groom and launches one batched step on the slot's
//      streams (tile plan -> pinned descriptors -> upload kernel -> one
//      tile-list kernel per launch class, classes side by side).
// Slots overlap on the device, so requests that arrive while a batch runs
// start as soon as a slot frees, not when the device drains.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <list>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "ftl.h"
#include "ftl_internal.h"
#include "ftl_launch.h"
#include "ftl_server.h"

namespace {

using Clock = std::chrono::steady_clock;

const int kMaxSlots = 4;
const uint64_t kDefaultBatchCap = 8ull << 30;

struct Req {
  uint64_t ticket;
  int id;
  ftl_step_request r;
  Clock::time_point t0;
  uint64_t cps;  // control points of the creature (set when batched)
};

struct Creature {
  ftl_groom* g = nullptr;  // NULL while waiting for memory
  int nstrands = 0, npoints = 0;
  ftl_params params;
  uint64_t bytes = 0;  // device bytes it holds (or is expected to hold)
  bool inflight = false;  // a batch (or a host copy) uses it
  // A step left non-finite values at a strand tip: the creature is left out
  // of batches (its requests complete as FTL_ERR_NONFINITE) until its state
  // is replaced with ftl_server_upload.
  bool quarantined = false;
  // host copies kept while waiting for memory
  std::vector<float> pos, vel, roots;
  bool has_roots = false;
};

struct Slot {
  hipStream_t s = nullptr;
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> side_ev;
  hipEvent_t fork = nullptr, done = nullptr;
  FtlStep* h_steps = nullptr;
  FtlStep* d_steps = nullptr;
  size_t steps_cap = 0;
  FtlSpan* h_spans = nullptr;
  FtlSpan* d_spans = nullptr;
  size_t spans_cap = 0;
  // Non-finite flag of each request of the batch: pinned, coherent host
  // memory the step kernels write directly (the groom's status word is
  // replaced by flags[i] in the batch's descriptors).
  unsigned* h_flags = nullptr;
  unsigned* d_flags = nullptr;
  size_t flags_cap = 0;
  std::vector<Req> reqs;
  bool busy = false;
};

double us_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::micro>(b - a).count();
}

}  // namespace

struct ftl_server {
  ftl_ctx* ctx = nullptr;
  uint64_t budget = 0, batch_cap = 0;
  int nslots = 2;
  // coalescing: a batch waits for this much work, or for its oldest request
  // to be coalesce_us old
  uint64_t batch_target = 0;
  int coalesce_us = 0;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::map<int, Creature> creatures;
  std::deque<int> waiting;  // ids, arrival order
  int next_id = 1;
  uint64_t resident_bytes = 0;
  std::list<Req> pending;
  std::unordered_set<uint64_t> open;
  // tickets that completed without a step (FTL_ERR_NONFINITE: quarantined
  // creature; FTL_ERR_CANCELLED: removed or not admitted), until
  // ftl_server_wait reports them
  std::unordered_map<uint64_t, ftl_status> failed;
  uint64_t next_ticket = 1;
  uint64_t work_gen = 0;  // bumped whenever something may have become launchable
  Slot slots[kMaxSlots];
  std::thread th;
  bool stop = false;
  ftl_status fatal = FTL_OK;
  std::string fatal_msg;
  // statistics
  std::vector<float> lat_us;
  uint64_t submitted = 0, completed = 0, batches = 0, cps = 0, batch_reqs = 0;
  uint64_t cps_finite = 0, dropped = 0, failed_admissions = 0;
  int busy_slots = 0;
  Clock::time_point busy_since, first_submit, last_complete, stats_since;
  double busy_total_s = 0.0;
  bool any_submit = false;
};

namespace {

uint64_t creature_bytes(int nstrands, int npoints, bool roots, const ftl_params& p) {
  uint64_t b = (uint64_t)nstrands * npoints * 32u;
  if (roots) b += (uint64_t)nstrands * 16u;
  if (p.wind_amplitude != 0.0f) b += (uint64_t)FTL_WIND_GRID * FTL_WIND_GRID * FTL_WIND_GRID * 32u;
  return b + 256u;
}

// Makes a waiting creature resident (server lock held).
ftl_status make_resident(ftl_server* s, Creature& c) {
  ftl_groom_desc d{c.nstrands, c.npoints, c.pos.data(), c.vel.data(), c.has_roots ? c.roots.data() : nullptr};
  ftl_groom* g = nullptr;
  ftl_status st = ftl_groom_create(s->ctx, &d, &g);
  if (st != FTL_OK) return st;
  if ((st = ftl_groom_set_params(g, &c.params)) != FTL_OK ||
      (c.has_roots && (st = ftl_groom_set_root_mode(g, FTL_ROOT_REST)) != FTL_OK)) {
    ftl_groom_destroy(g);
    return st;
  }
  c.g = g;
  s->resident_bytes += c.bytes;
  std::vector<float>().swap(c.pos);
  std::vector<float>().swap(c.vel);
  std::vector<float>().swap(c.roots);
  return FTL_OK;
}

// Removes the pending requests of creature id; their tickets complete
// unstepped, ftl_server_wait reporting `why` (lock held).
void drop_requests(ftl_server* s, int id, ftl_status why = FTL_ERR_CANCELLED) {
  for (auto it = s->pending.begin(); it != s->pending.end();) {
    if (it->id == id) {
      s->open.erase(it->ticket);
      s->failed[it->ticket] = why;
      if (it->t0 >= s->stats_since) s->dropped++;
      it = s->pending.erase(it);
    } else {
      ++it;
    }
  }
  s->cv_done.notify_all();
}

// Admits waiting creatures, in arrival order, while they fit (lock held).
void admit_waiting(ftl_server* s) {
  while (!s->waiting.empty()) {
    auto it = s->creatures.find(s->waiting.front());
    if (it == s->creatures.end()) {
      s->waiting.pop_front();
      continue;
    }
    if (s->resident_bytes + it->second.bytes > s->budget) break;
    const ftl_status st = make_resident(s, it->second);
    if (st == FTL_ERR_OUT_OF_MEMORY) break;  // device memory short: retry at the next removal
    s->waiting.pop_front();
    if (st != FTL_OK) {
      // Any other failure is the creature's own: it is dropped (its queued
      // requests complete as FTL_ERR_CANCELLED) and the creatures behind it
      // go on.
      const int id = it->first;
      drop_requests(s, id);
      s->creatures.erase(it);
      s->failed_admissions++;
    }
  }
  s->work_gen++;
  s->cv_work.notify_all();
}

// A batch's descriptors and tiles are uploaded by a kernel on the slot's
// stream, reading pinned coherent host memory (no copy-engine transfer ahead
// of the step kernels).
#ifndef FTL_KERNEL_UPLOAD
#define FTL_KERNEL_UPLOAD 1
#endif
constexpr unsigned kHostUpload = FTL_KERNEL_UPLOAD ? hipHostMallocCoherent : hipHostMallocDefault;

hipError_t ensure_slot(ftl_server* s, Slot& sl, size_t nsteps, size_t nspans, size_t nruns) {
  hipError_t e;
  if (nsteps > sl.steps_cap) {
    const size_t cap = std::max(nsteps, sl.steps_cap * 2);
    if (sl.h_steps) (void)hipHostFree(sl.h_steps);
    if (sl.d_steps) ftl_dev_free(s->ctx, sl.d_steps, sl.steps_cap * sizeof(FtlStep));
    sl.h_steps = nullptr;
    sl.d_steps = nullptr;
    sl.steps_cap = 0;
    if ((e = hipHostMalloc((void**)&sl.h_steps, cap * sizeof(FtlStep), kHostUpload)) != hipSuccess) return e;
    if (ftl_dev_alloc(s->ctx, (void**)&sl.d_steps, cap * sizeof(FtlStep), "server descriptors") != FTL_OK)
      return hipErrorOutOfMemory;
    sl.steps_cap = cap;
  }
  if (nspans > sl.spans_cap) {
    const size_t cap = std::max(nspans, sl.spans_cap * 2);
    if (sl.h_spans) (void)hipHostFree(sl.h_spans);
    if (sl.d_spans) ftl_dev_free(s->ctx, sl.d_spans, sl.spans_cap * sizeof(FtlSpan));
    sl.h_spans = nullptr;
    sl.d_spans = nullptr;
    sl.spans_cap = 0;
    if ((e = hipHostMalloc((void**)&sl.h_spans, cap * sizeof(FtlSpan), kHostUpload)) != hipSuccess) return e;
    if (ftl_dev_alloc(s->ctx, (void**)&sl.d_spans, cap * sizeof(FtlSpan), "server spans") != FTL_OK)
      return hipErrorOutOfMemory;
    sl.spans_cap = cap;
  }
  if (nsteps > sl.flags_cap) {
    const size_t cap = std::max(nsteps, sl.flags_cap * 2);
    if (sl.h_flags) (void)hipHostFree(sl.h_flags);
    sl.h_flags = nullptr;
    sl.d_flags = nullptr;
    sl.flags_cap = 0;
    if ((e = hipHostMalloc((void**)&sl.h_flags, cap * sizeof(unsigned), hipHostMallocCoherent)) != hipSuccess)
      return e;
    if ((e = hipHostGetDevicePointer((void**)&sl.d_flags, sl.h_flags, 0)) != hipSuccess) return e;
    sl.flags_cap = cap;
  }
  while (sl.side.size() + 1 < nruns) {
    hipStream_t st = nullptr;
    hipEvent_t ev = nullptr;
    if ((e = hipStreamCreateWithFlags(&st, hipStreamNonBlocking)) != hipSuccess) return e;
    sl.side.push_back(st);
    if ((e = hipEventCreateWithFlags(&ev, hipEventDisableTiming)) != hipSuccess) return e;
    sl.side_ev.push_back(ev);
  }
  return hipSuccess;
}

void set_fatal(ftl_server* s, ftl_status st, const std::string& msg) {
  if (s->fatal == FTL_OK) {
    s->fatal = st;
    s->fatal_msg = msg;
  }
}

// Launches one batch on slot sl (server lock NOT held; the batch's creatures
// are marked in flight, so nothing else touches their grooms).
ftl_status launch_batch(ftl_server* s, Slot& sl) {
  const size_t n = sl.reqs.size();
  std::vector<ftl_groom*> grooms(n);
  std::vector<float> dts(n);
  for (size_t i = 0; i < n; ++i) {
    const Req& q = sl.reqs[i];
    ftl_groom* g;
    {
      std::lock_guard<std::mutex> lk(s->mu);
      g = s->creatures.at(q.id).g;
    }
    grooms[i] = g;
    dts[i] = q.r.dt;
    if (q.r.has_transform) {
      std::memcpy(g->model, q.r.model, sizeof g->model);
      bool ident = true;
      for (int k = 0; k < 16; ++k) ident = ident && q.r.model[k] == ((k % 5 == 0) ? 1.0f : 0.0f);
      g->model_ident = ident ? 1 : 0;
    }
    if (q.r.ncolliders >= 0) {
      g->ncol = q.r.ncolliders;
      std::memset(g->sph, 0, sizeof g->sph);
      for (int k = 0; k < q.r.ncolliders; ++k)
        for (int c = 0; c < 4; ++c) g->sph[k][c] = q.r.spheres[4 * k + c];
    }
    g->time = q.r.time;
  }
  std::vector<FtlSpan> spans;
  std::vector<ClassRun> runs;
  ftl_status st = ftl_plan_tiles(s->ctx, grooms.data(), (int)n, spans, runs);
  if (st != FTL_OK) return st;
  hipError_t e = ensure_slot(s, sl, n, spans.size(), runs.size());
  if (e != hipSuccess) return ftl_hip_fail(e, "server slot buffers");
  for (size_t i = 0; i < n; ++i) {
    if ((st = ftl_make_step(grooms[i], dts[i], &sl.h_steps[i])) != FTL_OK) return st;
    sl.h_flags[i] = 0u;
    sl.h_steps[i].status = sl.d_flags + i;
  }
  if (!spans.empty()) std::memcpy(sl.h_spans, spans.data(), spans.size() * sizeof(FtlSpan));
  if (FTL_KERNEL_UPLOAD) {
    if ((e = ftl_launch_upload(sl.h_steps, sl.d_steps, n * sizeof(FtlStep), sl.s)) != hipSuccess)
      return ftl_hip_fail(e, "server descriptor upload");
    if ((e = ftl_launch_upload(sl.h_spans, sl.d_spans, spans.size() * sizeof(FtlSpan), sl.s)) != hipSuccess)
      return ftl_hip_fail(e, "server span upload");
  } else {
    if ((e = hipMemcpyAsync(sl.d_steps, sl.h_steps, n * sizeof(FtlStep), hipMemcpyHostToDevice, sl.s)) != hipSuccess)
      return ftl_hip_fail(e, "server descriptor upload");
    if (!spans.empty() && (e = hipMemcpyAsync(sl.d_spans, sl.h_spans, spans.size() * sizeof(FtlSpan),
                                              hipMemcpyHostToDevice, sl.s)) != hipSuccess)
      return ftl_hip_fail(e, "server span upload");
  }
  if ((e = ftl_enqueue_runs(runs, sl.d_steps, sl.d_spans, sl.s, &sl.side, &sl.side_ev, sl.fork)) != hipSuccess)
    return ftl_hip_fail(e, "server batch launch");
  if ((e = hipEventRecord(sl.done, sl.s)) != hipSuccess) return ftl_hip_fail(e, "hipEventRecord");
  return FTL_OK;
}

// Requests of one batch: in arrival order, one per creature, idle resident
// creatures only, up to the traffic cap (lock held).  Requests of a
// quarantined creature complete at once with FTL_ERR_NONFINITE.  With
// coalescing, the selection is taken only once it holds batch_target bytes
// or its oldest request is coalesce_us old; otherwise nothing is taken and
// *deadline says when to look again.
void select_batch(ftl_server* s, Slot& sl, Clock::time_point now, Clock::time_point* deadline) {
  sl.reqs.clear();
  uint64_t bytes = 0;
  std::unordered_set<int> blocked;
  std::vector<std::list<Req>::iterator> take;
  Clock::time_point oldest = now;
  for (auto it = s->pending.begin(); it != s->pending.end();) {
    const int id = it->id;
    Creature& c = s->creatures.at(id);
    if (c.quarantined && !c.inflight && !blocked.count(id)) {
      s->open.erase(it->ticket);
      s->failed[it->ticket] = FTL_ERR_NONFINITE;
      if (it->t0 >= s->stats_since) s->dropped++;
      it = s->pending.erase(it);
      s->cv_done.notify_all();
      continue;
    }
    if (blocked.count(id) || !c.g || c.inflight) {
      blocked.insert(id);  // keeps the creature's later requests behind this one
      ++it;
      continue;
    }
    const uint64_t b = 64ull * (uint64_t)c.nstrands * (uint64_t)c.npoints;
    if (!take.empty() && bytes + b > s->batch_cap) break;
    bytes += b;
    blocked.insert(id);
    oldest = std::min(oldest, it->t0);
    take.push_back(it);
    ++it;
  }
  if (take.empty()) return;
  // (Also with no batch on the device: a window that an idle device bypassed
  // would let one slot launch a small batch whenever the device drains, and
  // the batches would stay small under load.)
  if (s->coalesce_us > 0 && bytes < s->batch_target) {
    const Clock::time_point due = oldest + std::chrono::microseconds(s->coalesce_us);
    if (now < due) {
      *deadline = std::min(*deadline, due);
      return;
    }
  }
  for (auto it : take) {
    Creature& c = s->creatures.at(it->id);
    c.inflight = true;
    it->cps = (uint64_t)c.nstrands * (uint64_t)c.npoints;
    sl.reqs.push_back(*it);
    s->pending.erase(it);
  }
}

// Hands a finished batch's requests back.  Statistics count the requests
// submitted since the last reset, and only batches that ran (ok: a batch
// whose launch or execution failed stops the server, and its requests count
// as neither stepped nor dropped).  A request whose step flagged non-finite
// values quarantines its creature.
void complete_slot(ftl_server* s, Slot& sl, Clock::time_point now, bool ok = true) {
  for (size_t i = 0; i < sl.reqs.size(); ++i) {
    const Req& q = sl.reqs[i];
    const bool bad = ok && sl.h_flags && __atomic_load_n(&sl.h_flags[i], __ATOMIC_ACQUIRE) != 0u;
    if (ok && q.t0 >= s->stats_since) {
      s->lat_us.push_back((float)us_between(q.t0, now));
      s->completed++;
      s->cps += q.cps;
      if (!bad) s->cps_finite += q.cps;
      s->last_complete = now;
    }
    s->open.erase(q.ticket);
    auto it = s->creatures.find(q.id);
    if (it != s->creatures.end()) {
      it->second.inflight = false;
      if (bad) it->second.quarantined = true;
    }
  }
  sl.reqs.clear();
  sl.busy = false;
  if (--s->busy_slots == 0) s->busy_total_s += us_between(s->busy_since, now) * 1e-6;
  s->work_gen++;
  s->cv_done.notify_all();
}

void dispatcher(ftl_server* s) {
  (void)hipSetDevice(s->ctx->device);
  std::unique_lock<std::mutex> lk(s->mu);
  while (!s->stop) {
    const uint64_t gen = s->work_gen;
    bool progressed = false;
    for (int k = 0; k < s->nslots; ++k) {
      Slot& sl = s->slots[k];
      if (!sl.busy) continue;
      const hipError_t e = hipEventQuery(sl.done);
      if (e == hipSuccess) {
        complete_slot(s, sl, Clock::now());
        progressed = true;
      } else if (e != hipErrorNotReady) {
        set_fatal(s, ftl_hip_fail(e, "server batch"), "batch failed on the device");
        complete_slot(s, sl, Clock::now(), false);
      }
    }
    Clock::time_point deadline = Clock::time_point::max();
    if (s->fatal == FTL_OK && !s->pending.empty()) {
      for (int k = 0; k < s->nslots; ++k) {
        Slot& sl = s->slots[k];
        if (sl.busy) continue;
        select_batch(s, sl, Clock::now(), &deadline);
        if (sl.reqs.empty()) break;
        sl.busy = true;
        if (s->busy_slots++ == 0) s->busy_since = Clock::now();
        s->batches++;
        s->batch_reqs += sl.reqs.size();
        lk.unlock();
        const ftl_status st = launch_batch(s, sl);
        lk.lock();
        if (st != FTL_OK) {
          set_fatal(s, st, ftl_last_error());
          complete_slot(s, sl, Clock::now(), false);
        }
        progressed = true;
      }
    }
    if (!progressed) {
      // Batches in flight are polled every 20 us (sooner when a coalescing
      // window closes); otherwise sleep until a submission, admission or
      // release may have made work launchable.
      auto wake = [&] { return s->stop || s->work_gen != gen; };
      if (s->busy_slots > 0) {
        Clock::time_point until = Clock::now() + std::chrono::microseconds(20);
        s->cv_work.wait_until(lk, std::min(until, deadline), wake);
      } else if (deadline != Clock::time_point::max()) {
        s->cv_work.wait_until(lk, deadline, wake);
      } else {
        s->cv_work.wait(lk, wake);
      }
    }
  }
}

Creature* find(ftl_server* s, int id) {
  auto it = s->creatures.find(id);
  return it == s->creatures.end() ? nullptr : &it->second;
}

// Reserves a resident creature for host access: waits for its step in flight
// and keeps the dispatcher off it (lock held on entry and exit).
ftl_status reserve(ftl_server* s, std::unique_lock<std::mutex>& lk, int id, Creature** out) {
  Creature* c = find(s, id);
  if (!c) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "unknown creature %d", id);
  s->cv_done.wait(lk, [&] { c = find(s, id); return !c || !c->inflight; });
  if (!c) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "creature %d was removed", id);
  if (!c->g) return ftl_fail(FTL_ERR_NOT_INITIALIZED, "creature %d waits for memory", id);
  c->inflight = true;
  *out = c;
  return FTL_OK;
}

void release(ftl_server* s, int id) {
  Creature* c = find(s, id);
  if (c) c->inflight = false;
  s->work_gen++;
  s->cv_done.notify_all();
  s->cv_work.notify_all();
}

}  // namespace

extern "C" ftl_status ftl_server_create(ftl_ctx* ctx, const ftl_server_config* cfg, ftl_server** out) {
  if (!ctx || !out) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  *out = nullptr;
  ftl_server_config c{};
  if (cfg) c = *cfg;
  if (c.max_inflight < 0 || c.max_inflight > kMaxSlots)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "max_inflight must be in [0, %d] (got %d)", kMaxSlots, c.max_inflight);
  if (c.coalesce_us < 0 || c.coalesce_us > 1000000)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "coalesce_us must be in [0, 1000000] (got %d)", c.coalesce_us);
  hipError_t e;
  if ((e = hipSetDevice(ctx->device)) != hipSuccess) return ftl_hip_fail(e, "hipSetDevice");
  ftl_server* s = new ftl_server();
  s->ctx = ctx;
  s->nslots = c.max_inflight ? c.max_inflight : 2;
  s->batch_cap = c.batch_cap_bytes ? c.batch_cap_bytes : kDefaultBatchCap;
  s->coalesce_us = c.coalesce_us;
  s->batch_target = c.batch_target_bytes ? std::min(c.batch_target_bytes, s->batch_cap) : s->batch_cap;
  s->budget = c.hbm_budget_bytes;
  if (!s->budget) {
    size_t free_b = 0, total_b = 0;
    if ((e = hipMemGetInfo(&free_b, &total_b)) != hipSuccess) {
      delete s;
      return ftl_hip_fail(e, "hipMemGetInfo");
    }
    s->budget = (uint64_t)(0.8 * (double)free_b);
  }
  for (int k = 0; k < s->nslots; ++k) {
    Slot& sl = s->slots[k];
    ftl_status st = FTL_OK;
    if ((e = hipStreamCreateWithFlags(&sl.s, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&sl.fork, hipEventDisableTiming)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming)) != hipSuccess)
      st = ftl_hip_fail(e, "server streams");
    if (st != FTL_OK) {
      ftl_server_destroy(s);
      return st;
    }
  }
  {
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->live_objects++;
  }
  s->th = std::thread(dispatcher, s);
  *out = s;
  return FTL_OK;
}

extern "C" void ftl_server_destroy(ftl_server* s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
    s->work_gen++;
  }
  s->cv_work.notify_all();
  const bool started = s->th.joinable();
  if (started) s->th.join();
  (void)hipSetDevice(s->ctx->device);
  for (int k = 0; k < kMaxSlots; ++k) {
    Slot& sl = s->slots[k];
    if (sl.s) (void)hipStreamSynchronize(sl.s);
    for (hipStream_t st : sl.side) (void)hipStreamSynchronize(st);
  }
  for (auto& kv : s->creatures)
    if (kv.second.g) ftl_groom_destroy(kv.second.g);
  for (int k = 0; k < kMaxSlots; ++k) {
    Slot& sl = s->slots[k];
    for (hipStream_t st : sl.side) (void)hipStreamDestroy(st);
    for (hipEvent_t ev : sl.side_ev) (void)hipEventDestroy(ev);
    if (sl.s) (void)hipStreamDestroy(sl.s);
    if (sl.fork) (void)hipEventDestroy(sl.fork);
    if (sl.done) (void)hipEventDestroy(sl.done);
    if (sl.h_steps) (void)hipHostFree(sl.h_steps);
    if (sl.h_spans) (void)hipHostFree(sl.h_spans);
    if (sl.h_flags) (void)hipHostFree(sl.h_flags);
    ftl_dev_free(s->ctx, sl.d_steps, sl.steps_cap * sizeof(FtlStep));
    ftl_dev_free(s->ctx, sl.d_spans, sl.spans_cap * sizeof(FtlSpan));
  }
  if (started) {
    std::lock_guard<std::mutex> lk(s->ctx->mu);
    s->ctx->live_objects--;
  }
  delete s;
}

extern "C" ftl_status ftl_server_add(ftl_server* s, const ftl_groom_desc* d, const ftl_params* p, int* id,
                                     int* resident) {
  if (!s || !d || !p || !id) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  if (d->nstrands <= 0 || d->npoints < 2 || d->npoints > FTL_MAX_POINTS || !d->positions)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "bad groom description (nstrands %d, npoints %d)", d->nstrands,
                    d->npoints);
  // Validated now, whether or not the creature becomes resident at once: a
  // queued creature with bad parameters would otherwise fail only when
  // admitted, and block every creature behind it.
  {
    const ftl_status pst = ftl_check_params(p);
    if (pst != FTL_OK) return pst;
  }
  Creature c;
  c.nstrands = d->nstrands;
  c.npoints = d->npoints;
  c.params = *p;
  c.has_roots = d->rest_roots != nullptr;
  c.bytes = creature_bytes(c.nstrands, c.npoints, c.has_roots, c.params);
  std::lock_guard<std::mutex> lk(s->mu);
  if (c.bytes > s->budget)
    return ftl_fail(FTL_ERR_BUDGET, "creature needs %llu bytes, more than the whole budget (%llu)",
                    (unsigned long long)c.bytes, (unsigned long long)s->budget);
  const size_t n4 = (size_t)c.nstrands * c.npoints * 4;
  c.pos.assign(d->positions, d->positions + n4);
  if (d->velocities)
    c.vel.assign(d->velocities, d->velocities + n4);
  else
    c.vel.assign(n4, 0.0f);
  if (c.has_roots) c.roots.assign(d->rest_roots, d->rest_roots + (size_t)c.nstrands * 4);
  const int cid = s->next_id++;
  Creature& ref = s->creatures.emplace(cid, std::move(c)).first->second;
  bool now = false;
  if (s->waiting.empty() && s->resident_bytes + ref.bytes <= s->budget) {
    const ftl_status st = make_resident(s, ref);
    if (st != FTL_OK && st != FTL_ERR_OUT_OF_MEMORY) {
      s->creatures.erase(cid);
      return st;
    }
    now = st == FTL_OK;
  }
  if (!now) s->waiting.push_back(cid);
  *id = cid;
  if (resident) *resident = now ? 1 : 0;
  return FTL_OK;
}

extern "C" ftl_status ftl_server_remove(ftl_server* s, int id) {
  if (!s) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "server is NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  Creature* c = find(s, id);
  if (!c) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "unknown creature %d", id);
  s->cv_done.wait(lk, [&] { c = find(s, id); return !c || !c->inflight; });
  if (!c) return FTL_OK;
  drop_requests(s, id);
  if (c->g) {
    ftl_groom_destroy(c->g);
    s->resident_bytes -= c->bytes;
  } else {
    s->waiting.erase(std::remove(s->waiting.begin(), s->waiting.end(), id), s->waiting.end());
  }
  s->creatures.erase(id);
  admit_waiting(s);
  s->cv_done.notify_all();
  return FTL_OK;
}

extern "C" ftl_status ftl_server_resident(ftl_server* s, int id, int* resident) {
  if (!s || !resident) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  Creature* c = find(s, id);
  if (!c) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "unknown creature %d", id);
  *resident = c->g ? 1 : 0;
  return FTL_OK;
}

static ftl_status check_request(const ftl_step_request& r) {
  if (!(r.dt > 0.0f) || !std::isfinite(r.dt)) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "dt must be > 0");
  if (!std::isfinite(r.time)) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "non-finite time");
  if (r.has_transform) {
    if (r.model[3] != 0.0f || r.model[7] != 0.0f || r.model[11] != 0.0f || r.model[15] != 1.0f)
      return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "root transform must be affine (last row 0,0,0,1)");
    for (int i = 0; i < 16; ++i)
      if (!std::isfinite(r.model[i])) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "non-finite transform");
  }
  if (r.ncolliders < -1 || r.ncolliders > FTL_MAX_COLLIDERS)
    return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "ncolliders must be in [-1, %d]", FTL_MAX_COLLIDERS);
  for (int i = 0; i < 4 * r.ncolliders; ++i)
    if (!std::isfinite(r.spheres[i])) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "non-finite collider");
  return FTL_OK;
}

extern "C" ftl_status ftl_server_submit(ftl_server* s, int n, const int* ids, const ftl_step_request* reqs,
                                        uint64_t* tickets) {
  if (!s || n < 0 || (n > 0 && (!ids || !reqs))) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "bad arguments");
  for (int i = 0; i < n; ++i) {
    const ftl_status st = check_request(reqs[i]);
    if (st != FTL_OK) return st;
  }
  const Clock::time_point now = Clock::now();
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->fatal != FTL_OK) return ftl_fail(s->fatal, "server stopped: %s", s->fatal_msg.c_str());
    for (int i = 0; i < n; ++i)
      if (!find(s, ids[i])) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "unknown creature %d", ids[i]);
    if (!s->any_submit) {
      s->any_submit = true;
      s->first_submit = now;
    }
    for (int i = 0; i < n; ++i) {
      const uint64_t t = s->next_ticket++;
      s->pending.push_back(Req{t, ids[i], reqs[i], now});
      s->open.insert(t);
      if (tickets) tickets[i] = t;
    }
    s->submitted += (uint64_t)n;
    s->work_gen++;
  }
  s->cv_work.notify_all();
  return FTL_OK;
}

extern "C" ftl_status ftl_server_wait(ftl_server* s, uint64_t ticket, int timeout_ms) {
  if (!s) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "server is NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  auto done = [&] { return !s->open.count(ticket) || s->fatal != FTL_OK; };
  if (timeout_ms < 0)
    s->cv_done.wait(lk, done);
  else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done))
    return ftl_fail(FTL_ERR_BUSY, "request %llu still running", (unsigned long long)ticket);
  if (s->fatal != FTL_OK) return ftl_fail(s->fatal, "server stopped: %s", s->fatal_msg.c_str());
  auto f = s->failed.find(ticket);
  if (f != s->failed.end()) {
    const ftl_status st = f->second;
    s->failed.erase(f);
    return ftl_fail(st, "request %llu was not stepped: %s", (unsigned long long)ticket,
                    st == FTL_ERR_NONFINITE ? "its creature is quarantined (non-finite state)"
                                            : "its creature was removed or could not be made resident");
  }
  return FTL_OK;
}

extern "C" ftl_status ftl_server_drain(ftl_server* s, int timeout_ms) {
  if (!s) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "server is NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  // Requests of creatures still waiting for memory cannot complete.
  auto done = [&] {
    if (s->fatal != FTL_OK) return true;
    for (const Req& q : s->pending) {
      Creature* c = find(s, q.id);
      if (c && c->g) return false;
    }
    for (int k = 0; k < s->nslots; ++k)
      if (s->slots[k].busy) return false;
    return true;
  };
  if (timeout_ms < 0)
    s->cv_done.wait(lk, done);
  else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), done))
    return ftl_fail(FTL_ERR_BUSY, "server still busy after %d ms", timeout_ms);
  if (s->fatal != FTL_OK) return ftl_fail(s->fatal, "server stopped: %s", s->fatal_msg.c_str());
  return FTL_OK;
}

extern "C" ftl_status ftl_server_download(ftl_server* s, int id, float* pos, float* vel) {
  if (!s) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "server is NULL");
  std::unique_lock<std::mutex> lk(s->mu);
  Creature* c = nullptr;
  ftl_status st = reserve(s, lk, id, &c);
  if (st != FTL_OK) return st;
  ftl_groom* g = c->g;
  lk.unlock();
  st = ftl_groom_download(g, pos, vel);
  lk.lock();
  release(s, id);
  return st;
}

extern "C" ftl_status ftl_server_status(ftl_server* s, int id, int* nonfinite) {
  if (!s || !nonfinite) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  Creature* c = nullptr;
  ftl_status st = reserve(s, lk, id, &c);
  if (st != FTL_OK) return st;
  *nonfinite = c->quarantined ? 1 : 0;
  release(s, id);
  return FTL_OK;
}

extern "C" ftl_status ftl_server_upload(ftl_server* s, int id, const float* pos, const float* vel) {
  if (!s || !pos || !vel) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  std::unique_lock<std::mutex> lk(s->mu);
  Creature* c = nullptr;
  ftl_status st = reserve(s, lk, id, &c);
  if (st != FTL_OK) return st;
  ftl_groom* g = c->g;
  lk.unlock();
  st = ftl_groom_upload(g, pos, vel);
  lk.lock();
  c = find(s, id);
  if (c && st == FTL_OK) c->quarantined = false;
  release(s, id);
  return st;
}

extern "C" ftl_status ftl_server_get_stats(ftl_server* s, ftl_server_stats* out) {
  if (!s || !out) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "NULL argument");
  std::lock_guard<std::mutex> lk(s->mu);
  ftl_server_stats r{};
  r.submitted = s->submitted;
  r.completed = s->completed;
  r.batches = s->batches;
  r.cps_stepped = s->cps;
  if (!s->lat_us.empty()) {
    std::vector<float> v = s->lat_us;
    auto pct = [&](double q) {
      const size_t k = std::min(v.size() - 1, (size_t)std::floor(q * (double)(v.size() - 1) + 0.5));
      std::nth_element(v.begin(), v.begin() + k, v.end());
      return (double)v[k];
    };
    r.p50_us = pct(0.50);
    r.p99_us = pct(0.99);
    double sum = 0.0, mx = 0.0;
    for (float x : s->lat_us) {
      sum += x;
      mx = std::max(mx, (double)x);
    }
    r.mean_us = sum / (double)s->lat_us.size();
    r.max_us = mx;
  }
  r.mean_batch_requests = s->batches ? (double)s->batch_reqs / (double)s->batches : 0.0;
  r.busy_s = s->busy_total_s + (s->busy_slots > 0 ? us_between(s->busy_since, Clock::now()) * 1e-6 : 0.0);
  r.span_s = s->any_submit && s->completed ? us_between(s->first_submit, s->last_complete) * 1e-6 : 0.0;
  r.resident_bytes = s->resident_bytes;
  int res = 0;
  for (const auto& kv : s->creatures) res += kv.second.g ? 1 : 0;
  r.resident = res;
  r.waiting = (int)s->waiting.size();
  r.cps_finite = s->cps_finite;
  r.dropped = s->dropped;
  r.failed_admissions = s->failed_admissions;
  int q = 0;
  for (const auto& kv : s->creatures) q += kv.second.quarantined ? 1 : 0;
  r.quarantined = q;
  *out = r;
  return FTL_OK;
}

extern "C" ftl_status ftl_server_reset_stats(ftl_server* s) {
  if (!s) return ftl_fail(FTL_ERR_INVALID_ARGUMENT, "server is NULL");
  std::lock_guard<std::mutex> lk(s->mu);
  s->lat_us.clear();
  s->submitted = s->completed = s->batches = s->cps = s->batch_reqs = 0;
  s->cps_finite = s->dropped = 0;
  s->busy_total_s = 0.0;
  if (s->busy_slots > 0) s->busy_since = Clock::now();
  s->any_submit = false;
  s->stats_since = Clock::now();
  return FTL_OK;
}
