// Serving loop: independent solve requests arrive one at a time and are
// served with continuous batching.
//
// The reference is one synchronous call per matrix (ts::zquatev,
// zquatev.cc:42-100); a caller with a stream of requests loops over it.
// Here zq_server_submit only queues a request (device pointers, as
// zq_zquatev_device takes them).  A scheduler thread takes whatever has
// arrived and runs it as one batched call on a free lane: the oldest request
// picks the batch's jobz and size class (half-dimension up to 192, up to 640,
// up to 1024, up to 1450, larger: a lockstep batch runs as long as its
// largest member, so requests are not held behind much larger ones), then
// further requests of that jobz and class join in arrival order, up to
// max_batch problems and max_bytes of workspace.  Batches of the large
// classes (half-dimension above 640) occupy at most all lanes but one, so
// small requests always find a lane soon.
//
// Device memory: the lanes' workspaces together stay within a budget (by
// default 80% of the device memory free when the server is created).  A lane
// whose batch needs a larger workspace allocates it at once and keeps the old
// one (retired) until no lane has work in flight: hipFree waits for the whole
// device, so freeing while other lanes run would stall them.  Retired bytes
// count against the budget until they are freed.  A batch that would exceed
// the budget is cut to what fits; if not even its first request fits while
// other lanes run, it waits for them, and with the device idle the retired
// and idle workspaces are freed first.  A request that alone needs more than
// the budget completes at once with ZQ_ERR_BUDGET.  Each lane has its own
// thread that enqueues its batch (a large batch is thousands of launches),
// so lanes never wait for each other's host work.  A
// lane is a context with its own stream and resident workspace, so batches on
// different lanes overlap on the device.  A completion thread watches each
// batch's event and stamps its requests; zq_server_poll hands the records
// out.  Results are those of the batched call, so each request's output is
// bitwise that of its own zq_zquatev_device call with the same panel width.
#include "zq_internal.h"
#include "zq.h"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <list>
#include <mutex>
#include <thread>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

// pinned descriptor staging each lane allocates ahead (per buffer): the task
// tables of a batch of 128 problems of 2n <= 4096 take well under this
constexpr size_t kStageReserve = size_t(16) << 20;

struct Req {
  unsigned long long ticket;
  int n2, ld2, jobz;
  double* D;
  double* eig;
  double t_submit;
};

struct Batch {
  std::vector<Req> reqs;
  int lane = -1;
  int cls = 0;
  int rc = 0;
  double t_start = 0.0;
};

}  // namespace

struct zq_server {
  int device = 0;
  int max_batch = 64;
  size_t max_bytes = 0;
  size_t budget = 0;                 // device workspace bytes all lanes together may hold
  size_t default_budget = 0;
  std::vector<zq_context*> lanes;
  std::vector<size_t> lane_ws;       // workspace bytes each lane holds (or is about to)
  size_t retired = 0;                // replaced workspaces not freed yet (all lanes)
  std::vector<hipEvent_t> lane_ev;
  std::vector<int*> lane_info;       // pinned, max_batch ints per lane
  std::vector<int> lane_busy;
  std::vector<int> lane_cls;         // size class of the lane's batch, -1 when free
  std::vector<Batch> lane_job;       // batch handed to the lane's thread
  std::vector<int> lane_has_job;
  std::vector<std::thread> lane_threads;
  std::condition_variable cv_lane;   // lane threads: a batch was handed over
  std::mutex mu;
  std::condition_variable cv_work;   // scheduler: requests or a free lane
  std::condition_variable cv_fly;    // completion thread: a batch in flight
  std::condition_variable cv_done;   // drain
  std::list<Req> pending;            // a list: erasing a batch's members keeps the other iterators valid
  std::deque<Batch> inflight;
  std::deque<zq_completion> done;
  unsigned long long next_ticket = 1;
  long outstanding = 0;
  bool stop = false;
  std::thread sched, compl_;
  Clock::time_point t0;

  double now_us() const { return std::chrono::duration<double, std::micro>(Clock::now() - t0).count(); }

  // classes 2 and above are the large ones
  static int size_class(int n2) { return n2 <= 384 ? 0 : (n2 <= 1280 ? 1 : (n2 <= 2048 ? 2 : (n2 <= 2900 ? 3 : 4))); }

  int free_lane() const {
    for (size_t k = 0; k < lane_busy.size(); ++k)
      if (!lane_busy[k]) return (int)k;
    return -1;
  }

  // Workspace a batch of the first k requests of n2s needs (grows with k).
  size_t need_of(const std::vector<int>& n2s, int k, int jobz) const {
    return zq::batched_workspace(lanes[0], k, n2s.data(), jobz);
  }
  // Largest k in [0, kmax] with need_of(k) <= cap (0 if even one does not fit).
  int fit_prefix(const std::vector<int>& n2s, int kmax, int jobz, size_t cap) const {
    int lo = 0, hi = kmax;
    while (lo < hi) {
      const int mid = (lo + hi + 1) / 2;
      if (need_of(n2s, mid, jobz) <= cap) lo = mid;
      else hi = mid - 1;
    }
    return lo;
  }

  void complete_now(const Req& r, int info) {
    const double t = now_us();
    done.push_back(zq_completion{r.ticket, info, 0, -1, r.t_submit, t, t});
    --outstanding;
    cv_done.notify_all();
  }

  bool all_idle() const {
    for (int b : lane_busy)
      if (b) return false;
    return true;
  }

  // With no work in flight: free every retired workspace, and the workspaces
  // of the lanes other than `keep` (-1: of every lane).
  void free_idle_memory(int keep) {
    for (size_t q = 0; q < lanes.size(); ++q) {
      (void)zq::context_free_retired(lanes[q]);
      if ((int)q != keep && lane_ws[q] > 0) {
        (void)zq::context_release_workspace(lanes[q]);
        lane_ws[q] = 0;
      }
    }
    retired = 0;
  }

  void run_scheduler() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      cv_work.wait(lk, [&] { return stop || (!pending.empty() && free_lane() >= 0); });
      if (retired > 0 && all_idle()) {
        for (zq_context* c : lanes) (void)zq::context_free_retired(c);
        retired = 0;
      }
      if (pending.empty() || free_lane() < 0) {
        if (stop) break;
        continue;
      }
      // the head request picks the batch; large-class batches may hold at
      // most lanes - 1 lanes, so one lane always stays open to small requests
      int busy = 0, big_busy = 0;
      for (size_t k = 0; k < lanes.size(); ++k) {
        busy += lane_busy[k];
        big_busy += lane_cls[k] >= 2 ? 1 : 0;
      }
      auto head = pending.begin();
      if (lanes.size() > 1 && big_busy + 1 >= (int)lanes.size()) {
        while (head != pending.end() && size_class(head->n2) >= 2) ++head;
        if (head == pending.end()) {
          // only large requests wait: let the running batches finish first
          cv_work.wait_for(lk, std::chrono::microseconds(200));
          continue;
        }
      }
      const int jobz = head->jobz;
      const int cls = size_class(head->n2);
      std::vector<std::list<Req>::iterator> cand;
      std::vector<int> n2s;
      for (auto it = head; it != pending.end() && (int)cand.size() < max_batch; ++it) {
        if (it->jobz != jobz || size_class(it->n2) != cls) continue;
        cand.push_back(it);
        n2s.push_back(it->n2);
      }
      if (need_of(n2s, 1, jobz) > budget) {   // can never be served
        complete_now(*head, ZQ_ERR_BUDGET);
        pending.erase(head);
        continue;
      }
      int k = (int)cand.size();
      if (max_bytes > 0 && k > 1) k = std::max(1, fit_prefix(n2s, k, jobz, max_bytes));
      // the free lane holding the largest workspace
      int L = -1;
      size_t total = retired;
      for (size_t q = 0; q < lanes.size(); ++q) {
        total += lane_ws[q];
        if (!lane_busy[q] && (L < 0 || lane_ws[q] > lane_ws[L])) L = (int)q;
      }
      if (need_of(n2s, k, jobz) > lane_ws[L]) {
        // the lane grows: the new workspace is allocated beside the retired old one
        int k2 = total < budget ? fit_prefix(n2s, k, jobz, budget - total) : 0;
        if (k2 == 0 && busy == 0) {
          // the device is idle: free every workspace, the lane's own included
          // (it is about to be replaced), and the retired ones
          free_idle_memory(-1);
          k2 = std::max(1, fit_prefix(n2s, k, jobz, budget));
        }
        if (k2 == 0) {
          // no room to grow beside the running batches: take what the lane holds
          k2 = fit_prefix(n2s, k, jobz, lane_ws[L]);
          if (k2 == 0) {
            // even the head needs more: wait for the running batches (nothing
            // else is dispatched meanwhile, so the lanes drain)
            cv_work.wait(lk, [&] { return stop || all_idle(); });
            continue;
          }
        }
        k = k2;
      }
      if (need_of(n2s, k, jobz) > lane_ws[L]) {
        retired += lane_ws[L];
        lane_ws[L] = need_of(n2s, k, jobz);
      }
      lane_busy[L] = 1;
      Batch b;
      b.lane = L;
      b.cls = cls;
      b.reqs.reserve(k);
      for (int q = 0; q < k; ++q) b.reqs.push_back(*cand[q]);
      for (int q = 0; q < k; ++q) pending.erase(cand[q]);
      lane_cls[L] = cls;
      lane_job[L] = std::move(b);
      lane_has_job[L] = 1;
      cv_lane.notify_all();
    }
  }

  void run_lane(int L) {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      cv_lane.wait(lk, [&] { return stop || lane_has_job[L]; });
      if (!lane_has_job[L]) {
        if (stop) break;
        continue;
      }
      Batch b = std::move(lane_job[L]);
      lane_has_job[L] = 0;
      lk.unlock();
      const int cnt = (int)b.reqs.size();
      b.t_start = now_us();
      try {
        std::vector<int> n2s(cnt), ld2(cnt);
        std::vector<double*> D(cnt), eig(cnt);
        for (int k = 0; k < cnt; ++k) {
          n2s[k] = b.reqs[k].n2;
          ld2[k] = b.reqs[k].ld2;
          D[k] = b.reqs[k].D;
          eig[k] = b.reqs[k].eig;
        }
        b.rc = zq::batched_async(lanes[L], cnt, n2s.data(), D.data(), ld2.data(), eig.data(), b.reqs[0].jobz,
                                 lane_info[L]);
      } catch (...) {   // the batch's requests complete with this status
        b.rc = exception_status();
      }
      if (hipEventRecord(lane_ev[L], zq::context_stream(lanes[L])) != hipSuccess && b.rc == 0) b.rc = ZQ_ERR_HIP;
      lk.lock();
      inflight.push_back(std::move(b));
      cv_fly.notify_one();
    }
  }

  void run_completion() {
    (void)hipSetDevice(device);
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
      cv_fly.wait(lk, [&] { return stop || !inflight.empty(); });
      if (inflight.empty()) {
        if (stop) break;
        continue;
      }
      // find a finished batch (lanes may finish out of launch order)
      int hit = -1;
      for (size_t k = 0; k < inflight.size() && hit < 0; ++k) {
        const hipError_t q = hipEventQuery(lane_ev[inflight[k].lane]);
        if (q == hipSuccess) hit = (int)k;
        else if (q != hipErrorNotReady) {
          if (inflight[k].rc == 0) inflight[k].rc = ZQ_ERR_HIP;
          hit = (int)k;
        }
      }
      if (hit < 0) {
        lk.unlock();
        std::this_thread::sleep_for(std::chrono::microseconds(20));
        lk.lock();
        continue;
      }
      Batch b = std::move(inflight[hit]);
      inflight.erase(inflight.begin() + hit);
      const double t = now_us();
      const int* info = lane_info[b.lane];
      for (size_t k = 0; k < b.reqs.size(); ++k) {
        const Req& r = b.reqs[k];
        zq_completion c;
        c.ticket = r.ticket;
        c.info = b.rc != 0 ? b.rc : info_code(info[k]);
        c.batch_size = (int)b.reqs.size();
        c.lane = b.lane;
        c.t_submit_us = r.t_submit;
        c.t_start_us = b.t_start;
        c.t_done_us = t;
        done.push_back(c);
      }
      outstanding -= (long)b.reqs.size();
      lane_busy[b.lane] = 0;
      lane_cls[b.lane] = -1;
      cv_work.notify_all();
      cv_done.notify_all();
    }
  }
};

extern "C" {

int zq_server_create(zq_server** out, int device, int lanes, int max_batch, double max_batch_gb) try {
  if (!out) return ZQ_ERR_ARG;
  *out = nullptr;
  if (lanes < 1 || lanes > 8) return -3;
  if (max_batch < 1) return -4;
  if (max_batch_gb < 0.0) return -5;
  if (hipSetDevice(device) != hipSuccess) return ZQ_ERR_HIP;
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) return ZQ_ERR_HIP;
  zq_server* s = new zq_server();
  s->device = device;
  s->max_batch = max_batch;
  s->max_bytes = (size_t)(max_batch_gb * 1073741824.0);
  s->default_budget = (size_t)(0.8 * (double)fr);
  s->budget = s->default_budget;
  s->t0 = Clock::now();
  try {   // the per-lane containers, sized once so that the loop below cannot throw
    s->lanes.reserve(lanes);
    s->lane_ws.reserve(lanes);
    s->lane_ev.reserve(lanes);
    s->lane_info.reserve(lanes);
    s->lane_busy.reserve(lanes);
    s->lane_cls.reserve(lanes);
    s->lane_has_job.reserve(lanes);
    s->lane_job.reserve(lanes);
    s->lane_threads.reserve(lanes);
  } catch (...) {
    delete s;
    throw;
  }
  int rc = ZQ_OK;
  for (int k = 0; k < lanes && rc == ZQ_OK; ++k) {
    zq_context* c = nullptr;
    rc = zq_create(&c, device);
    if (rc) break;
    s->lanes.push_back(c);
    s->lane_ws.push_back(0);
    zq::context_set_defer_free(c, true);
    if ((rc = zq::context_reserve(c, max_batch, kStageReserve))) break;
    hipEvent_t e = nullptr;
    void* h = nullptr;
    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess ||
        hipHostMalloc(&h, (size_t)max_batch * sizeof(int), hipHostMallocDefault) != hipSuccess) {
      if (e) (void)hipEventDestroy(e);
      rc = ZQ_ERR_HIP;
      break;
    }
    s->lane_ev.push_back(e);
    s->lane_info.push_back(static_cast<int*>(h));
    s->lane_busy.push_back(0);
    s->lane_cls.push_back(-1);
    s->lane_has_job.push_back(0);
  }
  s->lane_job.resize(s->lanes.size());
  auto release = [s] {
    for (auto c : s->lanes) zq_destroy(c);
    for (auto e : s->lane_ev) (void)hipEventDestroy(e);
    for (auto h : s->lane_info) (void)hipHostFree(h);
    delete s;
  };
  if (rc) {
    release();
    return rc;
  }
  try {
    for (int k = 0; k < (int)s->lanes.size(); ++k) s->lane_threads.emplace_back([s, k] { s->run_lane(k); });
    s->sched = std::thread([s] { s->run_scheduler(); });
    s->compl_ = std::thread([s] { s->run_completion(); });
  } catch (...) {   // a thread could not be started: stop and join those that were
    rc = exception_status();
    {
      std::lock_guard<std::mutex> lk(s->mu);
      s->stop = true;
    }
    s->cv_work.notify_all();
    s->cv_fly.notify_all();
    s->cv_lane.notify_all();
    for (auto& t : s->lane_threads)
      if (t.joinable()) t.join();
    if (s->sched.joinable()) s->sched.join();
    if (s->compl_.joinable()) s->compl_.join();
    release();
    return rc;
  }
  *out = s;
  return ZQ_OK;
} catch (...) {
  return exception_status();
}

int zq_server_submit(zq_server* s, int n2, double* dD, int ld2, double* deig, int jobz, unsigned long long* ticket) try {
  if (!s) return ZQ_ERR_ARG;
  if (n2 < 0 || (n2 & 1)) return -2;
  if (n2 > 0 && !dD) return -3;
  if (ld2 < (n2 > 1 ? n2 : 1)) return -4;
  if (n2 > 0 && !deig) return -5;
  if (jobz != 'V' && jobz != 'v' && jobz != 'N' && jobz != 'n') return -6;
  std::lock_guard<std::mutex> lk(s->mu);
  const unsigned long long t = s->next_ticket++;
  if (ticket) *ticket = t;
  const double now = s->now_us();
  if (n2 == 0) {
    s->done.push_back(zq_completion{t, 0, 0, -1, now, now, now});
    return ZQ_OK;
  }
  s->pending.push_back(Req{t, n2, ld2, jobz, dD, deig, now});
  ++s->outstanding;
  s->cv_work.notify_one();
  return ZQ_OK;
} catch (...) {
  return exception_status();
}

int zq_server_poll(zq_server* s, zq_completion* out, int max) {
  if (!s || (max > 0 && !out)) return ZQ_ERR_ARG;
  std::lock_guard<std::mutex> lk(s->mu);
  int k = 0;
  while (k < max && !s->done.empty()) {
    out[k++] = s->done.front();
    s->done.pop_front();
  }
  return k;
}

int zq_server_drain(zq_server* s) {
  if (!s) return ZQ_ERR_ARG;
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->outstanding == 0; });
  return ZQ_OK;
}

double zq_server_clock_us(zq_server* s) { return s ? s->now_us() : 0.0; }

int zq_server_set_budget(zq_server* s, double budget_gb) {
  if (!s) return ZQ_ERR_ARG;
  if (budget_gb < 0.0) return -2;
  std::lock_guard<std::mutex> lk(s->mu);
  s->budget = budget_gb > 0.0 ? (size_t)(budget_gb * 1073741824.0) : s->default_budget;
  s->cv_work.notify_all();
  return ZQ_OK;
}

double zq_server_budget_gb(zq_server* s) {
  if (!s) return 0.0;
  std::lock_guard<std::mutex> lk(s->mu);
  return (double)s->budget / 1073741824.0;
}

int zq_server_destroy(zq_server* s) {
  if (!s) return ZQ_OK;
  zq_server_drain(s);
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  s->cv_fly.notify_all();
  s->cv_lane.notify_all();
  s->sched.join();
  for (auto& t : s->lane_threads) t.join();
  s->compl_.join();
  for (auto c : s->lanes) zq_destroy(c);
  for (auto e : s->lane_ev) (void)hipEventDestroy(e);
  for (auto h : s->lane_info) (void)hipHostFree(h);
  delete s;
  return ZQ_OK;
}

}  // extern "C"
