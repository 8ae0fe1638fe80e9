// Batched stepping and the serving loop (include/mts.h: mts_batch_run,
// mts_server_*), built on the instance entry points.
//
// Every instance runs on its own non-blocking stream, so the steps of
// different instances overlap on the device.  The host issues one step per
// instance per round and keeps at most kDepth steps of an instance queued
// (an event ring per instance), so a long solve cannot starve the others and
// new solves join between rounds (continuous batching).  Solves are admitted
// in arrival order while the sum of their device footprints stays under the
// HBM cap and the in-flight count under its limit.
#include "mts.h"

#include "abi_internal.hpp"
#include "host_parallel.hpp"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

constexpr int kDepth = 4;    // queued steps per instance
constexpr int kWorkers = 2;  // serving: instance creation / finishing threads

using Clock = std::chrono::steady_clock;
double ms_between(Clock::time_point a, Clock::time_point b) {
  return std::chrono::duration<double, std::milli>(b - a).count();
}

// (int)(t_finish / dt) steps with t += dt from t0, as the reference's loops.
int step_count(double dt, double t_finish) { return (int)(t_finish / dt); }

// One solve in flight: instance, event ring, progress.
struct Flight {
  mts_instance *inst = nullptr;
  hipStream_t stream = nullptr;
  hipEvent_t ev[kDepth] = {};
  int steps = 0, issued = 0;
  double t = 0.0, dt = 0.0;

  int start(mts_instance *in, double dt_, double t0, int nsteps) {
    inst = in;
    dt = dt_;
    t = t0;
    steps = nsteps;
    issued = 0;
    void *s = nullptr;
    int rc = mts_field_device(in, nullptr, nullptr, &s);
    if (rc) return rc;
    stream = static_cast<hipStream_t>(s);
    for (int i = 0; i < kDepth; i++)
      if (hipEventCreateWithFlags(&ev[i], hipEventDisableTiming) != hipSuccess)
        return mts::set_error(MTS_E_DEVICE, "hipEventCreateWithFlags failed");
    return MTS_OK;
  }
  ~Flight() {
    for (int i = 0; i < kDepth; i++)
      if (ev[i]) (void)hipEventDestroy(ev[i]);
  }
  // Issues the next step unless kDepth steps are still queued.  1: issued,
  // 0: throttled or finished, <0: error.
  int advance() {
    if (issued >= steps) return 0;
    if (issued >= kDepth && hipEventQuery(ev[issued % kDepth]) == hipErrorNotReady) return 0;
    t += dt;
    int rc = mts_step(inst, t);
    if (rc) return rc;
    if (hipEventRecord(ev[issued % kDepth], stream) != hipSuccess)
      return mts::set_error(MTS_E_DEVICE, "hipEventRecord failed");
    issued++;
    return 1;
  }
  bool issued_all() const { return issued >= steps; }
  bool done() const {
    if (!issued_all()) return false;
    if (steps == 0) return true;
    return hipEventQuery(ev[(steps - 1) % kDepth]) != hipErrorNotReady;
  }
};

}  // namespace

extern "C" int mts_batch_run(mts_instance *const *insts, int count, const double *t0,
                             const double *t_finish, int *steps, double *t_end) {
  if (count < 0 || (count > 0 && (!insts || !t0 || !t_finish)))
    return mts::set_error(MTS_E_ARG, "null argument or negative count");
  std::vector<std::unique_ptr<Flight>> fl;
  for (int i = 0; i < count; i++) {
    double dt = 0.0;
    int rc = mts_instance_dt(insts[i], &dt);
    if (rc) return rc;
    std::unique_ptr<Flight> f(new Flight());
    rc = f->start(insts[i], dt, t0[i], step_count(dt, t_finish[i]));
    if (rc) return rc;
    fl.push_back(std::move(f));
  }
  for (;;) {
    bool pending = false, progressed = false;
    for (auto &f : fl) {
      int rc = f->advance();
      if (rc < 0) return rc;
      progressed |= rc == 1;
      pending |= !f->issued_all();
    }
    if (!pending) break;
    if (!progressed) std::this_thread::sleep_for(std::chrono::microseconds(20));
  }
  for (int i = 0; i < count; i++) {
    int rc = mts_sync(insts[i]);
    if (rc) return rc;
    if (steps) steps[i] = fl[i]->steps;
    if (t_end) t_end[i] = fl[i]->t;
  }
  return MTS_OK;
}

extern "C" int mts_batch_run_graph(mts_instance *const *insts, int count, const double *t0,
                                   const double *t_finish, int *steps, double *t_end) {
  if (count < 0 || (count > 0 && (!insts || !t0 || !t_finish)))
    return mts::set_error(MTS_E_ARG, "null argument or negative count");
  std::vector<int> st(count);
  std::vector<double> te(count);
  for (int i = 0; i < count; i++) {
    const int rc = mts_run_graph(insts[i], t0[i], t_finish[i], &st[i], &te[i]);
    if (rc) return rc;
  }
  for (int i = 0; i < count; i++) {
    const int rc = mts_sync(insts[i]);
    if (rc) return rc;
    if (steps) steps[i] = st[i];
    if (t_end) t_end[i] = te[i];
  }
  return MTS_OK;
}

// The serving loop runs three kinds of threads:
//   setup threads    Mesh + Intersections, MIB tables and stage records on the host;
//   worker threads   create an admitted solve's instance (allocate, upload,
//                    initialize) and finish a completed one (error norms, free);
//   stepping thread  issues one step per round for every solve in flight and
//                    polls their events; it never blocks on the device.
// Admission (arrival order, HBM cap, in-flight limit) runs under the lock
// whenever a planned solve arrives or a finished one frees its memory.
struct mts_server {
  struct Pending {
    long long id;
    mts_request req;
    Clock::time_point submitted;
  };
  // one solve from host setup to its result
  struct Solve {
    long long id = 0;
    mts_request req{};
    mts::PlanPtr plan;  // until the instance exists
    size_t footprint = 0;
    double setup_ms = 0.0;
    int status = MTS_OK;
    char message[160] = {};
    Clock::time_point submitted, admitted;
    Flight f;
    ~Solve() {
      if (f.inst) mts_instance_destroy(f.inst);
    }
  };
  using SolvePtr = std::unique_ptr<Solve>;

  mts_ctx *ctx = nullptr;
  size_t cap = 0;
  int max_inflight = 1;
  std::mutex mu;
  std::condition_variable cv_setup, cv_work, cv_step, cv_done;
  std::deque<Pending> pending;     // submitted, waiting for a setup thread
  std::deque<SolvePtr> ready;      // planned, waiting for admission
  std::deque<SolvePtr> to_make;    // admitted, waiting for a worker to create the instance
  std::deque<SolvePtr> made;       // created, waiting for the stepping thread
  std::deque<SolvePtr> to_finish;  // all steps done (or failed), waiting for a worker
  std::vector<SolvePtr> active;    // stepping thread only
  std::vector<mts_result> done;
  long long next_id = 1;
  long long outstanding = 0;
  int admitted_count = 0;  // admitted and not yet finished
  size_t used = 0;         // device bytes of the admitted solves
  bool stop = false;
  std::vector<std::thread> setup_threads, workers;
  std::thread stepper;

  static void fill_failure(mts_result &r, long long id, const mts_request &q, int code,
                           const char *msg) {
    std::memset(&r, 0, sizeof(r));
    r.id = id;
    r.tag = q.tag;
    r.status = code;
    std::snprintf(r.message, sizeof(r.message), "%s", msg);
  }

  // Caller holds mu.
  void publish_locked(const mts_result &r) {
    done.push_back(r);
    outstanding--;
    cv_done.notify_all();
  }

  // Caller holds mu.  Arrival order: a solve that does not fit waits, and so
  // does everything behind it; one larger than the whole cap is refused.
  void admit_locked() {
    bool any = false;
    while (!ready.empty() && admitted_count < max_inflight) {
      SolvePtr &q = ready.front();
      if (q->footprint > cap) {
        mts_result r;
        fill_failure(r, q->id, q->req, MTS_E_NOMEM, "solve larger than the HBM admission cap");
        publish_locked(r);
        ready.pop_front();
        continue;
      }
      if (used + q->footprint > cap) break;
      used += q->footprint;
      admitted_count++;
      q->admitted = Clock::now();
      to_make.push_back(std::move(q));
      ready.pop_front();
      any = true;
    }
    if (any) cv_work.notify_all();
  }

  // Caller holds mu.  Releases an admitted solve's slot and memory.
  void release_locked(const Solve &q) {
    used -= q.footprint;
    admitted_count--;
    admit_locked();
  }

  void setup_loop() {
    for (;;) {
      Pending p;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_setup.wait(lk, [&] { return stop || !pending.empty(); });
        if (stop) return;
        p = pending.front();
        pending.pop_front();
      }
      const Clock::time_point t0 = Clock::now();
      mts_setup *s = nullptr;
      mts::PlanPtr plan;
      int rc = mts_setup_create(&p.req.grid, &p.req.surface, &p.req.physics, p.req.mib, 2, &s);
      if (!rc) {
        mts::Lanes lanes;
        const bool tuned =
            mts::ctx_tuned_lanes(ctx, p.req.grid.n, p.req.method, p.req.precision, lanes);
        rc = mts::plan_instance(s, &p.req.physics, p.req.method, p.req.dt, p.req.precision, plan,
                                tuned ? &lanes : nullptr);
      }
      mts_setup_destroy(s);  // the plan holds what the device needs
      std::lock_guard<std::mutex> lk(mu);
      if (rc) {
        mts_result r;
        fill_failure(r, p.id, p.req, rc, mts_last_error());
        publish_locked(r);
        continue;
      }
      SolvePtr q(new Solve());
      q->id = p.id;
      q->req = p.req;
      q->footprint = mts::plan_device_bytes(*plan);
      q->plan = std::move(plan);
      q->setup_ms = ms_between(t0, Clock::now());
      q->submitted = p.submitted;
      ready.push_back(std::move(q));
      admit_locked();
    }
  }

  // Instance creation: allocation, upload and the initial field, on the
  // solve's own stream; the stepping thread takes it from there.
  void make(Solve &q) {
    mts_instance *inst = nullptr;
    int rc = mts::instance_from_plan(ctx, std::move(q.plan), &inst);
    if (!rc) q.f.inst = inst;
    if (!rc) rc = mts_initialize(inst, q.req.t0);
    if (!rc) rc = q.f.start(inst, q.req.dt, q.req.t0, step_count(q.req.dt, q.req.t_finish));
    if (!rc) rc = mts_sync(inst);
    if (rc) {
      q.status = rc;
      std::snprintf(q.message, sizeof(q.message), "%s", mts_last_error());
    }
  }

  // Error norms against the exact solution at the final time, then the
  // instance's memory goes back.
  void finish(Solve &q, mts_result &r) {
    std::memset(&r, 0, sizeof(r));
    r.id = q.id;
    r.tag = q.req.tag;
    r.status = q.status;
    if (!r.status && q.f.inst) {
      r.status = mts_error_norms(q.f.inst, q.f.t, &r.linf, &r.l2, r.argmax);
      if (r.status) std::snprintf(q.message, sizeof(q.message), "%s", mts_last_error());
      r.steps = q.f.steps;
      r.t_end = q.f.t;
    }
    std::snprintf(r.message, sizeof(r.message), "%s", q.message);
    const Clock::time_point now = Clock::now();
    r.setup_ms = q.setup_ms;
    r.queue_ms = ms_between(q.submitted, q.admitted);
    r.run_ms = ms_between(q.admitted, now);
    r.latency_ms = ms_between(q.submitted, now);
    r.footprint = q.footprint;
    if (q.f.inst) mts_instance_destroy(q.f.inst);
    q.f.inst = nullptr;
  }

  void worker_loop() {
    for (;;) {
      SolvePtr q;
      bool finishing = false;
      {
        std::unique_lock<std::mutex> lk(mu);
        cv_work.wait(lk, [&] { return stop || !to_finish.empty() || !to_make.empty(); });
        if (stop) return;
        // finishing first: it frees memory for the next admission
        finishing = !to_finish.empty();
        std::deque<SolvePtr> &src = finishing ? to_finish : to_make;
        q = std::move(src.front());
        src.pop_front();
      }
      if (finishing) {
        mts_result r;
        finish(*q, r);
        std::lock_guard<std::mutex> lk(mu);
        release_locked(*q);
        publish_locked(r);
      } else {
        make(*q);
        std::lock_guard<std::mutex> lk(mu);
        if (q->status) {
          to_finish.push_back(std::move(q));
          cv_work.notify_one();
        } else {
          made.push_back(std::move(q));
          cv_step.notify_one();
        }
      }
    }
  }

  void step_loop() {
    for (;;) {
      {
        std::unique_lock<std::mutex> lk(mu);
        if (active.empty()) cv_step.wait(lk, [&] { return stop || !made.empty(); });
        if (stop) return;
        while (!made.empty()) {
          active.push_back(std::move(made.front()));
          made.pop_front();
        }
      }
      // one round: at most one step per solve in flight
      bool progressed = false;
      std::vector<SolvePtr> finished;
      for (size_t i = 0; i < active.size();) {
        Solve &q = *active[i];
        const int rc = q.f.advance();
        progressed |= rc == 1;
        if (rc >= 0 && !q.f.done()) {
          i++;
          continue;
        }
        if (rc < 0) {
          q.status = rc;
          std::snprintf(q.message, sizeof(q.message), "%s", mts_last_error());
        }
        finished.push_back(std::move(active[i]));
        active.erase(active.begin() + i);
      }
      if (!finished.empty()) {
        std::lock_guard<std::mutex> lk(mu);
        for (auto &q : finished) to_finish.push_back(std::move(q));
        cv_work.notify_all();
        progressed = true;
      }
      if (!progressed && !active.empty())
        std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
  }
};

extern "C" {

int mts_server_create(mts_ctx *ctx, size_t hbm_cap_bytes, int max_inflight, int setup_threads,
                      mts_server **out) {
  if (!ctx || !out || max_inflight < 1 || setup_threads < 1 || hbm_cap_bytes == 0)
    return mts::set_error(MTS_E_ARG,
                          "null argument, or a zero cap, in-flight limit or setup thread count");
  std::unique_ptr<mts_server> s(new mts_server());
  s->ctx = ctx;
  s->cap = hbm_cap_bytes;
  s->max_inflight = max_inflight;
  // the setup threads share the setup's thread budget: each setup splits its
  // own work over its share
  const int share = std::max(1, mts::setup_threads() / setup_threads);
  for (int i = 0; i < setup_threads; i++)
    s->setup_threads.emplace_back([p = s.get(), share] {
      mts::set_setup_thread_share(share);
      p->setup_loop();
    });
  for (int i = 0; i < kWorkers; i++) s->workers.emplace_back([p = s.get()] { p->worker_loop(); });
  s->stepper = std::thread([p = s.get()] { p->step_loop(); });
  *out = s.release();
  return MTS_OK;
}

int mts_server_submit(mts_server *s, const mts_request *req, long long *id) {
  if (!s || !req) return mts::set_error(MTS_E_ARG, "null argument");
  {
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->stop) return mts::set_error(MTS_E_STATE, "server is stopping");
    const long long i = s->next_id++;
    s->pending.push_back({i, *req, Clock::now()});
    s->outstanding++;
    if (id) *id = i;
  }
  s->cv_setup.notify_one();
  return MTS_OK;
}

int mts_server_poll(mts_server *s, mts_result *out, int capacity, int *count) {
  if (!s || !count || capacity < 0 || (capacity > 0 && !out))
    return mts::set_error(MTS_E_ARG, "null argument or negative capacity");
  std::lock_guard<std::mutex> lk(s->mu);
  const int n = std::min<int>(capacity, (int)s->done.size());
  for (int i = 0; i < n; i++) out[i] = s->done[i];
  s->done.erase(s->done.begin(), s->done.begin() + n);
  *count = n;
  return MTS_OK;
}

int mts_server_wait(mts_server *s, long long max_outstanding) {
  if (!s) return mts::set_error(MTS_E_ARG, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] { return s->outstanding <= max_outstanding; });
  return MTS_OK;
}

void mts_server_destroy(mts_server *s) {
  if (!s) return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_setup.notify_all();
  s->cv_work.notify_all();
  s->cv_step.notify_all();
  for (auto &t : s->setup_threads) t.join();
  for (auto &t : s->workers) t.join();
  s->stepper.join();
  delete s;  // unfinished solves free their instances and plans here
}

}  // extern "C"
