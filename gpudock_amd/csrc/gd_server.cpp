/*
 * gd_server.cpp -- the serving loop: continuous batching of docking jobs on
 * one engine.
 *
 * gd_server_submit (any thread) copies a job into a host image and queues
 * it.  One worker thread owns the engine and repeats:
 *   1. admit queued jobs in FIFO order while their HBM footprint fits the
 *      arena (and the resident-replica cap): upload + initial energies are
 *      queued on the engine stream behind the cycles already there;
 *   2. queue a tick (cycles_per_tick cycles of every resident job) while
 *      fewer than max_ticks_in_flight ticks are ahead of the GPU;
 *   3. a job whose steps are all queued gets its FP64 best-pose re-score and
 *      its state read-back queued into pinned host memory;
 *   4. a read-back that has landed is parsed, the job retired (its arena
 *      block is freed behind the queued work) and its waiters woken;
 *   5. wait for the oldest tick, a read-back or a submission.
 * The GPU never waits for the host while work is resident: the next tick
 * is already queued when the current one ends.
 *
 * Lanes: a cycle lasts as long as its slowest workgroup task, and a task's
 * time grows with its complex's lna x pnp, so one engine holds every job to
 * the cycle of the largest one resident.  With lanes > 1 each lane is an
 * engine of its own (stream, arena share, pinned read-back area, worker
 * thread running the loop above), jobs are routed by lna x pnp, and the
 * lanes' launches run side by side on the device: a lane of small jobs
 * cycles at its own pace, and its launches fill the CUs that another
 * lane's launch tail leaves idle.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "gd_internal.h"
#include "gd_engine.h"

/* arena bytes kept free beside a job's footprint: task lists, the
 * initial-energy tasks and the FP64 re-score of its read-back */
#define GD_ADMIT_MARGIN (16u << 10)

namespace {

struct SrvJob {
  long long ticket = 0;
  int lane = 0;
  std::unique_ptr<GdImage> img;     /* until admission */
  gd_server_job_info info{};
  int handle = -1;
  bool collecting = false;
  bool cancel_req = false;
  GdCollect col;
  size_t rb_bytes = 0;
  std::vector<char> rb_pageable;    /* read-back larger than the pinned area */
  /* results */
  std::vector<gd_replica_result> res;
  long long ac_mc = 0, ac_x = 0;
};

double pct(std::vector<double> v, double q)
{
  if (v.empty())
    return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double) (v.size() - 1);
  const size_t i = (size_t) pos;
  const double f = pos - (double) i;
  return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

}  // namespace

struct gd_server;

/* One engine and the worker thread that owns it.  Fields marked (mu) are
 * read or written under the server's mutex; the rest belong to the worker. */
struct SrvLane {
  gd_server *srv = nullptr;
  int index = 0;
  gd_engine eng;
  GdArena rb;                       /* pinned read-back area */
  std::thread worker;
  std::vector<SrvJob *> running;
  int admitting = 0;                /* (mu) a job between the queue and the engine */
  int n_pending = 0;                /* (mu) queued jobs routed here */
  int n_resident = 0;               /* (mu) jobs admitted and not yet retired */
  size_t arena_used = 0;            /* (mu) eng.arena.used() after the last change */
  long long stat_cycles = 0, stat_rebuilds = 0, n_completed = 0;   /* (mu) */
  int stat_split = 1;                                               /* (mu) */
  bool split_lane = false;          /* runs several wavefronts per replica */
  struct Batch {
    hipEvent_t done;
    std::vector<SrvJob *> jobs;
  };
  std::vector<Batch> batches;       /* read-backs in flight */

  void loop();
  bool admit_some();
  void start_collects(std::vector<SrvJob *> &ready);
  bool reap_collects();
  void drop_running(SrvJob *j);
};

struct gd_server {
  gd_server_config cfg{};
  std::vector<std::unique_ptr<SrvLane>> lanes;
  std::mutex mu;
  std::condition_variable cv_work, cv_done;
  std::deque<SrvJob *> pending;
  std::unordered_map<long long, SrvJob *> jobs;   /* every job not yet forgotten */
  long long next_ticket = 1;
  bool stop = false;
  int worker_rc = GD_OK;
  std::chrono::steady_clock::time_point t0;
  /* statistics (under mu) */
  long long n_submitted = 0, n_completed = 0, n_failed = 0, n_cancelled = 0;
  int n_running = 0, running_peak = 0;
  std::vector<double> lat, wait;
  double pose_evals = 0.0, first_submit = -1.0, last_done = 0.0;
  size_t arena_peak = 0;
  long long resident_reps = 0;

  double now() const
  {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  }
  void finish(SrvJob *j, int state, int status);
  int first_lane(const gd_job &job) const;
  int admitting() const;
  void note_arena(SrvLane &l);
};

/* The first lane a job may take: the first k < lanes - 1 with lna x pnp <=
 * lane_max_pairs[k], else the last. */
int gd_server::first_lane(const gd_job &job) const
{
  const long long pairs = (long long) job.size.lna * job.size.pnp;
  const int n = (int) lanes.size();
  for (int k = 0; k + 1 < n; ++k)
    if (pairs <= cfg.lane_max_pairs[k])
      return k;
  return n - 1;
}

/* Jobs between the queue and an engine (lock held). */
int gd_server::admitting() const
{
  int n = 0;
  for (auto &l : lanes)
    n += l->admitting;
  return n;
}

/* Record a lane's arena use after a change, and the peak over all lanes
 * (lock held; called by the lane's own worker). */
void gd_server::note_arena(SrvLane &l)
{
  l.arena_used = l.eng.arena.used();
  size_t total = 0;
  for (auto &x : lanes)
    total += x->arena_used;
  arena_peak = std::max(arena_peak, total);
}

/* Mark a job finished (lock held by the caller). */
void gd_server::finish(SrvJob *j, int state, int status)
{
  j->info.state = state;
  j->info.status = status;
  j->info.t_done = now();
  j->info.latency = j->info.t_done - j->info.t_submit;
  if (state == GD_JOB_DONE) {
    ++n_completed;
    ++lanes[j->lane]->n_completed;
    lat.push_back(j->info.latency);
    wait.push_back(j->info.t_admit - j->info.t_submit);
    pose_evals += (double) j->info.n_rep * (double) j->info.steps_total;
    last_done = std::max(last_done, j->info.t_done);
  } else if (state == GD_JOB_CANCELLED) {
    ++n_cancelled;
  } else {
    ++n_failed;
  }
  if (!cfg.keep_results && state == GD_JOB_DONE) {
    jobs.erase(j->ticket);
    delete j;
  }
  cv_done.notify_all();
}

/* Admit this lane's queued jobs in FIFO order while they fit.  Returns true
 * if any. */
bool SrvLane::admit_some()
{
  gd_server &S = *srv;
  bool any = false;
  for (;;) {
    SrvJob *j = nullptr;
    {
      std::lock_guard<std::mutex> lk(S.mu);
      if (!n_pending)
        return any;
      /* the first queued job of this lane; cancelled ones before it end here */
      auto it = S.pending.begin();
      while (it != S.pending.end()) {
        if ((*it)->lane != index) {
          ++it;
          continue;
        }
        if (!(*it)->cancel_req)
          break;
        SrvJob *c = *it;
        it = S.pending.erase(it);
        --n_pending;
        S.finish(c, GD_JOB_CANCELLED, GD_ERR_CANCELLED);
      }
      if (it == S.pending.end())
        return any;
      j = *it;
      if (S.cfg.max_resident_replicas > 0 && !running.empty() &&
          S.resident_reps + j->info.n_rep > S.cfg.max_resident_replicas)
        return any;
      /* footprint + its initial-energy tasks must fit beside the residents */
      if (j->info.footprint + GD_ADMIT_MARGIN > eng.free_bytes() && !running.empty())
        return any;
      S.pending.erase(it);
      --n_pending;
      admitting = 1;
    }
    int h = -1;
    int rc = eng.add_image(j->img, &h);
    std::lock_guard<std::mutex> lk(S.mu);
    admitting = 0;
    if (rc == GD_ERR_NOMEM && !running.empty()) {
      /* fragmented arena: back to the head of the queue (it was this lane's
       * oldest) until a resident job retires */
      S.pending.push_front(j);
      ++n_pending;
      return any;
    }
    if (rc) {
      S.finish(j, GD_JOB_FAILED, rc);
      continue;
    }
    j->handle = h;
    j->info.state = GD_JOB_RUNNING;
    j->info.t_admit = S.now();
    running.push_back(j);
    ++n_resident;
    S.resident_reps += j->info.n_rep;
    ++S.n_running;
    S.running_peak = std::max(S.running_peak, S.n_running);
    S.note_arena(*this);
    any = true;
  }
}

/* Remove a running job from the engine and the running list (lock held). */
void SrvLane::drop_running(SrvJob *j)
{
  (void) eng.remove(j->handle);
  running.erase(std::find(running.begin(), running.end(), j));
  --n_resident;
  srv->resident_reps -= j->info.n_rep;
  --srv->n_running;
  srv->note_arena(*this);
}

/* Queue one read-back (one FP64 re-score launch) for every job in `ready`
 * that finds room in the pinned read-back area; the rest wait for a later
 * round. */
void SrvLane::start_collects(std::vector<SrvJob *> &ready)
{
  std::vector<GdCollect> cs;
  std::vector<SrvJob *> js;
  for (SrvJob *j : ready) {
    const size_t need = eng.collect_bytes(j->handle);
    size_t off = rb.alloc(need);
    if (off == (size_t) -1 && need > rb.size()) {
      j->rb_pageable.resize(need);
      j->col.host = j->rb_pageable.data();
      j->rb_bytes = 0;
    } else if (off == (size_t) -1) {
      continue;   /* read-back area full: retry after others land */
    } else {
      j->col.host = rb.base() + off;
      j->col.host_off = off;
      j->rb_bytes = need;
    }
    j->col.handle = j->handle;
    j->col.host_bytes = need;
    cs.push_back(j->col);
    js.push_back(j);
  }
  if (js.empty())
    return;
  hipEvent_t done = nullptr;
  int rc = eng.collect_async(cs.data(), (int) cs.size(), &done);
  if (rc) {
    if (done)
      (void) hipEventDestroy(done);
    std::lock_guard<std::mutex> lk(srv->mu);
    for (SrvJob *j : js) {
      if (j->rb_bytes)
        rb.free(j->col.host_off, j->rb_bytes);
      drop_running(j);
      srv->finish(j, GD_JOB_FAILED, rc);
    }
    return;
  }
  for (SrvJob *j : js)
    j->collecting = true;
  batches.push_back(Batch{done, js});
}

/* Parse read-backs that have landed; retire their jobs.  Returns true if
 * any job finished. */
bool SrvLane::reap_collects()
{
  bool any = false;
  for (size_t b = 0; b < batches.size();) {
    if (hipEventQuery(batches[b].done) != hipSuccess) {
      ++b;
      continue;
    }
    for (SrvJob *j : batches[b].jobs) {
      GdSlot *s = eng.find(j->handle);
      if (s && srv->cfg.keep_results && !j->cancel_req) {
        j->res.resize(s->img->P.n_rep);
        gd_mclog lg;
        memset(&lg, 0, sizeof lg);
        gd_parse_state(*s->img, s->blk, j->col.host,
                       (const double *) (j->col.host + GdArena::round(s->state_bytes())), &lg,
                       j->res.data());
        j->ac_mc = lg.ac_mc;
        j->ac_x = lg.ac_temp_exchg;
      }
      if (j->rb_bytes)
        rb.free(j->col.host_off, j->rb_bytes);
      j->rb_pageable.clear();
      j->rb_pageable.shrink_to_fit();
      std::lock_guard<std::mutex> lk(srv->mu);
      drop_running(j);
      if (j->cancel_req)
        srv->finish(j, GD_JOB_CANCELLED, GD_ERR_CANCELLED);
      else
        srv->finish(j, GD_JOB_DONE, GD_OK);
      any = true;
    }
    (void) hipEventDestroy(batches[b].done);
    batches.erase(batches.begin() + b);
  }
  return any;
}

void SrvLane::loop()
{
  gd_server &S = *srv;
  if (hipSetDevice(S.cfg.device) != hipSuccess) {
    std::lock_guard<std::mutex> lk(S.mu);
    S.worker_rc = GD_ERR_HIP;
    S.stop = true;
    S.cv_work.notify_all();
    S.cv_done.notify_all();
    return;
  }
  std::deque<hipEvent_t> ticks;
  const int per_tick = S.cfg.cycles_per_tick > 0 ? S.cfg.cycles_per_tick : 1;
  const int max_flight = S.cfg.max_ticks_in_flight > 0 ? S.cfg.max_ticks_in_flight : 2;
  for (;;) {
    {
      std::lock_guard<std::mutex> lk(S.mu);
      if (S.stop)
        break;
    }
    admit_some();

    /* cancellations of running jobs: retire at once (queued work on them
     * finishes harmlessly; their memory is released behind it) */
    for (size_t i = 0; i < running.size();) {
      SrvJob *j = running[i];
      bool c;
      {
        std::lock_guard<std::mutex> lk(S.mu);
        c = j->cancel_req && !j->collecting;
      }
      if (!c) {
        ++i;
        continue;
      }
      std::lock_guard<std::mutex> lk(S.mu);
      drop_running(j);
      S.finish(j, GD_JOB_CANCELLED, GD_ERR_CANCELLED);
    }

    /* queue a tick if any resident job still has steps */
    bool work = false;
    for (SrvJob *j : running) {
      GdSlot *s = eng.find(j->handle);
      if (s && s->steps_done < s->total_steps)
        work = true;
    }
    if (work && (int) ticks.size() < max_flight) {
      int rc = eng.cycles(per_tick);
      {
        std::lock_guard<std::mutex> lk(S.mu);
        stat_cycles = eng.n_cycles;
        stat_rebuilds = eng.n_rebuilds;
        stat_split = eng.split_eff;
      }
      hipEvent_t ev = nullptr;
      if (!rc && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess) {
        if (hipEventRecord(ev, eng.stream) == hipSuccess)
          ticks.push_back(ev);
        else
          (void) hipEventDestroy(ev);
      }
      if (rc) {
        std::lock_guard<std::mutex> lk(S.mu);
        S.worker_rc = rc;
        S.stop = true;
        S.cv_work.notify_all();
        break;
      }
    }
    /* read-backs for jobs whose steps are all queued: one batch */
    {
      std::vector<SrvJob *> ready;
      for (SrvJob *j : running) {
        GdSlot *s = eng.find(j->handle);
        if (s && !j->collecting && s->steps_done >= s->total_steps)
          ready.push_back(j);
      }
      if (!ready.empty())
        start_collects(ready);
    }
    reap_collects();
    eng.reap();
    while (!ticks.empty() && hipEventQuery(ticks.front()) == hipSuccess) {
      (void) hipEventDestroy(ticks.front());
      ticks.pop_front();
    }

    /* wait for something to change */
    bool collecting = false;
    for (SrvJob *j : running)
      collecting |= j->collecting;
    if ((int) ticks.size() >= max_flight) {
      (void) hipEventSynchronize(ticks.front());
    } else if (!work && !batches.empty()) {
      (void) hipEventSynchronize(batches.front().done);
    } else if (!work && !collecting) {
      std::unique_lock<std::mutex> lk(S.mu);
      S.cv_work.wait_for(lk, std::chrono::milliseconds(20),
                         [&] { return S.stop || n_pending > 0; });
    }
  }
  (void) hipStreamSynchronize(eng.stream);
  for (auto e : ticks)
    (void) hipEventDestroy(e);
  for (auto &b : batches)
    (void) hipEventDestroy(b.done);
  batches.clear();
  std::lock_guard<std::mutex> lk(S.mu);
  for (SrvJob *j : running) {
    --S.n_running;
    S.finish(j, GD_JOB_CANCELLED, GD_ERR_CANCELLED);
  }
  running.clear();
  for (auto it = S.pending.begin(); it != S.pending.end();) {
    if ((*it)->lane != index) {
      ++it;
      continue;
    }
    SrvJob *j = *it;
    it = S.pending.erase(it);
    --n_pending;
    S.finish(j, GD_JOB_CANCELLED, GD_ERR_CANCELLED);
  }
  S.cv_done.notify_all();
}

/* ------------------------------------------------------------------------ */
/* C ABI                                                                     */
/* ------------------------------------------------------------------------ */
extern "C" int gd_server_create(const gd_server_config *cfg, gd_server **out)
{
  if (!out)
    return gd_fail(GD_ERR_ARG, "null server out pointer");
  *out = nullptr;
  std::unique_ptr<gd_server> s(new gd_server());
  if (cfg)
    s->cfg = *cfg;
  else
    s->cfg.keep_results = 1;
  const int n_lanes = s->cfg.lanes > 1 ? s->cfg.lanes : 1;
  if (n_lanes > GD_SERVER_MAX_LANES)
    return gd_fail(GD_ERR_ARG, "lanes %d: at most %d", s->cfg.lanes, GD_SERVER_MAX_LANES);
  for (int k = 0; k + 1 < n_lanes; ++k) {
    if (s->cfg.lane_max_pairs[k] <= 0 ||
        (k > 0 && s->cfg.lane_max_pairs[k] <= s->cfg.lane_max_pairs[k - 1]))
      return gd_fail(GD_ERR_ARG, "lane_max_pairs must be positive and increasing");
  }
  const size_t arena = s->cfg.arena_bytes ? s->cfg.arena_bytes : ((size_t) 4 << 30);
  const size_t rb = s->cfg.readback_bytes ? s->cfg.readback_bytes : ((size_t) 64 << 20);
  for (int k = 0; k < n_lanes; ++k) {
    std::unique_ptr<SrvLane> l(new SrvLane());
    l->srv = s.get();
    l->index = k;
    gd_engine_config ec;
    memset(&ec, 0, sizeof ec);
    ec.device = s->cfg.device;
    ec.arena_bytes = arena / (size_t) n_lanes;
    ec.wpg = s->cfg.wpg;
    ec.persistent = s->cfg.persistent;
    ec.use_graph = s->cfg.use_graph;
    ec.kde_policy = s->cfg.kde_policy;
    ec.occ = s->cfg.occ;
    ec.split = s->cfg.split;
    ec.prio = s->cfg.lane_prio[k];
    if (s->cfg.lane_split[k] > 1) {
      if (s->cfg.lane_split[k] != 4 && s->cfg.lane_split[k] != 8 && s->cfg.lane_split[k] != 16)
        return gd_fail(GD_ERR_ARG, "lane_split[%d] = %d: 0, 1, 4, 8 or 16", k,
                       s->cfg.lane_split[k]);
      ec.split = s->cfg.lane_split[k];
      ec.kde_policy = GD_KDE_STAGE_AUTO;   /* split launches read KDE points from LDS */
      l->split_lane = true;
    } else if (s->cfg.lane_split[k] == 1) {
      ec.split = 1;
    }
    int rc = l->eng.init(&ec);
    if (rc)
      return rc;
    GD_HIP_TRY(l->rb.init(rb / (size_t) n_lanes, true));
    s->lanes.push_back(std::move(l));
  }
  s->t0 = std::chrono::steady_clock::now();
  for (auto &l : s->lanes) {
    SrvLane *raw = l.get();
    l->worker = std::thread([raw] { raw->loop(); });
  }
  *out = s.release();
  return GD_OK;
}

extern "C" int gd_server_submit(gd_server *s, const gd_job *job, long long *ticket)
{
  if (!s || !job || !ticket)
    return gd_fail(GD_ERR_ARG, "null argument");
  std::unique_ptr<SrvJob> j(new SrvJob());
  /* the image follows its lane's KDE placement; a lane passes a job on when
   * it holds lane_max_jobs already or the server lane_max_total, and a split
   * lane one it cannot split */
  const int n_lanes = (int) s->lanes.size();
  for (int k = s->first_lane(*job);; ++k) {
    if (k + 1 < n_lanes && (s->cfg.lane_max_jobs[k] > 0 || s->cfg.lane_max_total[k] > 0)) {
      std::lock_guard<std::mutex> lk(s->mu);
      const SrvLane &l = *s->lanes[k];
      if (s->cfg.lane_max_jobs[k] > 0 &&
          l.n_pending + l.n_resident + l.admitting >= s->cfg.lane_max_jobs[k])
        continue;
      if (s->cfg.lane_max_total[k] > 0 &&
          (int) s->pending.size() + s->n_running + s->admitting() >= s->cfg.lane_max_total[k])
        continue;
    }
    j->img.reset(new GdImage());
    int rc = s->lanes[k]->eng.build_image(*j->img, *job);
    if (rc)
      return rc;
    j->lane = k;
    if (k + 1 == n_lanes || !s->lanes[k]->split_lane || s->lanes[k]->eng.can_split(*j->img))
      break;
  }
  j->info.lane = j->lane;
  const gd_engine &eng = s->lanes[j->lane]->eng;
  j->info.footprint = gd_image_footprint(*j->img);
  if (j->info.footprint + GD_ADMIT_MARGIN > eng.arena.size())
    return gd_fail(GD_ERR_NOMEM, "job needs %zu B of HBM; the server arena%s has %zu B",
                   j->info.footprint, s->lanes.size() > 1 ? " lane" : "", eng.arena.size());
  j->info.n_rep = j->img->P.n_rep;
  j->info.steps_total = j->img->P.steps_run;
  j->info.state = GD_JOB_QUEUED;
  std::lock_guard<std::mutex> lk(s->mu);
  if (s->stop)
    return gd_fail(GD_ERR_CANCELLED, "server is stopping");
  j->ticket = s->next_ticket++;
  j->info.ticket = j->ticket;
  j->info.t_submit = s->now();
  if (s->first_submit < 0.0)
    s->first_submit = j->info.t_submit;
  ++s->n_submitted;
  *ticket = j->ticket;
  s->jobs[j->ticket] = j.get();
  ++s->lanes[j->lane]->n_pending;
  s->pending.push_back(j.release());
  s->cv_work.notify_all();
  return GD_OK;
}

static bool finished(int state)
{
  return state == GD_JOB_DONE || state == GD_JOB_FAILED || state == GD_JOB_CANCELLED;
}

extern "C" int gd_server_wait(gd_server *s, long long ticket, double timeout_s,
                              gd_server_job_info *info)
{
  if (!s)
    return gd_fail(GD_ERR_ARG, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  auto it = s->jobs.find(ticket);
  if (it == s->jobs.end())
    return gd_fail(GD_ERR_ARG, "unknown ticket %lld", ticket);
  auto ready = [&] {
    auto i2 = s->jobs.find(ticket);
    return i2 == s->jobs.end() || finished(i2->second->info.state);
  };
  if (timeout_s < 0.0)
    s->cv_done.wait(lk, ready);
  else if (!s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), ready))
    return gd_fail(GD_ERR_TIMEOUT, "job %lld still %s", ticket,
                   s->jobs[ticket]->info.state == GD_JOB_QUEUED ? "queued" : "running");
  it = s->jobs.find(ticket);
  if (it == s->jobs.end())
    return GD_OK;   /* done and forgotten (keep_results = 0) */
  if (info)
    *info = it->second->info;
  return GD_OK;
}

extern "C" int gd_server_fetch(gd_server *s, long long ticket, gd_mclog *mclog,
                               gd_replica_result *results, gd_server_job_info *info)
{
  if (!s)
    return gd_fail(GD_ERR_ARG, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(ticket);
  if (it == s->jobs.end())
    return gd_fail(GD_ERR_ARG, "unknown ticket %lld", ticket);
  SrvJob *j = it->second;
  if (!finished(j->info.state))
    return gd_fail(GD_ERR_TIMEOUT, "job %lld has not finished", ticket);
  if (info)
    *info = j->info;
  const int status = j->info.status;
  if (j->info.state == GD_JOB_DONE) {
    if (results && !j->res.empty())
      memcpy(results, j->res.data(), j->res.size() * sizeof(gd_replica_result));
    if (mclog) {
      mclog->ac_mc += j->ac_mc;
      mclog->ac_temp_exchg += j->ac_x;
      for (size_t r = 0; r < j->res.size(); ++r) {
        if (mclog->acs_mc)
          mclog->acs_mc[r] += j->res[r].acs_mc;
        if (mclog->acs_temp_exchg)
          mclog->acs_temp_exchg[r] += j->res[r].acs_exchg;
      }
    }
  }
  s->jobs.erase(it);
  delete j;
  if (status != GD_OK)
    return gd_fail(status, "job %lld did not complete", ticket);
  return GD_OK;
}

extern "C" int gd_server_cancel(gd_server *s, long long ticket)
{
  if (!s)
    return gd_fail(GD_ERR_ARG, "null server");
  std::lock_guard<std::mutex> lk(s->mu);
  auto it = s->jobs.find(ticket);
  if (it == s->jobs.end())
    return gd_fail(GD_ERR_ARG, "unknown ticket %lld", ticket);
  if (!finished(it->second->info.state))
    it->second->cancel_req = true;
  s->cv_work.notify_all();
  return GD_OK;
}

extern "C" int gd_server_drain(gd_server *s)
{
  if (!s)
    return gd_fail(GD_ERR_ARG, "null server");
  std::unique_lock<std::mutex> lk(s->mu);
  s->cv_done.wait(lk, [&] {
    return s->stop || (s->pending.empty() && s->n_running == 0 && !s->admitting());
  });
  if (s->worker_rc)
    return gd_fail(s->worker_rc, "serving loop stopped on an error");
  return GD_OK;
}

extern "C" int gd_server_get_stats(gd_server *s, gd_server_stats *st)
{
  if (!s || !st)
    return gd_fail(GD_ERR_ARG, "null argument");
  std::lock_guard<std::mutex> lk(s->mu);
  memset(st, 0, sizeof *st);
  st->submitted = s->n_submitted;
  st->completed = s->n_completed;
  st->failed = s->n_failed;
  st->cancelled = s->n_cancelled;
  st->queued = (int) s->pending.size();
  st->running = s->n_running;
  st->p50_latency = pct(s->lat, 0.50);
  st->p99_latency = pct(s->lat, 0.99);
  double sum = 0.0, mx = 0.0;
  for (double v : s->lat) {
    sum += v;
    mx = std::max(mx, v);
  }
  st->mean_latency = s->lat.empty() ? 0.0 : sum / (double) s->lat.size();
  st->max_latency = mx;
  st->p50_wait = pct(s->wait, 0.50);
  st->p99_wait = pct(s->wait, 0.99);
  st->pose_evals = s->pose_evals;
  st->first_submit = s->first_submit < 0.0 ? 0.0 : s->first_submit;
  st->last_done = s->last_done;
  const double span = st->last_done - st->first_submit;
  st->pose_evals_per_s = span > 0.0 ? st->pose_evals / span : 0.0;
  st->arena_peak = s->arena_peak;
  st->running_peak = s->running_peak;
  st->lanes = (int) s->lanes.size();
  for (size_t k = 0; k < s->lanes.size(); ++k) {
    const SrvLane &l = *s->lanes[k];
    st->arena_size += l.eng.arena.size();
    st->cycles += l.stat_cycles;
    st->rebuilds += l.stat_rebuilds;
    st->lane_cycles[k] = l.stat_cycles;
    st->lane_completed[k] = l.n_completed;
    st->lane_split[k] = l.stat_split;
  }
  return GD_OK;
}

extern "C" void gd_server_destroy(gd_server *s)
{
  if (!s)
    return;
  {
    std::lock_guard<std::mutex> lk(s->mu);
    s->stop = true;
  }
  s->cv_work.notify_all();
  for (auto &l : s->lanes)
    if (l->worker.joinable())
      l->worker.join();
  for (auto &kv : s->jobs)
    delete kv.second;
  s->jobs.clear();
  delete s;
}
