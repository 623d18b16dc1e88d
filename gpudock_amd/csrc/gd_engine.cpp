/*
 * gd_engine.cpp -- the resident-complex engine behind every entry point.
 *
 * An engine owns one HIP stream, one device arena and a set of resident
 * complexes.  A cycle is three launches on the stream -- the MC kernel over
 * all tasks of all complexes, the exchange kernel over all temperature
 * chains, the advance kernel -- with each complex's step index kept in
 * device memory, so the cycle can be captured once into a hipGraph and
 * replayed.  Complexes are admitted and retired between cycles without
 * waiting for queued work: their memory comes from the arena (no
 * hipMalloc/hipFree), a new complex is uploaded into a block no queued
 * launch refers to, and a block (or task list, or graph) that queued
 * launches still refer to is released only after an event recorded behind
 * them has completed.
 *
 *   gd_run            = engine with one complex, run to completion
 *                       (the reference's Run(), run.cu:40-352)
 *   gd_session_*      = engine with one complex, driven by the caller
 *   gd_engine_*       = the engine itself (batched complexes)
 *   gd_batch_run      = engine with many complexes, autotuned, graph-run
 *   gd_server_*       = engine fed by a request loop (gd_server.cpp)
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <vector>

#include "gd_internal.h"
#include "gd_engine.h"

namespace {
template <class T> size_t vbytes(const std::vector<T> &v) { return v.size() * sizeof(T); }
}  // namespace

/* ------------------------------------------------------------------------ */
/* engine                                                                    */
/* ------------------------------------------------------------------------ */
int gd_engine::init(const gd_engine_config *cfg)
{
  gd_engine_config c;
  memset(&c, 0, sizeof c);
  if (cfg)
    c = *cfg;
  if (c.arena_bytes == 0)
    c.arena_bytes = (size_t) 1 << 30;
  const int ndev = gd_device_count();
  if (ndev < 1)
    return gd_fail(GD_ERR_NODEV, "no HIP device visible");
  if (c.device < 0 || c.device >= ndev)
    return gd_fail(GD_ERR_ARG, "device %d of %d", c.device, ndev);
  conf = c;
  lc.wpg = (c.wpg == 4 || c.wpg == 16) ? c.wpg : 8;
  lc.persistent = c.persistent ? 1 : 0;
  lc.occ = (c.occ == 6 || c.occ == 8) ? c.occ : 0;
  lc.split = (c.split == 1 || c.split == 4 || c.split == 8 || c.split == 16) ? c.split : 0;
  lc.prio = c.prio > 0 ? std::min(c.prio, 3) : 0;
  lc.stage_bytes = 16;
  GD_HIP_TRY(hipSetDevice(c.device));
  GD_HIP_TRY(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                   c.device));
  GD_HIP_TRY(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, c.device));
  GD_HIP_TRY(arena.init(c.arena_bytes));
  GD_HIP_TRY(pinned.init(c.staging_bytes ? c.staging_bytes : ((size_t) 64 << 20), true));
  GD_HIP_TRY(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
  /* control block: the persistent work-queue counter */
  ctl_off = arena.alloc(256, true);
  if (ctl_off == (size_t) -1)
    return gd_fail(GD_ERR_NOMEM, "arena too small");
  std::vector<char> zeros(256, 0);
  GD_HIP_TRY(hipMemcpy(arena.base() + ctl_off, zeros.data(), 256, hipMemcpyHostToDevice));
  return GD_OK;
}

gd_engine::~gd_engine()
{
  (void) hipSetDevice(conf.device);
  if (stream)
    (void) hipStreamSynchronize(stream);
  drop_graph();
  for (auto &d : deferred) {
    if (d.gexec) (void) hipGraphExecDestroy(d.gexec);
    if (d.graph) (void) hipGraphDestroy(d.graph);
    if (d.ev) (void) hipEventDestroy(d.ev);
  }
  deferred.clear();
  for (auto &e : ev_mc) (void) hipEventDestroy(e);
  for (auto &e : ev_x) (void) hipEventDestroy(e);
  for (auto *s : slots)
    delete s;
  if (stream)
    (void) hipStreamDestroy(stream);
}

int *gd_engine::queue_ptr() { return (int *) (arena.base() + ctl_off); }
/* the exchange launch's finished-workgroup counter: the last int of the
 * control block (the queue counters, one per class, come first) */
int *gd_engine::done_ptr() { return (int *) (arena.base() + ctl_off) + 63; }

void gd_engine::defer(size_t off, size_t bytes, hipGraph_t g, hipGraphExec_t x, bool host)
{
  Deferred d{nullptr, off, bytes, g, x, host};
  if (hipEventCreateWithFlags(&d.ev, hipEventDisableTiming) != hipSuccess ||
      hipEventRecord(d.ev, stream) != hipSuccess) {
    /* no event: release once the stream has drained */
    if (d.ev)
      (void) hipEventDestroy(d.ev);
    (void) hipStreamSynchronize(stream);
    if (bytes)
      (host ? pinned : arena).free(off, bytes);
    if (x) (void) hipGraphExecDestroy(x);
    if (g) (void) hipGraphDestroy(g);
    return;
  }
  deferred.push_back(d);
}

void gd_engine::reap()
{
  size_t k = 0;
  for (size_t i = 0; i < deferred.size(); ++i) {
    Deferred &d = deferred[i];
    if (hipEventQuery(d.ev) == hipSuccess) {
      if (d.bytes)
        (d.host ? pinned : arena).free(d.off, d.bytes);
      if (d.gexec) (void) hipGraphExecDestroy(d.gexec);
      if (d.graph) (void) hipGraphDestroy(d.graph);
      (void) hipEventDestroy(d.ev);
    } else {
      deferred[k++] = d;
    }
  }
  deferred.resize(k);
}

void gd_engine::drop_graph()
{
  if (gexec || graph)
    defer(0, 0, graph, gexec);
  gexec = nullptr;
  graph = nullptr;
}

/* Arena block of `bytes` (complex blocks low, small task lists and scratch
 * high); when the arena is full, wait for queued work and release what it
 * held before giving up. */
static size_t arena_take(gd_engine &E, size_t bytes, bool high = false)
{
  size_t off = E.arena.alloc(bytes, high);
  if (off == (size_t) -1 && !E.deferred.empty()) {
    (void) hipStreamSynchronize(E.stream);
    E.reap();
    off = E.arena.alloc(bytes, high);
  }
  return off;
}

int gd_engine::upload(char *dev, const void *src, size_t bytes)
{
  if (!bytes)
    return GD_OK;
  size_t off = pinned.alloc(bytes);
  if (off == (size_t) -1 && !deferred.empty()) {
    reap();
    off = pinned.alloc(bytes);
  }
  if (off != (size_t) -1) {
    memcpy(pinned.base() + off, src, bytes);
    hipError_t e = hipMemcpyAsync(dev, pinned.base() + off, bytes, hipMemcpyHostToDevice,
                                  stream);
    defer(off, bytes, nullptr, nullptr, true);
    GD_HIP_TRY(e);
    return GD_OK;
  }
  /* larger than the staging area: pageable copy, then wait for it */
  GD_HIP_TRY(hipMemcpyAsync(dev, src, bytes, hipMemcpyHostToDevice, stream));
  GD_HIP_TRY(hipStreamSynchronize(stream));
  return GD_OK;
}

int gd_engine::upload_image(GdImage &I, size_t off, GdBlock &b)
{
  const size_t bytes = gd_image_footprint(I);
  char *dev = arena.base() + off;
  size_t hoff = pinned.alloc(bytes);
  if (hoff == (size_t) -1 && !deferred.empty()) {
    reap();
    hoff = pinned.alloc(bytes);
  }
  if (hoff != (size_t) -1) {
    gd_pack_image(I, pinned.base() + hoff, dev, b);
    hipError_t e = hipMemcpyAsync(dev, pinned.base() + hoff, bytes, hipMemcpyHostToDevice,
                                  stream);
    defer(hoff, bytes, nullptr, nullptr, true);
    GD_HIP_TRY(e);
    return GD_OK;
  }
  std::vector<char> host(bytes);
  gd_pack_image(I, host.data(), dev, b);
  GD_HIP_TRY(hipMemcpyAsync(dev, host.data(), bytes, hipMemcpyHostToDevice, stream));
  GD_HIP_TRY(hipStreamSynchronize(stream));
  return GD_OK;
}

/* Wavefronts per replica.  One, unless the resident replicas are too few to
 * give every SIMD of the device two wavefronts: then 16, 8 or 4 per replica
 * (the most that keeps the total within that), one replica per workgroup.
 * Complexes with more position restraints than the split area holds keep
 * one wavefront per replica. */
/* Split launches stage every table, the pair list, the restraint targets
 * and the PSP rows in LDS beside the area: restraints beyond the area's
 * slots, a complex whose extras do not fit, or KDE points kept in HBM mean
 * one wavefront per replica. */
bool gd_engine::can_split(const GdImage &I) const
{
  const gd_pack &P = I.P;
  return P.pos <= GD_SPLIT_MAX_POS && P.split_extra && !kde_global &&
         P.split_lds_bytes + GD_WSLOT_FLOATS * 4 + (int) GD_SPLIT_AREA_BYTES <= max_lds;
}

int gd_engine::choose_split() const
{
  long long reps = 0;
  for (auto *s : slots) {
    if (!can_split(*s->img))
      return 1;
    reps += s->img->P.n_rep;
  }
  if (lc.split == 1 || lc.split == 4 || lc.split == 8 || lc.split == 16)
    return lc.split;
  const long long waves = 2LL * 4 * n_cu;
  for (int s : {16, 8, 4})
    if (reps > 0 && reps * s <= waves)
      return s;
  return 1;
}

/* New device copies of the pack, task and chain arrays for the current
 * members (fresh blocks: queued launches keep reading the old ones, which
 * are released behind them). */
int gd_engine::rebuild()
{
  drop_graph();
  split_eff = choose_split();
  gd_launch_cfg lce = lc;
  lce.split = split_eff;
  std::vector<gd_pack> packs;
  std::vector<gd_chain> chains;
  /* occupancy class of every complex: workgroups per CU at its LDS need */
  std::map<int, int> occ_cache;
  auto per_cu_of = [&](int stage) -> int {
    auto it = occ_cache.find(stage);
    if (it != occ_cache.end())
      return it->second;
    int k = 0;
    if (gd_mc_occupancy(lce, stage, &k) != hipSuccess || k < 1)
      k = 1;
    occ_cache[stage] = k;
    return k;
  };
  struct Pending {
    std::vector<gd_task> tasks;
    std::vector<double> cost;
    int stage = 16;
  };
  std::map<int, Pending> by_class;   /* key: workgroups per CU */
  std::vector<gd_task> t1;
  for (auto *s : slots) {
    const int cx = (int) packs.size();
    s->cx_index = cx;
    gd_pack p = s->img->P;
    if (kde_global && p.kde_in_lds) {
      p.kde_in_lds = 0;
      p.lds_bytes = p.off_kde;
      p.split_extra = 0;
      p.split_lds_bytes = p.lds_bytes;
    }
    packs.push_back(p);
    const int stage = std::max(16, split_eff > 1 ? std::max(p.lds_bytes, p.split_lds_bytes)
                                                 : p.lds_bytes);
    Pending &cl = by_class[per_cu_of(stage)];
    cl.stage = std::max(cl.stage, stage);
    t1.clear();
    gd_complex_tasks(*s->img, cx, split_eff > 1 ? 1 : lc.wpg, t1);
    for (const gd_task &t : t1) {
      cl.tasks.push_back(t);
      cl.cost.push_back((double) t.r_count * p.lna * p.pnp_pad);
    }
    const int nc = p.n_prt * p.n_lig;
    for (int c = 0; c < nc; ++c)
      chains.push_back(gd_chain{cx, c});
  }
  /* per class, longest tasks first: the queue hands out big tasks before
   * small ones; classes with the most work launch first */
  std::vector<gd_task> sorted;
  classes.clear();
  std::vector<std::pair<double, int>> order_cls;
  for (auto &kv : by_class) {
    double w = 0.0;
    for (double c : kv.second.cost)
      w += c;
    order_cls.push_back({-w, kv.first});
  }
  std::sort(order_cls.begin(), order_cls.end());
  int max_stage = 16;
  for (auto &oc : order_cls) {
    Pending &cl = by_class[oc.second];
    std::vector<int> order(cl.tasks.size());
    for (size_t i = 0; i < order.size(); ++i)
      order[i] = (int) i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return cl.cost[a] > cl.cost[b]; });
    Class c;
    c.stage = cl.stage;
    c.off = (int) sorted.size();
    c.n = (int) order.size();
    c.per_cu = oc.second;
    c.grid = lc.persistent ? std::min(c.n, c.per_cu * n_cu) : c.n;
    for (int i : order)
      sorted.push_back(cl.tasks[i]);
    classes.push_back(c);
    if (classes.size() > 62)
      return gd_fail(GD_ERR_ARG, "more than 62 occupancy classes");
    max_stage = std::max(max_stage, c.stage);
  }

  if (packs_bytes) defer(packs_off, packs_bytes);
  if (tasks_bytes) defer(tasks_off, tasks_bytes);
  if (chains_bytes) defer(chains_off, chains_bytes);
  packs_bytes = tasks_bytes = chains_bytes = 0;
  packs_off = tasks_off = chains_off = 0;
  auto place = [&](size_t &off, size_t &cap, const void *src, size_t bytes) -> int {
    if (!bytes)
      return GD_OK;
    const size_t want = GdArena::round(bytes);
    off = arena_take(*this, want, true);
    if (off == (size_t) -1)
      return gd_fail(GD_ERR_NOMEM, "arena too small for the task lists");
    cap = want;
    return upload(arena.base() + off, src, bytes);
  };
  int rc = place(packs_off, packs_bytes, packs.data(), vbytes(packs));
  if (!rc) rc = place(tasks_off, tasks_bytes, sorted.data(), vbytes(sorted));
  if (!rc) rc = place(chains_off, chains_bytes, chains.data(), vbytes(chains));
  if (rc)
    return rc;
  n_cx = (int) packs.size();
  n_tasks = (int) sorted.size();
  n_chain = (int) chains.size();
  /* initial energies of newly admitted complexes (MonteCarlo_Init_d,
   * montecarlo_d.cu:12-55): one grid launch over their tasks, queued before
   * the first cycle that includes them */
  std::vector<gd_task> init_tasks;
  int init_stage = 16;
  for (int h : pending_init) {
    GdSlot *s = find(h);
    if (!s)
      continue;
    gd_complex_tasks(*s->img, s->cx_index, lc.wpg, init_tasks);
    init_stage = std::max(init_stage, packs[s->cx_index].lds_bytes);
  }
  pending_init.clear();
  lc.stage_bytes = max_stage;
  GD_HIP_TRY(gd_set_kernel_attrs(std::max(max_stage, init_stage)));
  if (!init_tasks.empty()) {
    const size_t tb = GdArena::round(vbytes(init_tasks));
    const size_t toff = arena_take(*this, tb, true);
    if (toff == (size_t) -1)
      return gd_fail(GD_ERR_NOMEM, "arena too small for the initial-energy tasks");
    rc = upload(arena.base() + toff, init_tasks.data(), vbytes(init_tasks));
    gd_launch_cfg ic = lc;
    ic.persistent = 0;
    ic.split = 1;
    ic.stage_bytes = init_stage;
    hipError_t he = rc ? hipSuccess
                       : gd_launch_mc(ic, (const gd_pack *) (arena.base() + packs_off),
                                      (const gd_task *) (arena.base() + toff),
                                      (int) init_tasks.size(), nullptr,
                                      (int) init_tasks.size(), 1, stream);
    defer(toff, tb);
    if (rc)
      return rc;
    GD_HIP_TRY(he);
  }
  dirty = false;
  ++n_rebuilds;
  return GD_OK;
}

int gd_engine::ensure_built()
{
  if (!dirty)
    return GD_OK;
  GD_HIP_TRY(hipSetDevice(conf.device));
  return rebuild();
}

int gd_engine::grid() const
{
  int g = 0;
  for (const Class &c : classes)
    g += c.grid;
  return g;
}

int gd_engine::resident() const
{
  int r = 0;
  for (const Class &c : classes)
    r += c.per_cu * n_cu;
  return r;
}

/* MC launches of one cycle, one per occupancy class */
static hipError_t launch_classes(gd_engine &E, int *queue_base)
{
  const gd_pack *packs = (const gd_pack *) (E.arena.base() + E.packs_off);
  const gd_task *tasks = (const gd_task *) (E.arena.base() + E.tasks_off);
  for (size_t i = 0; i < E.classes.size(); ++i) {
    const gd_engine::Class &c = E.classes[i];
    gd_launch_cfg C = E.lc;
    C.split = E.split_eff;
    C.stage_bytes = c.stage;
    hipError_t e = gd_launch_mc(C, packs, tasks + c.off, c.n,
                                E.lc.persistent ? queue_base + i : nullptr, c.grid, 0,
                                E.stream);
    if (e != hipSuccess)
      return e;
  }
  return hipSuccess;
}

/* the exchange and advance of a cycle */
static hipError_t launch_exchange(gd_engine &E)
{
  const gd_pack *packs = (const gd_pack *) (E.arena.base() + E.packs_off);
  const gd_chain *chains = (const gd_chain *) (E.arena.base() + E.chains_off);
  return gd_launch_exchange(packs, chains, E.n_chain, E.n_cx, E.queue_ptr(),
                            (int) E.classes.size(), E.done_ptr(), E.stream);
}

int gd_engine::build_image(GdImage &I, const gd_job &job) const
{
  gd_options o;
  if (job.opt) {
    o = *job.opt;
  } else {
    memset(&o, 0, sizeof o);
    o.seed = 1;
    o.exchange = 1;
  }
  const gd_complex_size &sz = job.size;
  int rc = gd_check_inputs(job.lig, job.prt, job.psp, job.kde, job.mcs, job.enepara, sz);
  if (rc)
    return rc;
  if (!job.temp || !job.mcpara)
    return gd_fail(GD_ERR_ARG, "null temp/mcpara");
  if (sz.n_tmp < 1 || sz.n_tmp > 64)
    return gd_fail(GD_ERR_ARG, "n_tmp=%d outside [1, 64]", sz.n_tmp);
  if ((long long) sz.n_rep != (long long) sz.n_lig * sz.n_prt * sz.n_tmp)
    return gd_fail(GD_ERR_ARG, "n_rep=%d != n_lig*n_prt*n_tmp", sz.n_rep);
  const gd_mcpara *mc = job.mcpara;
  if (mc->steps_total < 0 || mc->steps_per_exchange < 1)
    return gd_fail(GD_ERR_ARG, "steps_total=%d steps_per_exchange=%d", mc->steps_total,
                   mc->steps_per_exchange);
  return gd_build_image(I, job.lig, job.prt, job.psp, job.kde, job.mcs, job.enepara, job.temp,
                        job.replica, mc, sz, o, max_lds, conf.kde_policy);
}

int gd_engine::add(const gd_ligand *lig, const gd_protein *prt, const gd_psp *psp,
                   const gd_kde *kde, const gd_mcs *mcs, const gd_enepara *ep,
                   const gd_temp *temp, const gd_replica *replica, const gd_mcpara *mc,
                   const gd_complex_size &sz, const gd_options &o, int *handle)
{
  gd_job job;
  memset(&job, 0, sizeof job);
  job.lig = lig; job.prt = prt; job.psp = psp; job.kde = kde; job.mcs = mcs;
  job.enepara = ep; job.temp = temp; job.replica = replica; job.mcpara = mc;
  job.size = sz; job.opt = &o;
  std::unique_ptr<GdImage> img(new GdImage());
  int rc = build_image(*img, job);
  if (rc)
    return rc;
  return add_image(img, handle);
}

int gd_engine::add_image(std::unique_ptr<GdImage> &img, int *handle)
{
  GD_HIP_TRY(hipSetDevice(conf.device));
  if (!img)
    return gd_fail(GD_ERR_ARG, "no image");
  const size_t bytes = gd_image_footprint(*img);
  size_t off = arena_take(*this, bytes);
  if (off == (size_t) -1 && slots.empty() && (packs_bytes || tasks_bytes || chains_bytes)) {
    /* nothing resident: the task lists of retired complexes go first */
    rebuild();
    (void) hipStreamSynchronize(stream);
    reap();
    off = arena_take(*this, bytes);
  }
  if (off == (size_t) -1)
    return gd_fail(GD_ERR_NOMEM, "complex needs %zu B; arena has %zu of %zu B in use", bytes,
                   arena.used(), arena.size());
  std::unique_ptr<GdSlot> s(new GdSlot());
  s->img = std::move(img);
  GdImage &I = *s->img;
  int rc = upload_image(I, off, s->blk);
  if (rc) {
    arena.free(off, bytes);
    return rc;
  }
  s->off = off;
  s->bytes = bytes;
  s->pack_dev = (gd_pack *) (arena.base() + off + s->blk.pack);
  s->handle = next_handle++;
  s->total_steps = I.P.steps_run;
  s->spe = I.P.spe;
  /* initial energies: queued with the next task-list build, one launch for
   * every complex admitted since the last one */
  pending_init.push_back(s->handle);
  *handle = s->handle;
  slots.push_back(s.release());
  dirty = true;
  return GD_OK;
}

GdSlot *gd_engine::find(int handle)
{
  for (auto *s : slots)
    if (s->handle == handle)
      return s;
  return nullptr;
}

int gd_engine::remove(int handle)
{
  for (size_t i = 0; i < slots.size(); ++i) {
    if (slots[i]->handle == handle) {
      GD_HIP_TRY(hipSetDevice(conf.device));
      pending_init.erase(std::remove(pending_init.begin(), pending_init.end(), handle),
                         pending_init.end());
      defer(slots[i]->off, slots[i]->bytes);
      delete slots[i];
      slots.erase(slots.begin() + i);
      dirty = true;
      return GD_OK;
    }
  }
  return gd_fail(GD_ERR_ARG, "no complex with handle %d", handle);
}

int gd_engine::enqueue_cycle()
{
  GD_HIP_TRY(launch_classes(*this, queue_ptr()));
  GD_HIP_TRY(launch_exchange(*this));
  return GD_OK;
}

int gd_engine::cycles(int n)
{
  if (n < 0)
    return gd_fail(GD_ERR_ARG, "negative cycle count");
  int rc = ensure_built();
  if (rc)
    return rc;
  if (n_tasks == 0)
    return GD_OK;
  GD_HIP_TRY(hipSetDevice(conf.device));
  for (int c = 0; c < n; ++c) {
    if (timing) {
      hipEvent_t a, b, x0, x1;
      GD_HIP_TRY(hipEventCreate(&a));
      GD_HIP_TRY(hipEventCreate(&b));
      GD_HIP_TRY(hipEventCreate(&x0));
      GD_HIP_TRY(hipEventCreate(&x1));
      GD_HIP_TRY(hipEventRecord(a, stream));
      GD_HIP_TRY(launch_classes(*this, queue_ptr()));
      GD_HIP_TRY(hipEventRecord(b, stream));
      GD_HIP_TRY(hipEventRecord(x0, stream));
      GD_HIP_TRY(launch_exchange(*this));
      GD_HIP_TRY(hipEventRecord(x1, stream));
      ev_mc.push_back(a);
      ev_mc.push_back(b);
      ev_x.push_back(x0);
      ev_x.push_back(x1);
    } else if (conf.use_graph) {
      if (!gexec) {
        GD_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
        int rc2 = enqueue_cycle();
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(stream, &g);
        if (rc2) {
          if (g) (void) hipGraphDestroy(g);
          return rc2;
        }
        GD_HIP_TRY(e);
        graph = g;
        GD_HIP_TRY(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
        ++n_captures;
      }
      GD_HIP_TRY(hipGraphLaunch(gexec, stream));
    } else {
      int rc2 = enqueue_cycle();
      if (rc2)
        return rc2;
    }
    ++n_cycles;
    for (auto *s : slots) {
      const int left = s->total_steps - s->steps_done;
      const int k = left < s->spe ? left : s->spe;
      if (k > 0) {
        s->steps_done += k;
        s->cycles_done += 1;
      }
    }
  }
  return GD_OK;
}

/* One MC launch at the current step of every complex with the current
 * launch shape, then the queue re-armed; no step counter moves. */
int gd_engine::trial_cycle()
{
  const gd_pack *packs = (const gd_pack *) (arena.base() + packs_off);
  GD_HIP_TRY(launch_classes(*this, queue_ptr()));
  GD_HIP_TRY(gd_launch_advance(packs, 0, queue_ptr(), (int) classes.size(), stream));
  return GD_OK;
}

/* Times one MC launch for each launch shape (waves per group 4/8/16, grid
 * or persistent form) on the resident complexes and keeps the fastest.
 * The replica state is saved before and restored after, so tuning leaves
 * no trace in the results. */
int gd_engine::autotune(gd_tune_report *rep)
{
  GD_HIP_TRY(hipSetDevice(conf.device));
  int rc = ensure_built();
  if (rc)
    return rc;
  if (rep)
    memset(rep, 0, sizeof *rep);
  if (n_tasks == 0)
    return GD_OK;
  size_t snap_bytes = 0;
  for (auto *s : slots)
    snap_bytes += GdArena::round(s->state_bytes());
  const size_t snap = arena_take(*this, snap_bytes, true);
  if (snap == (size_t) -1)
    return gd_fail(GD_ERR_NOMEM, "arena too small to save the replica state for tuning");
  auto copy_state = [&](bool save) -> hipError_t {
    size_t o = snap;
    for (auto *s : slots) {
      char *st = arena.base() + s->off + s->blk.mv;
      char *sv = arena.base() + o;
      hipError_t e = hipMemcpyAsync(save ? sv : st, save ? st : sv, s->state_bytes(),
                                    hipMemcpyDeviceToDevice, stream);
      if (e != hipSuccess)
        return e;
      o += GdArena::round(s->state_bytes());
    }
    return hipSuccess;
  };
  hipError_t he = copy_state(true);
  bool any_kde_lds = false;
  for (auto *s : slots)
    any_kde_lds |= s->img->P.kde_in_lds != 0;
  float best = 3.0e38f;
  gd_launch_cfg best_c = lc;
  int best_kg = kde_global;
  int n = 0;
  hipEvent_t a = nullptr, b = nullptr;
  if (he == hipSuccess)
    he = hipEventCreate(&a);
  if (he == hipSuccess)
    he = hipEventCreate(&b);
  struct Cand {
    float ms;
    int wpg, pers, kg, occ, split;
  };
  std::vector<Cand> cands;
  /* the split the resident set would get by itself: when it is above one,
   * the replicas are few and the split launches are candidates too */
  const int split_auto = [&] {
    const int keep = lc.split;
    lc.split = 0;
    const int v = choose_split();
    lc.split = keep;
    return v;
  }();
  /* time `reps` cycles of a configuration after a warm-up cycle; returns
   * the mean cycle time in ms, or a negative value on error */
  auto time_cfg = [&](int wpg, int pers, int kg, int occ, int split, int reps) -> float {
    if (rc || he != hipSuccess)
      return -1.0f;
    kde_global = kg;
    lc.wpg = wpg;
    lc.persistent = pers;
    lc.occ = occ;
    lc.split = split;
    rc = rebuild();
    if (!rc)
      rc = trial_cycle();
    if (!rc)
      he = hipEventRecord(a, stream);
    for (int i = 0; i < reps && !rc && he == hipSuccess; ++i)
      rc = trial_cycle();
    if (!rc && he == hipSuccess)
      he = hipEventRecord(b, stream);
    if (!rc && he == hipSuccess)
      he = hipEventSynchronize(b);
    float ms = 0.0f;
    if (!rc && he == hipSuccess)
      he = hipEventElapsedTime(&ms, a, b);
    if (rc || he != hipSuccess)
      return -1.0f;
    reap();
    return ms / (float) reps;
  };
  /* one trial of the staged search */
  auto trial = [&](int wpg, int pers, int kg, int occ, int split) {
    const float ms = time_cfg(wpg, pers, kg, occ, split, 2);
    if (ms < 0.0f)
      return;
    if (rep && n < GD_TUNE_MAX) {
      rep->wpg[n] = wpg;
      rep->persistent[n] = pers;
      rep->kde_global[n] = kg;
      rep->occ[n] = occ;
      rep->split[n] = split_eff;
      rep->ms[n] = ms;
    }
    ++n;
    cands.push_back({ms, wpg, pers, kg, occ, split});
    if (ms < best) {
      best = ms;
      best_c = lc;
      best_kg = kg;
    }
  };
  /* staged search: launch shape (and, for few replicas, wavefronts per
   * replica); then KDE placement x register budget; then every launch shape
   * and register budget at the chosen placement */
  const int wpgs[3] = {4, 8, 16};
  const int kg0 = kde_global, occ0 = lc.occ;
  for (int wi = 0; wi < 3; ++wi)
    for (int pers = 0; pers < 2; ++pers)
      trial(wpgs[wi], pers, kg0, occ0, 1);
  if (split_auto > 1)
    for (int sp : {4, 8, 16})
      for (int pers = 0; pers < 2; ++pers)
        trial(lc.wpg, pers, kg0, occ0, sp);
  {
    /* KDE placement x register budget (one wavefront per replica) at the
     * best launch shape */
    const gd_launch_cfg c = best_c;
    for (int kg = 0; kg < (any_kde_lds ? 2 : 1); ++kg)
      for (int occ : {0, 6, 8})
        if ((kg != kg0 || occ != occ0) && (c.split <= 1 || occ == occ0))
          trial(c.wpg, c.persistent, any_kde_lds ? kg : kg0, occ, c.split);
  }
  if (best_c.split <= 1) {
    /* the best shape of the first stage may not be the best at another
     * register budget: all of them again (a shape far slower in the first
     * stage, as 4 waves/group on LDS-bound complexes, is left out) */
    const int kg = best_kg;
    auto tried = [&](int wpg, int pers, int occ) {
      for (const Cand &c : cands)
        if (c.wpg == wpg && c.pers == pers && c.kg == kg && c.occ == occ && c.split <= 1)
          return true;
      return false;
    };
    float first_best[3] = {3.0e38f, 3.0e38f, 3.0e38f};
    for (const Cand &c : cands)
      for (int wi = 0; wi < 3; ++wi)
        if (c.wpg == wpgs[wi] && c.split <= 1 && c.ms < first_best[wi])
          first_best[wi] = c.ms;
    for (int wi = 0; wi < 3; ++wi) {
      if (first_best[wi] > 1.25f * best)
        continue;
      for (int pers = 0; pers < 2; ++pers)
        for (int occ : {0, 6, 8})
          if (!tried(wpgs[wi], pers, occ))
            trial(wpgs[wi], pers, kg, occ, 1);
    }
  }
  /* the staged trials differ by a few per cent, about what one short trial
   * can resolve: the three fastest distinct configurations are timed again
   * over more cycles, in two interleaved passes (A B C A B C, so a drift of
   * the clocks does not favour one of them), and the one with the fastest
   * pass is kept */
  if (!rc && he == hipSuccess && cands.size() > 1) {
    std::stable_sort(cands.begin(), cands.end(),
                     [](const Cand &x, const Cand &y) { return x.ms < y.ms; });
    std::vector<Cand> top;
    for (const Cand &c : cands) {
      bool dup = false;
      for (const Cand &t : top)
        dup |= t.wpg == c.wpg && t.pers == c.pers && t.kg == c.kg && t.occ == c.occ &&
               t.split == c.split;
      if (!dup)
        top.push_back(c);
      if (top.size() == 3)
        break;
    }
    /* about 100 ms of cycles per configuration and pass (4 to 16 cycles):
     * the shapes left differ by a few per cent, and shorter windows read
     * the device's clock changes as well */
    const int reps = std::max(4, std::min(16, (int) (100.0f / std::max(top[0].ms, 1e-3f))));
    std::vector<float> fastest(top.size(), 3.0e38f);
    for (int pass = 0; pass < 2; ++pass)
      for (size_t i = 0; i < top.size(); ++i) {
        const Cand &c = top[i];
        const float ms = time_cfg(c.wpg, c.pers, c.kg, c.occ, c.split, reps);
        if (ms >= 0.0f && ms < fastest[i])
          fastest[i] = ms;
      }
    size_t pick = 0;
    for (size_t i = 1; i < top.size(); ++i)
      if (fastest[i] < fastest[pick])
        pick = i;
    if (fastest[pick] < 3.0e38f) {
      const Cand &c = top[pick];
      best_c = lc;   /* the fields time_cfg sets; rebuild() below applies them */
      best_c.wpg = c.wpg;
      best_c.persistent = c.pers;
      best_c.occ = c.occ;
      best_c.split = c.split;
      best_kg = c.kg;
    }
  }
  if (a) (void) hipEventDestroy(a);
  if (b) (void) hipEventDestroy(b);
  lc = best_c;
  kde_global = best_kg;
  if (he == hipSuccess)
    he = copy_state(false);
  defer(snap, snap_bytes);
  if (rc)
    return rc;
  if (he != hipSuccess)
    return gd_fail(GD_ERR_HIP, "autotune: %s", hipGetErrorString(he));
  if (rep) {
    rep->n = n;
    rep->n = std::min(n, GD_TUNE_MAX);
    rep->chosen_wpg = lc.wpg;
    rep->chosen_persistent = lc.persistent;
    rep->chosen_kde_global = kde_global;
    rep->chosen_occ = lc.occ;
    rep->chosen_split = lc.split;
  }
  return rebuild();
}

int gd_engine::sync()
{
  GD_HIP_TRY(hipSetDevice(conf.device));
  GD_HIP_TRY(hipStreamSynchronize(stream));
  drain_timing();
  reap();
  return GD_OK;
}

void gd_engine::drain_timing()
{
  for (size_t i = 0; i + 1 < ev_mc.size(); i += 2) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ev_mc[i], ev_mc[i + 1]) == hipSuccess) {
      mc_ms += ms;
      mc_launches++;
    }
    (void) hipEventDestroy(ev_mc[i]);
    (void) hipEventDestroy(ev_mc[i + 1]);
  }
  ev_mc.clear();
  for (size_t i = 0; i + 1 < ev_x.size(); i += 2) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, ev_x[i], ev_x[i + 1]) == hipSuccess) {
      x_ms += ms;
      x_launches++;
    }
    (void) hipEventDestroy(ev_x[i]);
    (void) hipEventDestroy(ev_x[i + 1]);
  }
  ev_x.clear();
}

size_t gd_engine::collect_bytes(int handle)
{
  GdSlot *s = find(handle);
  if (!s)
    return 0;
  return GdArena::round(s->state_bytes()) + (size_t) s->img->P.n_rep * 10 * sizeof(double);
}

int gd_engine::collect_async(GdCollect *cs, int n, hipEvent_t *done)
{
  *done = nullptr;
  GD_HIP_TRY(hipSetDevice(conf.device));
  int rc = ensure_built();   /* current cx indices; initial energies queued */
  if (rc)
    return rc;
  std::vector<int2> items;
  std::vector<GdSlot *> ss(n);
  for (int i = 0; i < n; ++i) {
    GdSlot *s = find(cs[i].handle);
    if (!s)
      return gd_fail(GD_ERR_ARG, "no complex with handle %d", cs[i].handle);
    if (!cs[i].host || cs[i].host_bytes < collect_bytes(cs[i].handle))
      return gd_fail(GD_ERR_ARG, "read-back buffer too small");
    ss[i] = s;
    for (int r = 0; r < s->img->P.n_rep; ++r)
      items.push_back(make_int2(s->cx_index, r));
  }
  const size_t nitem = items.size();
  const size_t ib = GdArena::round(nitem * sizeof(int2));
  const size_t eb = GdArena::round(nitem * 10 * sizeof(double));
  const size_t ioff = arena_take(*this, ib, true);
  const size_t eoff = ioff == (size_t) -1 ? (size_t) -1 : arena_take(*this, eb, true);
  if (eoff == (size_t) -1) {
    if (ioff != (size_t) -1)
      arena.free(ioff, ib);
    return gd_fail(GD_ERR_NOMEM, "arena too small for the FP64 re-score");
  }
  const double *e64 = (const double *) (arena.base() + eoff);
  rc = upload(arena.base() + ioff, items.data(), nitem * sizeof(int2));
  hipError_t he = rc ? hipSuccess
                     : gd_launch_rescore64_best((const gd_pack *) (arena.base() + packs_off),
                                                (const int2 *) (arena.base() + ioff),
                                                (int) nitem, (double *) e64, stream);
  size_t first = 0;
  for (int i = 0; i < n && !rc && he == hipSuccess; ++i) {
    GdSlot *s = ss[i];
    const size_t sb = s->state_bytes();
    const int nr = s->img->P.n_rep;
    he = hipMemcpyAsync(cs[i].host, arena.base() + s->off + s->blk.mv, sb,
                        hipMemcpyDeviceToHost, stream);
    if (he == hipSuccess)
      he = hipMemcpyAsync(cs[i].host + GdArena::round(sb), e64 + first * 10,
                          (size_t) nr * 10 * sizeof(double), hipMemcpyDeviceToHost, stream);
    first += nr;
  }
  if (!rc && he == hipSuccess)
    he = hipEventCreateWithFlags(done, hipEventDisableTiming);
  if (!rc && he == hipSuccess)
    he = hipEventRecord(*done, stream);
  defer(ioff, ib);
  defer(eoff, eb);
  if (rc)
    return rc;
  if (he != hipSuccess)
    return gd_fail(GD_ERR_HIP, "read-back: %s", hipGetErrorString(he));
  return GD_OK;
}

int gd_engine::results(int handle, gd_mclog *mclog, gd_replica_result *results)
{
  GdSlot *s = find(handle);
  if (!s)
    return gd_fail(GD_ERR_ARG, "no complex with handle %d", handle);
  std::vector<char> host(collect_bytes(handle));
  GdCollect c;
  c.handle = handle;
  c.host = host.data();
  c.host_bytes = host.size();
  hipEvent_t done = nullptr;
  int rc = collect_async(&c, 1, &done);
  hipError_t he = hipStreamSynchronize(stream);
  if (done)
    (void) hipEventDestroy(done);
  if (rc)
    return rc;
  if (he != hipSuccess)
    return gd_fail(GD_ERR_HIP, "read-back: %s", hipGetErrorString(he));
  drain_timing();
  reap();
  gd_parse_state(*s->img, s->blk, host.data(),
                 (const double *) (host.data() + GdArena::round(s->state_bytes())), mclog,
                 results);
  return GD_OK;
}

/* ------------------------------------------------------------------------ */
/* C ABI: engine                                                             */
/* ------------------------------------------------------------------------ */
static gd_options default_opts(const gd_options *opt)
{
  gd_options o;
  memset(&o, 0, sizeof o);
  if (opt)
    return *opt;
  o.seed = 1;
  o.complex_id = 0;
  o.exchange = 1;
  o.away = 0;
  o.term_order = GD_TERM_ORDER_FILE;
  o.device = 0;
  return o;
}

extern "C" int gd_engine_create(const gd_engine_config *cfg, gd_engine **out)
{
  if (!out)
    return gd_fail(GD_ERR_ARG, "null engine out pointer");
  *out = nullptr;
  gd_engine *e = new gd_engine();
  int rc = e->init(cfg);
  if (rc) {
    delete e;
    return rc;
  }
  *out = e;
  return GD_OK;
}

extern "C" void gd_engine_destroy(gd_engine *e) { delete e; }

extern "C" int gd_engine_add(gd_engine *e, const gd_job *job, int *handle)
{
  if (!e || !job || !handle)
    return gd_fail(GD_ERR_ARG, "null argument");
  const gd_options o = default_opts(job->opt);
  return e->add(job->lig, job->prt, job->psp, job->kde, job->mcs, job->enepara, job->temp,
                job->replica, job->mcpara, job->size, o, handle);
}

extern "C" int gd_engine_remove(gd_engine *e, int handle)
{
  if (!e)
    return gd_fail(GD_ERR_ARG, "null engine");
  return e->remove(handle);
}

extern "C" int gd_engine_cycles(gd_engine *e, int n)
{
  if (!e)
    return gd_fail(GD_ERR_ARG, "null engine");
  return e->cycles(n);
}

extern "C" int gd_engine_sync(gd_engine *e)
{
  if (!e)
    return gd_fail(GD_ERR_ARG, "null engine");
  return e->sync();
}

extern "C" int gd_engine_cycles_left(gd_engine *e, int handle)
{
  if (!e)
    return gd_fail(GD_ERR_ARG, "null engine");
  GdSlot *s = e->find(handle);
  if (!s)
    return gd_fail(GD_ERR_ARG, "no complex with handle %d", handle);
  const int left = s->total_steps - s->steps_done;
  return left <= 0 ? 0 : (left + s->spe - 1) / s->spe;
}

extern "C" int gd_engine_results(gd_engine *e, int handle, gd_mclog *mclog,
                                 gd_replica_result *results)
{
  if (!e)
    return gd_fail(GD_ERR_ARG, "null engine");
  return e->results(handle, mclog, results);
}

extern "C" int gd_engine_autotune(gd_engine *e, gd_tune_report *rep)
{
  if (!e)
    return gd_fail(GD_ERR_ARG, "null engine");
  return e->autotune(rep);
}

extern "C" int gd_engine_set_timing(gd_engine *e, int on)
{
  if (!e)
    return gd_fail(GD_ERR_ARG, "null engine");
  e->timing = on ? 1 : 0;
  return GD_OK;
}

extern "C" int gd_engine_timing(gd_engine *e, double *mc_ms, int *mc_launches, double *x_ms,
                                int *x_launches)
{
  if (!e)
    return gd_fail(GD_ERR_ARG, "null engine");
  e->drain_timing();
  if (mc_ms) *mc_ms = e->mc_ms;
  if (mc_launches) *mc_launches = e->mc_launches;
  if (x_ms) *x_ms = e->x_ms;
  if (x_launches) *x_launches = e->x_launches;
  return GD_OK;
}

extern "C" int gd_engine_layout(gd_engine *e, gd_engine_layout_info *info)
{
  if (!e || !info)
    return gd_fail(GD_ERR_ARG, "null argument");
  int rc = e->ensure_built();
  if (rc)
    return rc;
  info->stage_bytes = e->lc.stage_bytes;
  info->wpg = e->lc.wpg;
  info->persistent = e->lc.persistent;
  info->n_tasks = e->n_tasks;
  info->n_chains = e->n_chain;
  info->n_complexes = e->n_cx;
  info->grid = e->grid();
  info->resident_groups = e->resident();
  info->n_classes = (int) e->classes.size();
  info->kde_global = e->kde_global;
  info->occ = e->lc.occ;
  info->split = e->split_eff;
  info->arena_used = e->arena.used();
  info->arena_size = e->arena.size();
  int kde_lds = 1;
  for (auto *s : e->slots)
    kde_lds &= s->img->P.kde_in_lds;
  info->kde_in_lds = kde_lds;
  return GD_OK;
}

extern "C" int gd_job_footprint(const gd_job *job, int device, size_t *bytes)
{
  if (!job || !bytes)
    return gd_fail(GD_ERR_ARG, "null argument");
  int rc = gd_check_inputs(job->lig, job->prt, job->psp, job->kde, job->mcs, job->enepara,
                           job->size);
  if (rc)
    return rc;
  int max_lds = 65536;
  if (gd_device_count() > 0)
    (void) hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device);
  GdImage I;
  const gd_options o = default_opts(job->opt);
  rc = gd_build_image(I, job->lig, job->prt, job->psp, job->kde, job->mcs, job->enepara,
                      job->temp, job->replica, job->mcpara, job->size, o, max_lds, 0);
  if (rc)
    return rc;
  *bytes = gd_image_footprint(I);
  return GD_OK;
}

/* ------------------------------------------------------------------------ */
/* C ABI: gd_run (drop-in for Run(), run.cu:40-352) and the session          */
/* ------------------------------------------------------------------------ */
struct gd_session {
  gd_engine eng;
  int handle = -1;
  int spe = 1, total = 0;
};

extern "C" int gd_session_create(const gd_ligand *lig, const gd_protein *prt,
                                 const gd_psp *psp, const gd_kde *kde, const gd_mcs *mcs,
                                 const gd_enepara *enepara, const gd_temp *temp,
                                 const gd_replica *replica, const gd_mcpara *mcpara,
                                 gd_complex_size size, const gd_options *opt,
                                 gd_session **out)
{
  if (!out)
    return gd_fail(GD_ERR_ARG, "null session out pointer");
  *out = nullptr;
  int rc = gd_check_inputs(lig, prt, psp, kde, mcs, enepara, size);
  if (rc)
    return rc;
  if (!mcpara || !temp)
    return gd_fail(GD_ERR_ARG, "null temp/mcpara");
  const gd_options o = default_opts(opt);
  gd_job job;
  memset(&job, 0, sizeof job);
  job.lig = lig; job.prt = prt; job.psp = psp; job.kde = kde; job.mcs = mcs;
  job.enepara = enepara; job.temp = temp; job.replica = replica; job.mcpara = mcpara;
  job.size = size; job.opt = &o;
  size_t bytes = 0;
  rc = gd_job_footprint(&job, o.device, &bytes);
  if (rc)
    return rc;
  std::unique_ptr<gd_session> s(new gd_session());
  gd_engine_config c;
  memset(&c, 0, sizeof c);
  c.device = o.device;
  c.arena_bytes = bytes + ((size_t) 8 << 20) + (size_t) size.n_rep * 256;
  /* one complex: staging for its block and task lists only */
  c.staging_bytes = bytes + ((size_t) 1 << 20);
  c.wpg = 8;
  c.persistent = 0;
  c.use_graph = 0;
  rc = s->eng.init(&c);
  if (rc)
    return rc;
  rc = s->eng.add(lig, prt, psp, kde, mcs, enepara, temp, replica, mcpara, size, o,
                  &s->handle);
  if (rc)
    return rc;
  s->spe = mcpara->steps_per_exchange;
  s->total = s->eng.slots[0]->total_steps;
  *out = s.release();
  return GD_OK;
}

extern "C" void gd_session_destroy(gd_session *s) { delete s; }

/* initial energies happen at admission (gd_session_create); kept for the
 * create / init / cycles call sequence */
extern "C" int gd_session_init(gd_session *s)
{
  if (!s)
    return gd_fail(GD_ERR_ARG, "null session");
  return GD_OK;
}

extern "C" int gd_session_cycles(gd_session *s, int n_cycles)
{
  if (!s)
    return gd_fail(GD_ERR_ARG, "null session");
  return s->eng.cycles(n_cycles);
}

extern "C" int gd_session_sync(gd_session *s)
{
  if (!s)
    return gd_fail(GD_ERR_ARG, "null session");
  return s->eng.sync();
}

extern "C" long long gd_session_steps_done(const gd_session *s)
{
  if (!s || s->eng.slots.empty())
    return -1;
  return s->eng.slots[0]->steps_done;
}

extern "C" int gd_session_results(gd_session *s, gd_mclog *mclog, gd_replica_result *results)
{
  if (!s)
    return gd_fail(GD_ERR_ARG, "null session");
  return s->eng.results(s->handle, mclog, results);
}

extern "C" int gd_session_set_timing(gd_session *s, int on)
{
  if (!s)
    return gd_fail(GD_ERR_ARG, "null session");
  s->eng.timing = on ? 1 : 0;
  return GD_OK;
}

extern "C" int gd_session_timing(gd_session *s, double *mc_ms, int *mc_launches, double *x_ms,
                                 int *x_launches)
{
  if (!s)
    return gd_fail(GD_ERR_ARG, "null session");
  return gd_engine_timing(&s->eng, mc_ms, mc_launches, x_ms, x_launches);
}

extern "C" int gd_session_layout(const gd_session *s, int *lds_bytes, int *kde_in_lds,
                                 int *n_tasks, int *waves_per_group)
{
  if (!s)
    return gd_fail(GD_ERR_ARG, "null session");
  gd_engine_layout_info info;
  int rc = gd_engine_layout(const_cast<gd_engine *>(&s->eng), &info);
  if (rc)
    return rc;
  if (lds_bytes) *lds_bytes = info.stage_bytes;
  if (kde_in_lds) *kde_in_lds = info.kde_in_lds;
  if (n_tasks) *n_tasks = info.n_tasks;
  if (waves_per_group) *waves_per_group = info.wpg;
  return GD_OK;
}

/* Records of the current steps_per_dump window of a one-complex engine:
 * first step and records per replica, then [n_rep][n_step] records. */
static int window_records(gd_engine &E, GdSlot *sl, int *first_step, int *n_step,
                          gd_step_record *dst)
{
  const gd_pack &P = sl->img->P;
  if (P.rec_stride <= 0)
    return gd_fail(GD_ERR_ARG, "records were not requested (gd_options.record_stride = 0)");
  int rc = E.sync();
  if (rc)
    return rc;
  const int done = sl->steps_done;
  if (done == 0) {
    *first_step = 0;
    *n_step = 0;
    return GD_OK;
  }
  const int first = (done - 1) / P.rec_window * P.rec_window;
  const int n = (done - first + P.rec_stride - 1) / P.rec_stride;
  *first_step = first;
  *n_step = n;
  if (dst) {
    const size_t rb = GD_REC_WORDS * 4;
    GD_HIP_TRY(hipMemcpy2D(dst, (size_t) n * rb, E.arena.base() + sl->off + sl->blk.rec,
                           (size_t) P.rec_per_window * rb, (size_t) n * rb, P.n_rep,
                           hipMemcpyDeviceToHost));
  }
  return GD_OK;
}

extern "C" int gd_session_records(gd_session *s, int *first_step, int *n_step,
                                  gd_step_record *dst)
{
  static_assert(sizeof(gd_step_record) == GD_REC_WORDS * 4, "record layout");
  if (!s || !first_step || !n_step)
    return gd_fail(GD_ERR_ARG, "null argument");
  if (s->eng.slots.empty())
    return gd_fail(GD_ERR_ARG, "empty session");
  return window_records(s->eng, s->eng.slots[0], first_step, n_step, dst);
}

/* <prefix>_%04d.gdtraj for one window (the reference's DumpLigRecord,
 * hdf5io.C:10-30, called per steps_per_dump window, gpusingle.cu:26-34) */
static int dump_window(gd_session *s, const char *prefix, int seq)
{
  GdSlot *sl = s->eng.slots[0];
  int first = 0, n = 0;
  int rc = window_records(s->eng, sl, &first, &n, nullptr);
  if (rc)
    return rc;
  const int n_rep = sl->img->P.n_rep;
  std::vector<gd_step_record> recs((size_t) n_rep * n);
  rc = window_records(s->eng, sl, &first, &n, recs.data());
  if (rc)
    return rc;
  char path[4096];
  snprintf(path, sizeof path, "%s_%04d.gdtraj", prefix, seq);
  FILE *f = fopen(path, "wb");
  if (!f)
    return gd_fail(GD_ERR_IO, "cannot write %s", path);
  gd_traj_header h;
  memset(&h, 0, sizeof h);
  memcpy(h.magic, "GDTRAJ01", 8);
  h.n_rep = n_rep;
  h.n_step = n;
  h.record_bytes = (int32_t) sizeof(gd_step_record);
  h.first_step = first;
  h.stride = sl->img->P.rec_stride;
  const bool ok = fwrite(&h, sizeof h, 1, f) == 1 &&
                  (recs.empty() || fwrite(recs.data(), sizeof(gd_step_record), recs.size(), f) ==
                                       recs.size());
  if (fclose(f) != 0 || !ok)
    return gd_fail(GD_ERR_IO, "short write to %s", path);
  return GD_OK;
}

extern "C" int gd_run(const gd_ligand *lig, const gd_protein *prt, const gd_psp *psp,
                      const gd_kde *kde, const gd_mcs *mcs, const gd_enepara *enepara,
                      const gd_temp *temp, const gd_replica *replica,
                      const gd_mcpara *mcpara, gd_mclog *mclog, gd_complex_size size,
                      const gd_options *opt, gd_replica_result *results)
{
  auto t_begin = std::chrono::steady_clock::now();
  gd_session *s = nullptr;
  int rc = gd_session_create(lig, prt, psp, kde, mcs, enepara, temp, replica, mcpara, size,
                             opt, &s);
  if (rc)
    return rc;
  /* the reference's schedule: whole steps_per_dump windows of whole
   * launches (gpusingle.cu:8-20), steps_run / spe cycles in all */
  const int spe = mcpara->steps_per_exchange;
  const int n_cycles = s->total / spe;
  /* the reference's clocks cover its MC loop only (gpusingle.cu:8-41):
   * setup (here also the task lists and the launch shape, else built on
   * the first cycle), upload and the initial energies finish before they
   * start */
  rc = s->eng.ensure_built();
  if (!rc)
    rc = s->eng.sync();
  if (rc) {
    gd_session_destroy(s);
    return rc;
  }
  const auto t_loop = std::chrono::steady_clock::now();
  double t_io = 0.0;
  if (opt && opt->record_stride > 0 && opt->record_prefix) {
    /* one window at a time, dumped after it ends (records require spe to
     * divide steps_per_dump, so a window is steps_per_dump / spe cycles) */
    const int per_window = mcpara->steps_per_dump / spe;
    int seq = 0;
    for (int c = 0; c < n_cycles && !rc; c += per_window) {
      rc = s->eng.cycles(std::min(per_window, n_cycles - c));
      /* the window's cycles are compute: only the copy and the file count
       * as I/O (the reference syncs after every launch, cutil_yeah.h) */
      if (!rc)
        rc = s->eng.sync();
      auto t_w = std::chrono::steady_clock::now();
      if (!rc)
        rc = dump_window(s, opt->record_prefix, seq++);
      t_io += std::chrono::duration<double>(std::chrono::steady_clock::now() - t_w).count();
    }
  } else {
    rc = s->eng.cycles(n_cycles);
  }
  if (!rc)
    rc = s->eng.sync();
  auto t_comp = std::chrono::steady_clock::now();
  if (!rc)
    rc = s->eng.results(s->handle, mclog, results);
  auto t_end = std::chrono::steady_clock::now();
  if (!rc && mclog) {
    /* t0: the cycles (compute time), t1: the cycles and the record dumps
     * (wall time), as gpusingle.cu:10,31,41 take them; t2 (unused by the
     * reference): the whole call, from setup to the results on the host */
    const double t_mc = std::chrono::duration<double>(t_comp - t_loop).count();
    mclog->t0 += t_mc - t_io;
    mclog->t1 += t_mc;
    mclog->t2 += std::chrono::duration<double>(t_end - t_begin).count();
  }
  gd_session_destroy(s);
  return rc;
}

/* ------------------------------------------------------------------------ */
/* C ABI: batch (many complexes, one engine, autotuned, graph-run)           */
/* ------------------------------------------------------------------------ */
extern "C" int gd_batch_run(int n_jobs, const gd_job *jobs, const gd_batch_config *cfg,
                            gd_replica_result **results, gd_mclog *mclogs,
                            gd_batch_report *report)
{
  if (n_jobs < 0 || (n_jobs > 0 && !jobs))
    return gd_fail(GD_ERR_ARG, "bad job array");
  gd_batch_config bc;
  memset(&bc, 0, sizeof bc);
  if (cfg)
    bc = *cfg;
  size_t total = (size_t) 16 << 20;
  int max_cycles = 0;
  for (int i = 0; i < n_jobs; ++i) {
    size_t b = 0;
    int rc = gd_job_footprint(&jobs[i], bc.device, &b);
    if (rc)
      return rc;
    total += b + (size_t) jobs[i].size.n_rep * 256;
    int run = 0, lpw = 0;
    rc = gd_schedule(jobs[i].mcpara, &run, &lpw);
    if (rc)
      return rc;
    max_cycles = std::max(max_cycles, run / jobs[i].mcpara->steps_per_exchange);
  }
  gd_engine eng;
  gd_engine_config c;
  memset(&c, 0, sizeof c);
  c.device = bc.device;
  c.arena_bytes = total;
  c.wpg = bc.wpg ? bc.wpg : 8;
  c.persistent = bc.persistent;
  c.use_graph = bc.no_graph ? 0 : 1;
  c.kde_policy = bc.kde_policy;
  c.occ = bc.occ;
  c.split = bc.split;
  int rc = eng.init(&c);
  if (rc)
    return rc;
  std::vector<int> handles(n_jobs);
  auto t0 = std::chrono::steady_clock::now();
  for (int i = 0; i < n_jobs; ++i) {
    const gd_options o = default_opts(jobs[i].opt);
    rc = eng.add(jobs[i].lig, jobs[i].prt, jobs[i].psp, jobs[i].kde, jobs[i].mcs,
                 jobs[i].enepara, jobs[i].temp, jobs[i].replica, jobs[i].mcpara, jobs[i].size,
                 o, &handles[i]);
    if (rc)
      return rc;
  }
  rc = eng.sync();
  if (rc)
    return rc;
  auto t1 = std::chrono::steady_clock::now();
  gd_tune_report tr;
  memset(&tr, 0, sizeof tr);
  if (bc.autotune) {
    rc = eng.autotune(&tr);
    if (rc)
      return rc;
  }
  auto t2 = std::chrono::steady_clock::now();
  rc = eng.cycles(max_cycles);
  if (!rc)
    rc = eng.sync();
  if (rc)
    return rc;
  auto t3 = std::chrono::steady_clock::now();
  for (int i = 0; i < n_jobs; ++i) {
    rc = eng.results(handles[i], mclogs ? &mclogs[i] : nullptr, results ? results[i] : nullptr);
    if (rc)
      return rc;
  }
  auto t4 = std::chrono::steady_clock::now();
  if (report) {
    report->admit_s = std::chrono::duration<double>(t1 - t0).count();
    report->tune_s = std::chrono::duration<double>(t2 - t1).count();
    report->run_s = std::chrono::duration<double>(t3 - t2).count();
    report->results_s = std::chrono::duration<double>(t4 - t3).count();
    report->cycles = max_cycles;
    report->tune = tr;
    gd_engine_layout_info info;
    if (gd_engine_layout(&eng, &info) == GD_OK)
      report->layout = info;
  }
  return GD_OK;
}

/* ------------------------------------------------------------------------ */
/* gd_score_poses                                                            */
/* ------------------------------------------------------------------------ */
namespace {

/* Device layout of one pose-scoring call inside one region: the complex
 * image, the task list, the pose order and arrays, the FP32 and FP64
 * energies, each at a 256-byte aligned offset. */
struct ScoreLayout {
  size_t img, tasks, order, lc, pc, mv, e, e64, total;
};

ScoreLayout score_layout(size_t img_bytes, size_t n_tasks, int n_pose)
{
  ScoreLayout L;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    const size_t o = off;
    off += GdArena::round(bytes);
    return o;
  };
  L.img = take(img_bytes);
  L.tasks = take(n_tasks * sizeof(gd_task));
  L.order = take((size_t) n_pose * 4);
  L.lc = take((size_t) n_pose * 4);
  L.pc = take((size_t) n_pose * 4);
  L.mv = take((size_t) n_pose * 24);
  L.e = take((size_t) n_pose * 40);
  L.e64 = take((size_t) n_pose * 80);
  L.total = off;
  return L;
}

/* Argument checks, the host image of the complex (the pose arrays ride in it
 * as a one-temperature replica set) and the workgroup tasks: runs of up to 8
 * poses on one protein conformer, poses ordered by conformer. */
int score_prepare(const gd_ligand *lig, const gd_protein *prt, const gd_psp *psp,
                  const gd_kde *kde, const gd_mcs *mcs, const gd_enepara *enepara,
                  gd_complex_size size, int n_pose, const int *lig_conf, const int *prt_conf,
                  const float *move, int term_order, int device, float *energy,
                  double *energy_fp64, GdImage &I, std::vector<int> &order,
                  std::vector<gd_task> &tasks)
{
  int rc = gd_check_inputs(lig, prt, psp, kde, mcs, enepara, size);
  if (rc)
    return rc;
  if (n_pose < 0 || (n_pose > 0 && (!lig_conf || !prt_conf || !move)))
    return gd_fail(GD_ERR_ARG, "bad pose arrays");
  if (!energy && !energy_fp64)
    return gd_fail(GD_ERR_ARG, "no output requested");
  for (int i = 0; i < n_pose; ++i)
    if (lig_conf[i] < 0 || lig_conf[i] >= size.n_lig || prt_conf[i] < 0 ||
        prt_conf[i] >= size.n_prt)
      return gd_fail(GD_ERR_ARG, "pose %d: conformer index out of range", i);
  const int ndev = gd_device_count();
  if (ndev > 0 && (device < 0 || device >= ndev))
    return gd_fail(GD_ERR_ARG, "device %d of %d", device, ndev);
  int max_lds = 65536;
  if (ndev > 0)
    GD_HIP_TRY(hipDeviceGetAttribute(&max_lds, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                     device));
  gd_options o = default_opts(nullptr);
  o.term_order = term_order;
  o.device = device;
  gd_complex_size sz1 = size;
  sz1.n_tmp = 1;
  sz1.n_rep = size.n_lig * size.n_prt;
  rc = gd_build_image(I, lig, prt, psp, kde, mcs, enepara, nullptr, nullptr, nullptr, sz1, o,
                      max_lds, 0);
  if (rc)
    return rc;
  order.resize(n_pose);
  for (int i = 0; i < n_pose; ++i)
    order[i] = i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return prt_conf[a] < prt_conf[b]; });
  tasks.clear();
  for (int i = 0; i < n_pose;) {
    int j = i;
    while (j < n_pose && prt_conf[order[j]] == prt_conf[order[i]] && j - i < 8)
      ++j;
    tasks.push_back(gd_task{0, prt_conf[order[i]], i, j - i});
    i = j;
  }
  return GD_OK;
}

/* Uploads, launches and read-back, all ordered on st, into the device
 * region at B laid out as L; returns after st has drained. */
int score_run(char *B, const ScoreLayout &L, GdImage &I, const std::vector<gd_task> &tasks,
              const std::vector<int> &order, int n_pose, const int *lig_conf,
              const int *prt_conf, const float *move, float *energy, double *energy_fp64,
              hipStream_t st)
{
  const size_t img = gd_image_footprint(I);
  std::vector<char> host(img);
  GdBlock b;
  gd_pack_image(I, host.data(), B + L.img, b);
  const gd_pack *pack_dev = (const gd_pack *) (B + L.img + b.pack);
  /* the host buffers outlive the final synchronisation */
  hipError_t he = hipMemcpyAsync(B + L.img, host.data(), img, hipMemcpyHostToDevice, st);
  auto put = [&](const void *src, size_t bytes, size_t off) {
    if (he == hipSuccess)
      he = hipMemcpyAsync(B + off, src, bytes, hipMemcpyHostToDevice, st);
  };
  put(tasks.data(), vbytes(tasks), L.tasks);
  put(order.data(), (size_t) n_pose * 4, L.order);
  put(lig_conf, (size_t) n_pose * 4, L.lc);
  put(prt_conf, (size_t) n_pose * 4, L.pc);
  put(move, (size_t) n_pose * 24, L.mv);
  if (he == hipSuccess && energy) {
    he = gd_set_kernel_attrs(I.P.lds_bytes);
    if (he == hipSuccess)
      he = gd_launch_score(pack_dev, (const gd_task *) (B + L.tasks), (int) tasks.size(),
                           (const int *) (B + L.order), (const int *) (B + L.lc),
                           (const float *) (B + L.mv), I.P.lds_bytes, (float *) (B + L.e), st);
  }
  if (he == hipSuccess && energy_fp64)
    he = gd_launch_rescore64(pack_dev, 0, (const int *) (B + L.lc), (const int *) (B + L.pc),
                             (const float *) (B + L.mv), 6, n_pose, (double *) (B + L.e64), st);
  if (he == hipSuccess && energy)
    he = hipMemcpyAsync(energy, B + L.e, (size_t) n_pose * 40, hipMemcpyDeviceToHost, st);
  if (he == hipSuccess && energy_fp64)
    he = hipMemcpyAsync(energy_fp64, B + L.e64, (size_t) n_pose * 80, hipMemcpyDeviceToHost,
                        st);
  const hipError_t hs = hipStreamSynchronize(st);
  if (he == hipSuccess)
    he = hs;
  if (he != hipSuccess)
    return gd_fail(GD_ERR_HIP, "score: %s", hipGetErrorString(he));
  return GD_OK;
}

}  // namespace

extern "C" int gd_score_poses(const gd_ligand *lig, const gd_protein *prt,
                              const gd_psp *psp, const gd_kde *kde, const gd_mcs *mcs,
                              const gd_enepara *enepara, gd_complex_size size,
                              int n_pose, const int *lig_conf, const int *prt_conf,
                              const float *move, int term_order, int device,
                              float *energy, double *energy_fp64)
{
  GdImage I;
  std::vector<int> order;
  std::vector<gd_task> tasks;
  int rc = score_prepare(lig, prt, psp, kde, mcs, enepara, size, n_pose, lig_conf, prt_conf,
                         move, term_order, device, energy, energy_fp64, I, order, tasks);
  if (rc)
    return rc;
  if (gd_device_count() < 1)
    return gd_fail(GD_ERR_NODEV, "no HIP device visible");
  if (n_pose == 0)
    return GD_OK;
  GD_HIP_TRY(hipSetDevice(device));
  const ScoreLayout L = score_layout(gd_image_footprint(I), tasks.size(), n_pose);
  GdArena arena;
  GD_HIP_TRY(arena.init(L.total));
  struct StreamGuard {
    hipStream_t s = nullptr;
    ~StreamGuard() { if (s) (void) hipStreamDestroy(s); }
  } sg;
  GD_HIP_TRY(hipStreamCreateWithFlags(&sg.s, hipStreamNonBlocking));
  return score_run(arena.base(), L, I, tasks, order, n_pose, lig_conf, prt_conf, move, energy,
                   energy_fp64, sg.s);
}

/* The same scoring with no device allocation: the region is the static
 * per-GPU workspace (GD_SCRATCH_BYTES); work on the default stream. */
extern "C" int gd_score_poses_noalloc(const gd_ligand *lig, const gd_protein *prt,
                                      const gd_psp *psp, const gd_kde *kde,
                                      const gd_mcs *mcs, const gd_enepara *enepara,
                                      gd_complex_size size, int n_pose, const int *lig_conf,
                                      const int *prt_conf, const float *move, int term_order,
                                      int device, float *energy, double *energy_fp64)
{
  GdImage I;
  std::vector<int> order;
  std::vector<gd_task> tasks;
  int rc = score_prepare(lig, prt, psp, kde, mcs, enepara, size, n_pose, lig_conf, prt_conf,
                         move, term_order, device, energy, energy_fp64, I, order, tasks);
  if (rc)
    return rc;
  const ScoreLayout L = score_layout(gd_image_footprint(I), tasks.size(), n_pose);
  if (L.total > GD_SCRATCH_BYTES)
    return gd_fail(GD_ERR_ARG,
                   "gd_score_poses_noalloc: %d poses of this complex need %zu bytes of "
                   "workspace; the static workspace holds %zu (use gd_score_poses, or fewer "
                   "poses per call)",
                   n_pose, L.total, (size_t) GD_SCRATCH_BYTES);
  if (gd_device_count() < 1)
    return gd_fail(GD_ERR_NODEV, "no HIP device visible");
  if (n_pose == 0)
    return GD_OK;
  GD_HIP_TRY(hipSetDevice(device));
  char *base = nullptr;
  GD_HIP_TRY(gd_scratch_region(device, &base));
  return score_run(base, L, I, tasks, order, n_pose, lig_conf, prt_conf, move, energy,
                   energy_fp64, nullptr);
}

/* ------------------------------------------------------------------------ */
/* C ABI: VALU issue-rate probe                                              */
/* ------------------------------------------------------------------------ */
extern "C" int gd_valu_probe(int device, int waves_per_simd, double *ns_per_inst, double *tflops,
                             int *n_simd)
{
  if (!ns_per_inst || !tflops || !n_simd || waves_per_simd < 1 || waves_per_simd > 8)
    return gd_fail(GD_ERR_ARG, "waves_per_simd must be 1..8 and outputs non-null");
  if (gd_device_count() <= 0)
    return gd_fail(GD_ERR_NODEV, "no HIP device");
  GD_HIP_TRY(hipSetDevice(device));
  int n_cu = 0;
  GD_HIP_TRY(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device));
  /* 256-lane workgroups = 4 wavefronts, one per SIMD of a CU */
  const int blocks = n_cu * waves_per_simd;
  const int iters = 400;
  float *out = nullptr;
  GD_HIP_TRY(hipMalloc(&out, (size_t) blocks * 256 * sizeof(float)));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = gd_launch_valu_probe(out, blocks, iters, 0);   /* warm-up */
  if (e == hipSuccess) e = hipEventRecord(e0, 0);
  if (e == hipSuccess) e = gd_launch_valu_probe(out, blocks, iters, 0);
  if (e == hipSuccess) e = hipEventRecord(e1, 0);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  float ms = 0.0f;
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (e0) (void) hipEventDestroy(e0);
  if (e1) (void) hipEventDestroy(e1);
  (void) hipFree(out);
  if (e != hipSuccess)
    return gd_fail(GD_ERR_HIP, "valu probe: %s", hipGetErrorString(e));
  const double insts = (double) blocks * 4.0 * iters * 16.0 * 8.0;   /* wavefront FMAs */
  const double simds = (double) n_cu * 4.0;
  *ns_per_inst = (double) ms * 1e6 / (insts / simds);
  *tflops = insts * 64.0 * 2.0 / ((double) ms * 1e-3) / 1e12;
  *n_simd = n_cu * 4;
  return GD_OK;
}

static const char *const gd_valu_kind_names[GD_VALU_KINDS] = {
    "v_fmac_f32 (vector operands)", "v_fma_f32 (one scalar-register operand)", "v_mul_f32",
    "v_max_f32", "v_cndmask_b32 (scalar-register mask)", "v_lshl_add_u32",
    "v_cmp_gt_f32 (to vcc)", "v_add_f32_dpp", "v_readlane_b32",
    "v_fmac_f32 and v_max_f32 alternating"};

extern "C" const char *gd_valu_kind_name(int kind)
{
  return kind >= 0 && kind < GD_VALU_KINDS ? gd_valu_kind_names[kind] : nullptr;
}

extern "C" int gd_valu_probe_kinds(int device, int waves_per_simd, double *ns_per_inst,
                                   int n_kinds)
{
  if (!ns_per_inst || n_kinds != GD_VALU_KINDS || waves_per_simd < 1 || waves_per_simd > 8)
    return gd_fail(GD_ERR_ARG, "waves_per_simd must be 1..8, n_kinds %d and ns_per_inst "
                               "non-null", GD_VALU_KINDS);
  if (gd_device_count() <= 0)
    return gd_fail(GD_ERR_NODEV, "no HIP device");
  GD_HIP_TRY(hipSetDevice(device));
  int n_cu = 0;
  GD_HIP_TRY(hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, device));
  const int blocks = n_cu * waves_per_simd;
  const int iters = 200;
  float *out = nullptr;
  GD_HIP_TRY(hipMalloc(&out, (size_t) blocks * 256 * sizeof(float)));
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  for (int k = 0; k < GD_VALU_KINDS && e == hipSuccess; ++k) {
    e = gd_launch_valu_kind(k, out, blocks, 8, 0);   /* warm-up */
    if (e == hipSuccess) e = hipEventRecord(e0, 0);
    if (e == hipSuccess) e = gd_launch_valu_kind(k, out, blocks, iters, 0);
    if (e == hipSuccess) e = hipEventRecord(e1, 0);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    /* wavefront instructions per SIMD: 8 chains x 16 rounds per iteration,
     * two instructions per chain step in the alternating kind */
    const double per_simd = (double) blocks * 4.0 / (n_cu * 4.0) * iters * 16.0 * 8.0 *
                            (k == GD_VALU_KINDS - 1 ? 2.0 : 1.0);
    ns_per_inst[k] = (double) ms * 1e6 / per_simd;
  }
  if (e0) (void) hipEventDestroy(e0);
  if (e1) (void) hipEventDestroy(e1);
  (void) hipFree(out);
  if (e != hipSuccess)
    return gd_fail(GD_ERR_HIP, "valu kind probe: %s", hipGetErrorString(e));
  return GD_OK;
}
