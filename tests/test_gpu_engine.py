"""GPU tests of the engine: many complexes resident and advanced together.

Every job's trajectory must equal the CPU replay of that job alone (and so
gd_run of it alone), whatever else shares the launches, whatever the launch
shape (waves per workgroup, grid or persistent queue, graph or not), and
whenever it was admitted.
"""
import numpy as np
import pytest

import oracle_lib as O
from gpudock_amd import remc, synth

pytestmark = pytest.mark.gpu

KEYS = ("move_final", "energy_final", "move_best", "energy_best", "step_best",
        "idx_tmp_final", "acs_mc", "acs_exchg", "cmcc_final", "cmcc_best")


def _same(a, b, what=""):
    for k in KEYS:
        np.testing.assert_array_equal(a[k], b[k], err_msg=f"{what} {k}")


def _uneven_jobs(n, seed=0, steps=(30, 90)):
    """Jobs of uneven sizes, schedules and options (small enough for the replay)."""
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(n):
        lna = int(rng.integers(5, 100))
        cx, temps, mc, cfg = synth.make_config(
            "c0", seed=100 + i, lna=lna, radius=max(10.0, 3.0 * lna ** (1 / 3) + 4),
            pnp=int(rng.integers(40, 420)), pnk=int(rng.integers(0, 300)),
            n_lig=int(rng.integers(1, 4)), n_prt=int(rng.integers(1, 3)),
            n_tmp=int(rng.choice([1, 2, 3, 8])), steps=int(rng.integers(*steps)),
            spe=int(rng.choice([1, 5, 10])))
        opt = remc.default_options(seed=7 + i, complex_id=i, exchange=int(i % 4 != 3),
                                   away=int(i % 5 == 4), term_order=int(i % 3 == 2),
                                   mcc=int(i % 2 == 1))
        jobs.append(remc.JobSpec(cx, temps, mc, opt))
    return jobs


@pytest.fixture(scope="module")
def jobs():
    js = _uneven_jobs(6)
    replays = [O.replay_run(j.cx, j.temps, j.mc, j.opt) for j in js]
    return js, replays


def test_batch_equals_replay_of_each_job(jobs):
    """gd_batch_run (graph-replayed cycles, autotuned shape) on uneven jobs:
    each job's results are those of its replay alone."""
    js, replays = jobs
    res, rep = remc.batch_run(js, autotune=1)
    assert 6 <= rep["tune"]["n"] <= remc.TUNE_MAX
    assert rep["layout"]["n_complexes"] == len(js)
    assert rep["layout"]["n_classes"] >= 1
    for i, (g, r) in enumerate(zip(res, replays)):
        _same(g, r, f"job {i}")


@pytest.mark.parametrize("wpg,persistent,no_graph,kde_policy,occ", [
    (4, 0, 0, 0, 0), (8, 1, 0, 0, 6), (16, 1, 1, 0, 8), (16, 0, 0, 1, 0), (4, 1, 1, 1, 8),
    (8, 0, 0, 1, 6)])
def test_launch_shapes_identical(jobs, wpg, persistent, no_graph, kde_policy, occ):
    """Every launch shape, graph or not, KDE points in LDS or read from HBM,
    every register budget: the same bits."""
    js, replays = jobs
    res, rep = remc.batch_run(js, wpg=wpg, persistent=persistent, no_graph=no_graph,
                              kde_policy=kde_policy, occ=occ)
    assert rep["layout"]["occ"] == occ
    if kde_policy:
        assert rep["layout"]["kde_in_lds"] == 0
    for i, (g, r) in enumerate(zip(res, replays)):
        _same(g, r, f"wpg={wpg} persistent={persistent} job {i}")


@pytest.mark.parametrize("split,persistent,no_graph", [(4, 0, 0), (8, 1, 0), (16, 0, 1),
                                                       (16, 1, 0), (1, 0, 0)])
def test_split_replicas_identical(jobs, split, persistent, no_graph):
    """Each replica's pair loop spread over 4, 8 or 16 wavefronts (atom and
    KDE blocks evaluated in parallel, partials added in block order): the same
    bits as one wavefront per replica and as the replay."""
    js, replays = jobs
    res, rep = remc.batch_run(js, persistent=persistent, no_graph=no_graph, split=split)
    assert rep["layout"]["split"] == split
    for i, (g, r) in enumerate(zip(res, replays)):
        _same(g, r, f"split={split} job {i}")


def test_split_chosen_for_few_replicas(jobs):
    """With fewer resident replicas than SIMDs the engine spreads each replica
    over several wavefronts by itself, and the autotuner times those shapes
    too; the results stay the replay's."""
    js, replays = jobs
    eng = remc.Engine(arena_bytes=256 << 20)
    h = [eng.add(j) for j in js[:2]]
    assert eng.layout()["split"] > 1
    tune = eng.autotune()
    assert any(s > 1 for s in tune["split"][:tune["n"]])
    while any(eng.cycles_left(x) > 0 for x in h):
        eng.cycles(1)
    for i, x in enumerate(h):
        g, _ = eng.results(x)
        _same(g, replays[i], f"auto split job {i}")
    eng.close()


def test_alone_equals_mix(jobs):
    """gd_run of one job alone == the same job inside the mix."""
    js, replays = jobs
    for i in (0, 3):
        j = js[i]
        alone, _ = remc.run(j.cx, j.temps, j.mc, j.opt)
        _same(alone, replays[i], f"alone {i}")


def test_staggered_admission_and_retirement(jobs):
    """Jobs admitted between cycles start at their own step 0; retiring a
    finished job leaves the others untouched."""
    js, replays = jobs
    eng = remc.Engine(arena_bytes=256 << 20, wpg=8, persistent=1, use_graph=1)
    h = {}
    h[0] = eng.add(js[0])
    h[1] = eng.add(js[1])
    eng.cycles(2)
    h[2] = eng.add(js[2])
    eng.cycles(3)
    h[3] = eng.add(js[3])
    left = {k: eng.cycles_left(v) for k, v in h.items()}
    done = {}
    while any(eng.cycles_left(v) > 0 for k, v in h.items() if k not in done):
        eng.cycles(1)
        for k, v in list(h.items()):
            if k not in done and eng.cycles_left(v) == 0:
                done[k], _ = eng.results(v)
                eng.remove(v)
    for k in h:
        if k not in done:
            done[k], _ = eng.results(h[k])
    eng.close()
    assert max(left.values()) > 0
    for k, g in done.items():
        _same(g, replays[k], f"staggered job {k}")


def test_admission_refused_when_arena_full(jobs):
    js, replays = jobs
    need = js[0].footprint()
    slack = 32 << 10    # control block + task lists
    assert min(j.footprint() for j in js[1:]) > slack
    eng = remc.Engine(arena_bytes=need + slack, use_graph=0)
    h0 = eng.add(js[0])
    with pytest.raises(remc.GDError) as ei:
        eng.add(js[1])
    assert "arena" in str(ei.value)
    while eng.cycles_left(h0) > 0:
        eng.cycles(eng.cycles_left(h0))
    g, _ = eng.results(h0)
    eng.close()
    _same(g, replays[0], "after refused admission")


def test_engine_timing_and_layout(jobs):
    js, _ = jobs
    eng = remc.Engine(arena_bytes=256 << 20, use_graph=1)
    for j in js[:3]:
        eng.add(j)
    lay = eng.layout()
    assert lay["n_complexes"] == 3 and lay["n_tasks"] > 0 and lay["resident_groups"] > 0
    eng.set_timing(True)
    eng.cycles(3)
    eng.sync()
    ms, n, xms, xn = eng.timing()
    eng.close()
    assert n == 3 and xn == 3 and ms > 0.0


def test_server_results_equal_alone_under_continuous_batching(jobs):
    """Jobs submitted while others run (arrival spread over several ticks)
    are admitted as memory frees up; each result equals the job run alone."""
    import time
    js, replays = jobs
    srv = remc.Server(arena_bytes=64 << 20, wpg=8, persistent=1)
    tickets = []
    for i, j in enumerate(js):
        tickets.append(srv.submit(j))
        if i % 2 == 1:
            time.sleep(0.005)
    # the same complexes again, with other seeds, to keep the mix busy
    extra = []
    for i, j in enumerate(js[:3]):
        o = remc.default_options(seed=1000 + i, complex_id=50 + i)
        extra.append(remc.JobSpec(j.cx, j.temps, j.mc, o))
    t_extra = [srv.submit(e) for e in extra]
    srv.drain()
    for i, t in enumerate(tickets):
        info = srv.wait(t, 10.0)
        assert info["state"] == remc.JOB_DONE
        g, log, info = srv.fetch(t)
        _same(g, replays[i], f"served job {i}")
        assert info["latency"] >= info["t_admit"] - info["t_submit"] >= 0.0
    for e, t in zip(extra, t_extra):
        g, _, _ = srv.fetch(t)
        _same(g, O.replay_run(e.cx, e.temps, e.mc, e.opt), "served extra")
    st = srv.stats()
    srv.close()
    assert st["completed"] == len(js) + len(extra)
    assert st["failed"] == 0 and st["p99_latency"] >= st["p50_latency"] > 0.0
    # pose evaluations = the MC steps each job's schedule runs (whole launches,
    # gpusingle.cu:8-20) x its replicas
    runs = [O.steps_run(j.mc) for j in js + extra]
    assert st["pose_evals"] == sum(j.n_rep * r for j, r in zip(js + extra, runs))


def test_server_lanes_route_by_size(jobs):
    """Three lanes (bounds on ligand atoms x pocket points), the first two
    with raised wavefront priority: each job goes to the first lane whose
    bound holds its size, the lanes cycle on their own, and every result
    equals the replay of the job alone."""
    js, replays = jobs
    pairs = [j.cx.lig_xyz.shape[1] * j.cx.prt_xyz.shape[1] for j in js]
    lo, hi = sorted(pairs)[1], sorted(pairs)[3]
    srv = remc.Server(arena_bytes=96 << 20, wpg=8, lane_max_pairs=(lo, hi), lane_prio=(3, 1, 0))
    tickets = [srv.submit(j) for j in js]
    srv.drain()
    lanes = []
    for i, t in enumerate(tickets):
        g, _, info = srv.fetch(t)
        _same(g, replays[i], f"lane-served job {i}")
        lanes.append(info["lane"])
    want = [0 if p <= lo else 1 if p <= hi else 2 for p in pairs]
    assert lanes == want and set(want) == {0, 1, 2}
    st = srv.stats()
    srv.close()
    assert st["lanes"] == 3 and st["completed"] == len(js)
    assert st["lane_completed"][:3] == [want.count(k) for k in range(3)]
    assert all(c > 0 for c in st["lane_cycles"][:3])
    assert st["cycles"] == sum(st["lane_cycles"])
    with pytest.raises(remc.GDError):
        remc.Server(lane_max_pairs=(hi, lo))


def test_server_split_lane(jobs):
    """A first lane run with 16 wavefronts per replica (KDE points in LDS)
    beside a lane held to one wavefront per replica: small jobs go to it, a small job with more
    than 64 restraints (which no split launch holds) passes on to the next
    lane, and every result equals the replay of the job alone."""
    js, replays = jobs
    pairs = [j.cx.lig_xyz.shape[1] * j.cx.prt_xyz.shape[1] for j in js]
    bound = sorted(pairs)[3]
    cx, temps, mc, _ = synth.make_config("c0", seed=77, lna=12, pnp=90, pnk=40, pos=80,
                                         n_lig=1, n_prt=1, n_tmp=2, steps=40, spe=5)
    many = remc.JobSpec(cx, temps, mc, remc.default_options(seed=77, complex_id=40))
    assert 12 * 90 <= bound
    srv = remc.Server(arena_bytes=96 << 20, wpg=8, lane_max_pairs=(bound,), lane_split=(16, 1))
    tickets = [srv.submit(j) for j in js]
    t_many = srv.submit(many)
    srv.drain()
    lanes = []
    for i, t in enumerate(tickets):
        g, _, info = srv.fetch(t)
        _same(g, replays[i], f"split-lane job {i}")
        lanes.append(info["lane"])
    g, _, info = srv.fetch(t_many)
    _same(g, O.replay_run(many.cx, many.temps, many.mc, many.opt), "passed-on job")
    assert info["lane"] == 1
    assert lanes == [0 if p <= bound else 1 for p in pairs]
    st = srv.stats()
    srv.close()
    assert st["lane_split"][:2] == [16, 1]
    with pytest.raises(remc.GDError):
        remc.Server(lane_max_pairs=(bound,), lane_split=(3, 0))


def test_server_lane_overflow(jobs):
    """A split lane that holds at most one job passes the small jobs that
    arrive while it is busy to the next lane, and one gated on the whole
    server passes them on while any job is in it; every result still equals
    the replay of the job alone."""
    js, replays = jobs
    pairs = [j.cx.lig_xyz.shape[1] * j.cx.prt_xyz.shape[1] for j in js]
    bound = sorted(pairs)[3]
    srv = remc.Server(arena_bytes=96 << 20, wpg=8, lane_max_pairs=(bound,), lane_split=(16, 1),
                      lane_max_jobs=(1, 0))
    tickets = [srv.submit(j) for j in js]
    srv.drain()
    lanes = []
    for i, t in enumerate(tickets):
        g, _, info = srv.fetch(t)
        _same(g, replays[i], f"overflow job {i}")
        lanes.append(info["lane"])
    st = srv.stats()
    srv.close()
    first_small = next(i for i, p in enumerate(pairs) if p <= bound)
    assert lanes[first_small] == 0
    assert all(ln == 1 for ln, p in zip(lanes, pairs) if p > bound)
    assert st["lane_completed"][:2] == [lanes.count(0), lanes.count(1)]
    # gated on the whole server instead: while a long job is in the server
    # the split lane is closed to small jobs; in an empty server it is open
    small = [i for i, p in enumerate(pairs) if p <= bound]
    big = next(i for i, p in enumerate(pairs) if p > bound)
    long_job = remc.JobSpec(js[big].cx, js[big].temps,
                            remc.mcpara(200000, 10, list(js[big].mc.move_scale)), js[big].opt)
    srv = remc.Server(arena_bytes=96 << 20, wpg=8, lane_max_pairs=(bound,), lane_split=(16, 1),
                      lane_max_total=(1, 0))
    t_long = srv.submit(long_job)
    tickets = {i: srv.submit(js[i]) for i in small}
    for i, t in tickets.items():
        info = srv.wait(t, 60.0)
        assert info["state"] == remc.JOB_DONE and info["lane"] == 1
        g, _, _ = srv.fetch(t)
        _same(g, replays[i], f"gated job {i}")
    srv.cancel(t_long)
    srv.drain()
    assert srv.wait(t_long, 10.0)["state"] == remc.JOB_CANCELLED
    t = srv.submit(js[small[0]])
    assert srv.wait(t, 60.0)["state"] == remc.JOB_DONE
    g, _, info = srv.fetch(t)
    assert info["lane"] == 0
    _same(g, replays[small[0]], "gated job in an empty server")
    srv.close()


def test_server_lanes_cancel_and_close(jobs):
    """Lanes whose arenas hold one job each: a queued job cancelled in one
    lane never runs while the others complete; closing the server with jobs
    resident in several lanes and more queued ends every job as cancelled
    (no hang)."""
    js, replays = jobs
    pairs = [j.cx.lig_xyz.shape[1] * j.cx.prt_xyz.shape[1] for j in js]
    bound = sorted(pairs)[2]
    need = max(j.footprint() for j in js)
    srv = remc.Server(arena_bytes=2 * (need + need // 2), wpg=8, lane_max_pairs=(bound,))
    small = [i for i, p in enumerate(pairs) if p <= bound]
    t = [srv.submit(js[i]) for i in small]
    t_cancel = srv.submit(js[small[0]])
    srv.cancel(t_cancel)
    srv.drain()
    for i, x in zip(small, t):
        g, _, info = srv.fetch(x)
        _same(g, replays[i], f"lane 0 job {i}")
        assert info["lane"] == 0
    assert srv.wait(t_cancel, 10.0)["state"] == remc.JOB_CANCELLED
    # a long job in each lane, more queued behind them, then close
    long_jobs = []
    for i in (small[0], next(i for i, p in enumerate(pairs) if p > bound)):
        cx, temps, mc = js[i].cx, js[i].temps, remc.mcpara(200000, 10, list(js[i].mc.move_scale))
        long_jobs.append(remc.JobSpec(cx, temps, mc, js[i].opt))
    tl = [srv.submit(j) for j in long_jobs + long_jobs]
    st = srv.stats()
    assert st["submitted"] == len(small) + 1 + 4
    srv.close()


def test_server_admission_by_footprint_and_cancel(jobs):
    """An arena that holds one job at a time serialises admission (FIFO);
    a cancelled queued job never runs; a job larger than the arena is
    refused at submission."""
    js, replays = jobs
    need = max(j.footprint() for j in js[:3])
    fp = sorted(j.footprint() for j in js[:3])
    assert fp[0] + fp[1] > need + need // 2     # no two fit together
    srv = remc.Server(arena_bytes=need + need // 2, wpg=4)
    t = [srv.submit(j) for j in js[:3]]
    t_cancel = srv.submit(js[0])
    srv.cancel(t_cancel)
    srv.drain()
    infos = [srv.wait(x, 10.0) for x in t]
    for i, x in enumerate(t):
        g, _, _ = srv.fetch(x)
        _same(g, replays[i], f"footprint-admitted job {i}")
    # one resident at a time: each admission follows the previous completion
    for a, b in zip(infos, infos[1:]):
        assert b["t_admit"] >= a["t_done"] - 0.05
    assert srv.wait(t_cancel, 10.0)["state"] == remc.JOB_CANCELLED
    big = synth.make_config("c0", pnp=2000, n_lig=4, n_tmp=64, radius=16.0)
    with pytest.raises(remc.GDError):
        srv.submit(remc.JobSpec(big[0], big[1], big[2]))
    st = srv.stats()
    srv.close()
    assert st["running_peak"] == 1 and st["cancelled"] == 1
