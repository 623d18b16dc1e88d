"""The batched and served workloads on the GPU, at their own size
distributions.

Batched uneven complexes (ligand atoms U[12, 96], pocket points U[200, 1500],
ligand conformers U[4, 64], 8 or 16 temperatures; synth.make_batch_jobs)
and served jobs (lognormal ligand and pocket sizes, 16 ligand conformers x 8
temperatures, Poisson arrivals; synth.make_serving_pool):

- a sample of each, with few MC steps so that the CPU replay stays short:
  every job's results equal its replay alone, bit for bit, through the
  batch path (autotuned, graph-replayed) and through the server under
  continuous batching;
- the full batch of 256 complexes and a 200-jobs/s arrival stream, with
  fewer steps than the benchmark: properties every job must have (finite
  energies, acceptance strictly between 0 and 1 over the job, each chain's
  temperatures a permutation, the best energy no higher than the start
  and the final one, the FP32 best energies within the per-term tolerance of
  their FP64 re-score).
"""
import math
import time

import numpy as np
import pytest

import oracle_lib as O
from gpudock_amd import remc, synth
from test_gpu_engine import _same
from test_gpu_parity import tolerance

pytestmark = pytest.mark.gpu


def _with_steps(spec, steps):
    mc = remc.mcpara(steps, spec.mc.steps_per_exchange, list(spec.mc.move_scale))
    return remc.JobSpec(spec.cx, spec.temps, mc, spec.opt)


def _check_properties(spec, g):
    n_tmp = len(spec.temps)
    size = spec.size
    for k in ("energy_final", "energy_best", "move_final", "move_best", "energy_best_fp64"):
        assert np.all(np.isfinite(g[k])), k
    run = O.steps_run(spec.mc)
    acc = g["acs_mc"].sum() / (size.n_rep * run)
    assert 0.0 < acc < 1.0
    assert np.all(g["energy_best"][:, 9] <= g["energy_final"][:, 9])
    it = g["idx_tmp_final"].reshape(size.n_prt, n_tmp, size.n_lig)
    for p in range(size.n_prt):
        for l in range(size.n_lig):
            assert sorted(it[p, :, l]) == list(range(n_tmp))
    tol = tolerance(spec.cx, g["energy_best_fp64"], spec.opt.term_order)
    assert np.all(np.abs(g["energy_best"] - g["energy_best_fp64"]) <= tol)


def test_batched_sample_equals_replay():
    """10 complexes drawn from the batched workload's distributions, 20 steps:
    gd_batch_run (autotuned, graph-replayed) gives each the bits of its replay."""
    specs = [_with_steps(s, 20) for s in synth.make_batch_jobs(n=10, seed=5)]
    res, rep = remc.batch_run(specs, autotune=1)
    assert rep["layout"]["n_complexes"] == 10
    for i, (s, g) in enumerate(zip(specs, res)):
        _same(g, O.replay_run(s.cx, s.temps, s.mc, s.opt), f"batched job {i}")


def test_batched_full_size_properties():
    """All 256 complexes of the batched workload (one launch stream), 200 steps."""
    specs = [_with_steps(s, 200) for s in synth.make_batch_jobs(n=256, seed=1)]
    res, rep = remc.batch_run(specs, autotune=0)
    assert rep["layout"]["n_complexes"] == 256
    for s, g in zip(specs, res):
        _check_properties(s, g)


def _poisson_serve(srv, specs, rate, rng):
    tickets = []
    t_next = time.perf_counter()
    for s in specs:
        t_next += rng.exponential(1.0 / rate)
        while time.perf_counter() < t_next:
            time.sleep(0.0005)
        tickets.append(srv.submit(s))
    srv.drain()
    return tickets


def test_served_sample_equals_replay():
    """40 jobs of the serving mix arriving as a Poisson stream at 200/s, 30
    steps each: every served result equals the job's replay alone."""
    pool = synth.make_serving_pool(n=40, seed=4, steps=30)
    specs = [remc.JobSpec(cx, t, mc, remc.default_options(seed=11, complex_id=i))
             for i, (cx, t, mc) in enumerate(pool)]
    srv = remc.Server(wpg=8, persistent=1)
    tickets = _poisson_serve(srv, specs, 200.0, np.random.default_rng(8))
    for i, (s, t) in enumerate(zip(specs, tickets)):
        g, _, info = srv.fetch(t)
        assert info["state"] == remc.JOB_DONE
        _same(g, O.replay_run(s.cx, s.temps, s.mc, s.opt), f"served job {i}")
    st = srv.stats()
    srv.close()
    assert st["completed"] == 40 and st["failed"] == 0


def test_served_stream_properties():
    """400 jobs of the serving mix at 200/s (2 s of arrivals), 100 steps each:
    all complete, with the properties above and latency statistics."""
    pool = synth.make_serving_pool(n=64, seed=2, steps=100)
    rng = np.random.default_rng(9)
    picks = rng.integers(0, len(pool), 400)
    specs = [remc.JobSpec(*pool[k], remc.default_options(seed=3, complex_id=i))
             for i, k in enumerate(picks)]
    srv = remc.Server(wpg=8, persistent=1, use_graph=1)
    tickets = _poisson_serve(srv, specs, 200.0, rng)
    for s, t in zip(specs, tickets):
        g, _, info = srv.fetch(t)
        assert info["state"] == remc.JOB_DONE
        _check_properties(s, g)
    st = srv.stats()
    srv.close()
    assert st["completed"] == 400 and st["failed"] == 0
    assert 0.0 < st["p50_latency"] <= st["p99_latency"] <= st["max_latency"]
    assert st["running_peak"] > 1
