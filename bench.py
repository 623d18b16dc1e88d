"""Benchmark: REMC docking throughput on one MI355X.

A "step" of this benchmark is one Monte Carlo cycle of the whole replica
ensemble: steps_per_exchange (10) MC steps (Move + CalcEnergy + Accept) of
every replica in one fused kernel launch, followed by one temperature-
exchange launch -- the unit the reference's driver loop repeats
(gpusingle.cu:14-20).  Default workload: one synthetic complex with pocket
800 points (14 A sphere), ligand 48 atoms, 64 ligand x 16 protein
conformers x 16 temperatures = 16,384 replicas, 2,000 KDE points, exchange
every 10 steps; 500 cycles are the full 5,000-step run.

Prints one JSON line: pose energy evals/s (= MC replica-steps/s), the
roofline of the MC kernel (FP32 VALU), and the reference-order CPU oracle
timed on this host on a bounded sample of the same complex.
--workload batch (256 uneven complexes in one engine) and --workload serve
(Poisson job arrivals into the serving loop, p50/p99 latency) print their
own lines.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X vector FP32 (published)


def flops_per_pose(cx):
    """Algorithmic FP32 work of one pose evaluation (each sqrt/exp/div
    counted as one op): 60*lna*pnp + 12*M_kde + 10*pos*lna + 8*lna + 40,
    M_kde = type-matched (ligand atom, KDE point) pairs."""
    lna = cx.lig_xyz.shape[1]
    pnp = cx.prt_xyz.shape[1]
    pos = cx.mcs_xyz.shape[0]
    counts = np.bincount(cx.kde_t, minlength=24)
    m_kde = int(counts[cx.lig_t].sum())
    return 60 * lna * pnp + 12 * m_kde + 10 * pos * lna + 8 * lna + 40, m_kde


def cpu_baseline(cx, temps_all, scale, cfg, seed):
    """Reference-order FP32 oracle (the cpu_mic loop order, libm math) on
    this host: 1 thread and all cores, on a bounded sample of replicas x
    steps of the same complex.  Returns pose evals/s."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from gpudock_amd import remc, synth

    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    out = {}
    # about 10 s of CPU work each: 1 thread ~2.6e3 and 16 threads ~3.6e4
    # pose evals/s on this complex
    for threads, n_lig, steps in ((1, 8, 800), (cores, 64, 1500 * cores // 16)):
        sub = synth.subset(cx, n_lig, 1)
        t_sub = remc.temperatures(4, cfg["t_lo"], cfg["t_hi"])
        n_rep = n_lig * 4
        t0 = time.perf_counter()
        O.ref_run(sub, t_sub, scale, steps, cfg["spe"], exchange=1, seed=seed,
                  threads=threads)
        dt = time.perf_counter() - t0
        out[threads] = (n_rep * steps / dt, n_rep, steps, dt)
    return out, cores


def copy_calibration(mib):
    """One hipMemcpyDeviceToDevice of `mib` MiB: a plain copy kernel with a
    known byte count (reads mib MiB, writes mib MiB), to scale the HBM
    traffic counters."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), C.c_size_t(n)) == 0
    assert hip.hipMalloc(C.byref(b), C.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, C.c_size_t(n)) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, C.c_size_t(n), 3) == 0   # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"copy_calibration_bytes": n}))


def bench_batch(args):
    """Batched uneven workload: 256 complexes (ligand atoms U[12, 96], pocket
    points U[200, 1500], ligand conformers U[4, 64], 8|16 temperatures, 2,000
    steps) resident together, launch shape autotuned, cycles replayed from
    one captured hipGraph.  Timed region: the cycles, inputs resident in
    HBM.  Beside it, the plain path (gd_run, one complex per call) on every
    16th complex."""
    from gpudock_amd import remc, synth

    t0 = time.perf_counter()
    jobs = synth.make_batch_jobs(n=args.batch_n, seed=args.seed, steps=args.batch_steps)
    gen_s = time.perf_counter() - t0
    poses = sum(j.n_rep * j.mc.steps_total for j in jobs)
    flops = sum(flops_per_pose(j.cx)[0] * j.n_rep * j.mc.steps_total for j in jobs)
    pairs = sum(j.n_rep * j.mc.steps_total * j.size.lna * j.size.pnp for j in jobs)
    print(f"batch: {len(jobs)} complexes, {sum(j.n_rep for j in jobs)} replicas, "
          f"{poses:.3e} pose evals (generated in {gen_s:.1f} s)", file=sys.stderr, flush=True)
    res, rep = remc.batch_run(jobs, autotune=0 if args.no_tune else 1)
    value = poses / rep["run_s"]
    achieved = flops / rep["run_s"] / 1e12
    print(f"batch: run {rep['run_s']:.2f} s", file=sys.stderr, flush=True)
    plain_poses, plain_s = 0, 0.0
    for j in jobs[::16]:
        t1 = time.perf_counter()
        remc.run(j.cx, j.temps, j.mc, j.opt)
        plain_s += time.perf_counter() - t1
        plain_poses += j.n_rep * j.mc.steps_total
    plain = plain_poses / plain_s
    acc = float(sum(r["acs_mc"].sum() for r in res)) / poses
    out = {
        "metric": "pose energy evals/s (MC replica-steps/s)",
        "value": value,
        "unit": "pose evals/s",
        "n_gpus": 1,
        "steps": rep["cycles"],
        "warmup": 0,
        "ms_per_step": rep["run_s"] * 1e3 / max(1, rep["cycles"]),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, reference parameter files)",
        "config": {
            "workload": f"batched uneven: {len(jobs)} complexes, ligand atoms U[12,96], "
                        f"pocket points U[200,1500], ligand conformers U[4,64], 8|16 "
                        f"temperatures, 1 protein conformer, {args.batch_steps} steps, "
                        "exchange every 10",
            "replicas": sum(j.n_rep for j in jobs), "complexes": len(jobs),
            "parallelism": "single GPU, one engine",
        },
        "pair_interactions_per_s": pairs / rep["run_s"],
        "end_to_end_s": rep["admit_s"] + rep["tune_s"] + rep["run_s"] + rep["results_s"],
        "phases_s": {k: rep[k] for k in ("admit_s", "tune_s", "run_s", "results_s")},
        "plain_path": {"value": plain, "unit": "pose evals/s",
                       "sample": f"gd_run on {len(jobs[::16])} of the complexes (every 16th), "
                                 "one call each, wall time"},
        "speedup_vs_plain": value / plain,
        "mc_acceptance": acc,
        "autotune": rep["tune"],
        "layout": rep["layout"],
        "roofline": {
            "bound": "valu", "kernel": "gd_mc_kernel", "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": None,
            "note": "achieved = algorithmic FLOPs of all complexes / wall time of the "
                    "graph-replayed cycles (MC + exchange + advance launches)",
        },
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def bench_serve(args):
    """Serving loop: Poisson arrivals at --rate jobs/s for --duration s, each
    job a complex drawn from a pool of 256 (ligand atoms lognormal median 32
    in [10, 120], pocket points lognormal median 600 in [150, 2000]) with 16
    ligand conformers x 8 temperatures x 1,000 steps and its own Philox key.
    Then a burst of --burst jobs submitted at once measures the sustainable
    completion rate."""
    import numpy as np
    from gpudock_amd import remc, synth

    pool = synth.make_serving_pool(n=256, seed=args.seed + 1, steps=args.serve_steps)
    rng = np.random.default_rng(args.seed)
    gaps = rng.exponential(1.0 / args.rate, size=int(args.rate * args.duration * 1.5) + 16)
    arrivals = np.cumsum(gaps)
    arrivals = arrivals[arrivals < args.duration]
    picks = rng.integers(0, len(pool), size=len(arrivals) + args.burst)
    def job(i):
        cx, temps, mc = pool[picks[i]]
        return remc.JobSpec(cx, temps, mc, remc.default_options(seed=args.seed, complex_id=i))
    specs = [job(i) for i in range(len(arrivals))]
    fp = max(s.footprint() for s in specs[:64])
    srv = remc.Server(arena_bytes=args.arena_gib << 30, wpg=args.wpg,
                      persistent=args.persistent, keep_results=0, occ=args.occ,
                      kde_policy=args.kde_policy)
    print(f"serve: {len(specs)} jobs at {args.rate}/s over {args.duration} s", file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    behind = 0.0
    next_report = 10.0
    for i, (t, sp) in enumerate(zip(arrivals, specs)):
        now = time.perf_counter() - t0
        if t > now:
            time.sleep(t - now)
        else:
            behind = max(behind, now - t)
        srv.submit(sp)
        if t >= next_report:
            st = srv.stats()
            print(f"serve: t={t:.0f}s submitted {st['submitted']} completed {st['completed']} "
                  f"running {st['running']} queued {st['queued']}", file=sys.stderr, flush=True)
            next_report += 10.0
    srv.drain()
    st = srv.stats()
    wall = time.perf_counter() - t0
    srv.close()
    # burst: sustainable completion rate
    burst = None
    if args.burst:
        bs = [job(len(arrivals) + k) for k in range(args.burst)]
        srv = remc.Server(arena_bytes=args.arena_gib << 30, wpg=args.wpg,
                          persistent=args.persistent, keep_results=0, occ=args.occ,
                          kde_policy=args.kde_policy)
        tb = time.perf_counter()
        for sp in bs:
            srv.submit(sp)
        srv.drain()
        sb = srv.stats()
        srv.close()
        span = sb["last_done"] - sb["first_submit"]
        burst = {"jobs": args.burst, "jobs_per_s": args.burst / span,
                 "pose_evals_per_s": sb["pose_evals_per_s"], "wall_s": time.perf_counter() - tb}
        print(f"serve: burst {burst}", file=sys.stderr, flush=True)
    pairs_per_job = float(np.mean([pool[k][0].lig_xyz.shape[1] * pool[k][0].prt_xyz.shape[1]
                                   for k in picks[:len(arrivals)]])) * 16 * 8 * args.serve_steps
    out = {
        "metric": "pose energy evals/s (MC replica-steps/s); p50/p99 docking-job latency",
        "value": st["pose_evals_per_s"],
        "unit": "pose evals/s",
        "n_gpus": 1,
        "steps": int(st["completed"]),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, reference parameter files)",
        "config": {
            "workload": f"serving: Poisson arrivals {args.rate} jobs/s for {args.duration} s; "
                        "ligand atoms lognormal(median 32, sigma 0.5) in [10,120], pocket "
                        "points lognormal(median 600, sigma 0.5) in [150,2000]; 16 ligand "
                        f"conformers x 8 temperatures x {args.serve_steps} steps per job",
            "jobs": len(specs), "rate_jobs_per_s": args.rate, "duration_s": args.duration,
            "parallelism": "single GPU, continuous batching",
        },
        "latency_s": {"p50": st["p50_latency"], "p99": st["p99_latency"],
                      "mean": st["mean_latency"], "max": st["max_latency"],
                      "queue_wait_p50": st["p50_wait"], "queue_wait_p99": st["p99_wait"]},
        "completed": st["completed"], "failed": st["failed"],
        "offered_pairs_per_s": pairs_per_job * args.rate,
        "running_peak": st["running_peak"], "arena_peak_bytes": st["arena_peak"],
        "cycles": st["cycles"], "rebuilds": st["rebuilds"],
        "submit_lag_max_s": behind, "wall_s": wall, "job_footprint_max_bytes": fp,
        "burst": burst,
        "server": {"wpg": args.wpg, "persistent": args.persistent, "occ": args.occ,
                   "kde_policy": args.kde_policy},
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500, help="timed MC cycles")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c1")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-host-path", action="store_true",
                    help="skip the end-to-end gd_run (host buffers) line")
    ap.add_argument("--mcc", action="store_true",
                    help="compute the per-pose MCC contact score (IS_CALCU_MCC)")
    ap.add_argument("--no-tune", action="store_true",
                    help="keep the default launch shape (8 waves per workgroup, grid form)")
    ap.add_argument("--workload", choices=("single", "batch", "serve"), default="single",
                    help="single: one complex (--config, default the 16,384-replica "
                         "complex); batch: 256 uneven complexes; serve: request loop")
    ap.add_argument("--batch-n", type=int, default=256)
    ap.add_argument("--batch-steps", type=int, default=2000)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--duration", type=float, default=120.0)
    ap.add_argument("--serve-steps", type=int, default=1000)
    ap.add_argument("--burst", type=int, default=1000)
    ap.add_argument("--wpg", type=int, default=8)
    ap.add_argument("--persistent", type=int, default=1)
    ap.add_argument("--arena-gib", type=int, default=16)
    ap.add_argument("--occ", type=int, default=6,
                    help="serving: MC kernel register budget (0, 6 or 8 waves per SIMD)")
    ap.add_argument("--kde-policy", type=int, default=1,
                    help="serving: 1 = KDE points read from HBM, 0 = staged in LDS when they fit")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (calibrates "
                         "the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (--gpus {args.gpus} "
              "requested); the REMC chains shard by whole temperature ladders, "
              "see the sharding notes", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        bench_batch(args)
        return
    if args.workload == "serve":
        bench_serve(args)
        return

    from gpudock_amd import remc, synth

    cx, temps, mc, cfg = synth.make_config(args.config, seed=args.seed)
    n_rep = cfg["n_lig"] * cfg["n_prt"] * cfg["n_tmp"]
    spe = cfg["spe"]
    # every warm-up and timed cycle must do a full spe steps
    mc.steps_total = max(mc.steps_total, (args.warmup + args.steps) * spe)
    f_pose, m_kde = flops_per_pose(cx)
    opt = remc.default_options(seed=args.seed, mcc=int(args.mcc))
    job = remc.JobSpec(cx, temps, mc, opt)
    eng = remc.Engine(arena_bytes=job.footprint() + (64 << 20), use_graph=1)
    h = eng.add(job)
    tune = eng.autotune() if not args.no_tune else None
    if args.warmup:
        eng.cycles(args.warmup)
    eng.sync()
    eng.set_timing(True)
    mc0, n0, x0, nx0 = eng.timing()
    t0 = time.perf_counter()
    eng.cycles(args.steps)
    eng.sync()
    elapsed = time.perf_counter() - t0
    mc_ms, mc_n, x_ms, x_n = eng.timing()
    mc_ms, mc_n, x_ms, x_n = mc_ms - mc0, mc_n - n0, x_ms - x0, x_n - nx0
    layout = eng.layout()
    res, log = eng.results(h)
    eng.close()

    poses = n_rep * spe * args.steps
    value = poses / elapsed
    traffic = None
    valu_per_launch = None
    tfile = os.path.join(ROOT, "profiles", "gd_mc_kernel_16384rep_traffic.json")
    if args.config == "c1" and os.path.exists(tfile):
        # HBM bytes per cycle launch from the rocprofv3 FETCH_SIZE / WRITE_SIZE
        # passes (calibrated against a device copy), and the VALU instructions
        # per launch (SQ_INSTS_VALU), kept with the profile
        with open(tfile) as f:
            prof = json.load(f)
        traffic = prof["bytes_per_launch"]
        valu_per_launch = prof.get("valu_insts_per_launch")
    # the device's FP32 VALU issue ceiling, measured now (about a millisecond)
    probe_ns, probe_tflops, n_simd = remc.valu_probe(8)
    valu_issue = {"probe_ns_per_wave_inst": probe_ns, "probe_fma_tflops": probe_tflops,
                  "n_simd": n_simd}
    if valu_per_launch and mc_n:
        kern_ns = (mc_ms / mc_n) * 1e6 / (valu_per_launch / n_simd)
        valu_issue.update(kernel_valu_insts_per_launch=valu_per_launch,
                          kernel_ns_per_wave_inst=kern_ns, issue_frac=probe_ns / kern_ns)
    avg_mc_s = (mc_ms / mc_n) / 1e3 if mc_n else float("nan")
    achieved = f_pose * n_rep * spe / avg_mc_s / 1e12
    out = {
        "metric": "pose energy evals/s (MC replica-steps/s)",
        "value": value,
        "unit": "pose evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, reference parameter files)",
        "config": {
            "workload": f"one complex, pocket {cfg['pnp']} points "
                        f"({cfg['radius']} A sphere), ligand {cfg['lna']} atoms, "
                        f"{cfg['n_lig']} lig x {cfg['n_prt']} prt conformers x "
                        f"{cfg['n_tmp']} temperatures = {n_rep} replicas, "
                        f"{cfg['pnk']} KDE points, exchange every {spe} steps",
            "replicas": n_rep, "lna": cfg["lna"], "pnp": cfg["pnp"], "pnk": cfg["pnk"],
            "steps_per_exchange": spe, "mc_steps_timed": spe * args.steps,
            "mcc": bool(args.mcc),
            "parallelism": "single GPU",
        },
        "pair_interactions_per_s": value * cfg["lna"] * cfg["pnp"],
        "mc_acceptance": float(res["acs_mc"].sum()) / max(1, n_rep * spe * (args.steps + args.warmup)),
        "roofline": {
            "bound": "valu",
            "kernel": "gd_mc_kernel",
            "achieved": achieved,
            "peak": PEAK_FP32_TFLOPS,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32_TFLOPS,
            "traffic": traffic,
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE/WRITE_SIZE, "
                              "profiles/gd_mc_kernel_16384rep.md" if traffic else None,
            "hbm_gbps_at_traffic": traffic / (mc_ms / mc_n) / 1e6 if traffic and mc_n else None,
            "flops_per_pose": f_pose,
            "poses_per_launch": n_rep * spe,
            "avg_launch_ms": mc_ms / mc_n if mc_n else None,
            "exchange_avg_ms": x_ms / x_n if x_n else None,
            "lds_bytes": layout["stage_bytes"],
            "kde_in_lds": layout["kde_in_lds"],
            "waves_per_group": layout["wpg"],
            "persistent": layout["persistent"],
            "grid": layout["grid"],
            "valu_issue": valu_issue,
        },
        "cpu_baseline": None,
        "autotune": tune,
    }
    if not args.no_host_path:
        # the drop-in path end to end: gd_run from host buffers (upload, initial
        # energies, every cycle, FP64 re-score, results back to the host), the
        # default launch shape; t0 / t1 as the reference reports them (run.cu)
        mc_full = remc.mcpara(cfg["steps"], spe, list(mc.move_scale))
        _, rlog = remc.run(cx, temps, mc_full, opt)
        full = n_rep * cfg["steps"]
        out["host_path"] = {
            "entry": "gd_run", "mc_steps": cfg["steps"],
            "t0_s": rlog.t0, "t1_s": rlog.t1,
            "pose_evals_per_s_t0": full / rlog.t0, "pose_evals_per_s_t1": full / rlog.t1,
            "note": "t0 = upload + cycles (PCIe-inclusive inputs); t1 = t0 + FP64 re-score "
                    "and read-back of results; not the headline value",
        }
    if not args.no_cpu:
        scale = list(mc.move_scale)
        base, cores = cpu_baseline(cx, temps, scale, cfg, args.seed)
        v1 = base[1]
        vn = base[cores]
        out["cpu_baseline"] = {
            "value": vn[0], "unit": "pose evals/s", "cores": cores, "kind": "port",
            "sample": f"reference-order FP32 oracle, same complex (protein conformer 0), {vn[1]} replicas "
                      f"(4 T) x {vn[2]} MC steps on {cores} threads, {vn[3]:.1f} s",
            "single_thread_value": v1[0],
            "single_thread_sample": f"{v1[1]} replicas x {v1[2]} steps, {v1[3]:.1f} s",
        }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
