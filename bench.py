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
roofline of the MC kernel (FP32 VALU), and the reference-order CPU
restatement built for and timed on this host on a bounded sample of the
same complex.  The same line carries two more legs (--legs): "batched", the
256 uneven complexes in one engine with fewer steps and their speed-up over
the plain one-complex-per-call path, and "serving", Poisson job arrivals at
200 jobs/s into the serving loop for 30 s with p50/p99 job latency.
--workload batch and --workload serve run those in full and print their own
lines (serve adds the sustainable completion rate and p50/p99 there).
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

# The serving lanes launch on a stream each; HIP maps a process's streams onto
# GPU_MAX_HW_QUEUES hardware queues (4 by default), and two lanes that share a
# queue run their launches in turn.  The runtime reads it when it starts, so it
# is raised here to 8, before anything touches the device (a larger value set
# by the caller stays).
_queues = os.environ.get("GPU_MAX_HW_QUEUES", "")
if not _queues.isdigit() or int(_queues) < 8:
    os.environ["GPU_MAX_HW_QUEUES"] = "8"

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


def host_cpu():
    """The host's CPU as the CPU baseline reports it: model, logical CPUs in
    the affinity mask, the thread share used (OMP_NUM_THREADS when set: the
    CPU share that goes with the GPU), the clock and the FP32 FLOPs per cycle
    per core assumed for the host's peak."""
    model, mhz, flags = "unknown", None, set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                k, _, v = line.partition(":")
                k = k.strip()
                if k == "model name" and model == "unknown":
                    model = v.strip()
                elif k == "cpu MHz" and mhz is None:
                    mhz = float(v)
                elif k == "flags" and not flags:
                    flags = set(v.split())
    except OSError:
        pass
    try:
        with open("/sys/devices/system/cpu/cpu0/cpufreq/cpuinfo_max_freq") as f:
            mhz = int(f.read()) / 1000.0
        clock_src = "cpuinfo_max_freq"
    except (OSError, ValueError):
        clock_src = "/proc/cpuinfo cpu MHz"
    try:
        visible = len(os.sched_getaffinity(0))
    except AttributeError:
        visible = os.cpu_count() or 1
    share = os.environ.get("OMP_NUM_THREADS")
    threads = max(1, min(visible, int(share))) if share and share.isdigit() else visible
    # FP32 FLOPs per cycle per core: two FMA pipes of the widest vector the
    # core executes at full rate (AVX-512 on Zen 5 and on Xeon with two
    # 512-bit FMA units, 256-bit otherwise)
    import re
    epyc = re.search(r"EPYC (\d{4})", model)
    zen4 = bool(epyc) and epyc.group(1)[-1] == "4"    # 512-bit ops on 256-bit pipes
    wide = "avx512f" in flags and not zen4
    fpc = 64 if wide else (32 if "fma" in flags else 16)
    return {"model": model, "cpus_visible": visible, "threads": threads,
            "thread_limit": "OMP_NUM_THREADS" if share else "affinity mask",
            "clock_ghz": (mhz or 0.0) / 1000.0, "clock_source": clock_src,
            "fp32_flops_per_cycle_per_core": fpc,
            "peak_note": "threads x clock x FLOPs/cycle (2 FMA pipes x vector width); "
                         "SMT threads counted as cores"}


def native_oracle_lib():
    """The reference-order CPU restatement (oracle/gd_oracle.c) compiled for
    this host: gcc -O3 -march=native, contraction off, no fast-math, OpenMP.
    Built into a temporary directory at run time, because the shipped build
    is portable; returns (path, flags) or (None, reason)."""
    import shutil
    import subprocess
    import tempfile
    gcc = shutil.which("gcc")
    if not gcc:
        return None, "no gcc on this host: portable -O2 build used"
    d = tempfile.mkdtemp(prefix="gd_cpu_")
    out = os.path.join(d, "liboracle_native.so")
    flags = ["-O3", "-march=native", "-std=c11", "-ffp-contract=off", "-fno-fast-math",
             "-fopenmp", "-fPIC", "-shared"]
    cmd = [gcc] + flags + ["-I", os.path.join(ROOT, "gpudock_amd", "csrc"),
                           "-I", os.path.join(ROOT, "include"),
                           os.path.join(ROOT, "oracle", "gd_oracle.c"), "-o", out, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return None, "native build failed: portable -O2 build used"
    return out, " ".join(flags)


def cpu_baseline(cx, temps_all, scale, cfg, seed, f_pose):
    """Reference-order FP32 restatement (the cpu_mic loop order, libm math)
    on this host, built for it: 1 thread and the host's thread share, each on
    a bounded sample of replicas x steps of the same complex (about 10 s of
    CPU work each).  Returns the cpu_baseline record."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O
    from gpudock_amd import remc, synth

    lib, build = native_oracle_lib()
    if lib:
        O.ORACLE_SO = lib
    info = host_cpu()
    threads = info["threads"]
    runs = {}
    # 1 thread ~2.6e3 and 16 threads ~3.6e4 pose evals/s on this complex
    for n_thr, n_lig, steps in ((1, 8, 800), (threads, 64, max(100, 1500 * threads // 16))):
        sub = synth.subset(cx, n_lig, 1)
        t_sub = remc.temperatures(4, cfg["t_lo"], cfg["t_hi"])
        n_rep = n_lig * 4
        t0 = time.perf_counter()
        O.ref_run(sub, t_sub, scale, steps, cfg["spe"], exchange=1, seed=seed, threads=n_thr)
        dt = time.perf_counter() - t0
        runs[n_thr] = (n_rep * steps / dt, n_rep, steps, dt)
    vn, v1 = runs[threads], runs[1]
    pairs = cfg["lna"] * cfg["pnp"]
    peak = threads * info["clock_ghz"] * 1e9 * info["fp32_flops_per_cycle_per_core"]
    return {
        "value": vn[0], "unit": "pose evals/s", "cores": threads, "kind": "port",
        "sample": f"reference-order FP32 restatement, same complex (protein conformer 0), "
                  f"{vn[1]} replicas (4 T) x {vn[2]} MC steps on {threads} threads, "
                  f"{vn[3]:.1f} s",
        "single_thread_value": v1[0],
        "single_thread_sample": f"{v1[1]} replicas x {v1[2]} steps, {v1[3]:.1f} s",
        "pair_interactions_per_s": vn[0] * pairs,
        "single_thread_pair_interactions_per_s": v1[0] * pairs,
        "algorithmic_tflops": vn[0] * f_pose / 1e12,
        "host_peak_tflops": peak / 1e12 if peak else None,
        "host_peak_frac": vn[0] * f_pose / peak if peak else None,
        "build": build if lib else build,
        "host": info,
    }


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


def batch_leg(n, steps, seed, tune=True, plain_every=16):
    """Batched uneven workload: n complexes (ligand atoms U[12, 96], pocket
    points U[200, 1500], ligand conformers U[4, 64], 8|16 temperatures,
    `steps` MC steps each) resident together, launch shape autotuned, cycles
    replayed from one captured hipGraph.  Timed: the cycles, inputs resident
    in HBM.  Beside it, the plain path (gd_run, one complex per call, wall
    time from host buffers) on every plain_every-th complex."""
    from gpudock_amd import remc, synth

    t0 = time.perf_counter()
    jobs = synth.make_batch_jobs(n=n, seed=seed, steps=steps)
    gen_s = time.perf_counter() - t0
    poses = sum(j.n_rep * j.mc.steps_total for j in jobs)
    flops = sum(flops_per_pose(j.cx)[0] * j.n_rep * j.mc.steps_total for j in jobs)
    pairs = sum(j.n_rep * j.mc.steps_total * j.size.lna * j.size.pnp for j in jobs)
    print(f"batch: {len(jobs)} complexes, {sum(j.n_rep for j in jobs)} replicas, "
          f"{poses:.3e} pose evals (generated in {gen_s:.1f} s)", file=sys.stderr, flush=True)
    res, rep = remc.batch_run(jobs, autotune=1 if tune else 0)
    value = poses / rep["run_s"]
    achieved = flops / rep["run_s"] / 1e12
    print(f"batch: run {rep['run_s']:.2f} s, {value:.3e} pose evals/s", file=sys.stderr,
          flush=True)
    plain_poses, plain_s = 0, 0.0
    for j in jobs[::plain_every]:
        t1 = time.perf_counter()
        remc.run(j.cx, j.temps, j.mc, j.opt)
        plain_s += time.perf_counter() - t1
        plain_poses += j.n_rep * j.mc.steps_total
    plain = plain_poses / plain_s
    print(f"batch: plain path {plain:.3e} pose evals/s", file=sys.stderr, flush=True)
    return {
        "value": value, "unit": "pose evals/s",
        "workload": f"batched uneven: {len(jobs)} complexes, ligand atoms U[12,96], pocket "
                    f"points U[200,1500], ligand conformers U[4,64], 8|16 temperatures, "
                    f"1 protein conformer, {steps} steps, exchange every 10",
        "complexes": len(jobs), "replicas": sum(j.n_rep for j in jobs), "mc_steps": steps,
        "cycles": rep["cycles"], "run_s": rep["run_s"],
        "pair_interactions_per_s": pairs / rep["run_s"],
        "end_to_end_s": rep["admit_s"] + rep["tune_s"] + rep["run_s"] + rep["results_s"],
        "phases_s": {k: rep[k] for k in ("admit_s", "tune_s", "run_s", "results_s")},
        "plain_path": {"value": plain, "unit": "pose evals/s",
                       "sample": f"gd_run on {len(jobs[::plain_every])} of the complexes "
                                 f"(every {plain_every}th), one call each, wall time"},
        "speedup_vs_plain": value / plain,
        "mc_acceptance": float(sum(r["acs_mc"].sum() for r in res)) / poses,
        "autotune": rep["tune"], "layout": rep["layout"],
        "achieved_tflops": achieved, "frac_of_fp32_peak": achieved / PEAK_FP32_TFLOPS,
    }


def bench_batch(args):
    leg = batch_leg(args.batch_n, args.batch_steps, args.seed, tune=not args.no_tune)
    out = {
        "metric": "pose energy evals/s (MC replica-steps/s)",
        "value": leg["value"],
        "unit": "pose evals/s",
        "n_gpus": 1,
        "steps": leg["cycles"],
        "warmup": 0,
        "ms_per_step": leg["run_s"] * 1e3 / max(1, leg["cycles"]),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, reference parameter files)",
        "config": {"workload": leg["workload"], "replicas": leg["replicas"],
                   "complexes": leg["complexes"], "parallelism": "single GPU, one engine"},
        "batched": leg,
        "roofline": {
            "bound": "valu", "kernel": "gd_mc_kernel", "achieved": leg["achieved_tflops"],
            "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": leg["frac_of_fp32_peak"],
            "traffic": None,
            "note": "achieved = algorithmic FLOPs of all complexes / wall time of the "
                    "graph-replayed cycles (MC + exchange + advance launches)",
        },
        "cpu_baseline": None,
    }
    print(json.dumps(out))


def serve_leg(rate, duration, steps, seed, wpg=16, persistent=0, occ=8, kde_policy=1,
              arena_gib=16, tag="serve", lanes=(), lane_split=(), lane_max_jobs=(),
              lane_prio=(), lane_max_total=()):
    """Serving loop: Poisson arrivals at `rate` jobs/s for `duration` s, each
    job a complex drawn from a pool of 256 (ligand atoms lognormal median 32
    in [10, 120], pocket points lognormal median 600 in [150, 2000]) with 16
    ligand conformers x 8 temperatures x `steps` steps and its own Philox
    key; the server drains after the last arrival."""
    import numpy as np
    from gpudock_amd import remc, synth

    pool = serve_leg.pools.get((seed, steps))
    if pool is None:
        pool = synth.make_serving_pool(n=256, seed=seed + 1, steps=steps)
        serve_leg.pools[(seed, steps)] = pool
    rng = np.random.default_rng(seed + int(rate))
    gaps = rng.exponential(1.0 / rate, size=int(rate * duration * 1.5) + 16)
    arrivals = np.cumsum(gaps)
    arrivals = arrivals[arrivals < duration]
    picks = rng.integers(0, len(pool), size=len(arrivals))
    specs = [remc.JobSpec(*pool[picks[i]], remc.default_options(seed=seed, complex_id=i))
             for i in range(len(arrivals))]
    srv = remc.Server(arena_bytes=arena_gib << 30, wpg=wpg, persistent=persistent,
                      keep_results=0, occ=occ, kde_policy=kde_policy, lane_max_pairs=lanes,
                      lane_split=lane_split, lane_max_jobs=lane_max_jobs, lane_prio=lane_prio,
                      lane_max_total=lane_max_total)
    print(f"{tag}: {len(specs)} jobs at {rate:.1f}/s over {duration} s", file=sys.stderr,
          flush=True)
    t0 = time.perf_counter()
    behind = 0.0
    next_report = 10.0
    for t, sp in zip(arrivals, specs):
        now = time.perf_counter() - t0
        if t > now:
            time.sleep(t - now)
        else:
            behind = max(behind, now - t)
        srv.submit(sp)
        if t >= next_report:
            st = srv.stats()
            print(f"{tag}: t={t:.0f}s submitted {st['submitted']} completed "
                  f"{st['completed']} running {st['running']} queued {st['queued']}",
                  file=sys.stderr, flush=True)
            next_report += 10.0
    srv.drain()
    st = srv.stats()
    wall = time.perf_counter() - t0
    srv.close()
    pairs_per_job = float(np.mean([pool[k][0].lig_xyz.shape[1] * pool[k][0].prt_xyz.shape[1]
                                   for k in picks])) * 16 * 8 * steps
    print(f"{tag}: p50 {st['p50_latency']:.3f} s, p99 {st['p99_latency']:.3f} s, "
          f"{st['pose_evals_per_s']:.3e} pose evals/s", file=sys.stderr, flush=True)
    return {
        "value": st["pose_evals_per_s"], "unit": "pose evals/s",
        "workload": f"serving: Poisson arrivals {rate:.1f} jobs/s for {duration} s; ligand "
                    "atoms lognormal(median 32, sigma 0.5) in [10,120], pocket points "
                    "lognormal(median 600, sigma 0.5) in [150,2000]; 16 ligand conformers x "
                    f"8 temperatures x {steps} steps per job",
        "rate_jobs_per_s": rate, "duration_s": duration, "jobs": len(specs),
        "latency_s": {"p50": st["p50_latency"], "p99": st["p99_latency"],
                      "mean": st["mean_latency"], "max": st["max_latency"],
                      "queue_wait_p50": st["p50_wait"], "queue_wait_p99": st["p99_wait"]},
        "completed": st["completed"], "failed": st["failed"],
        "offered_pairs_per_s": pairs_per_job * rate,
        "running_peak": st["running_peak"], "arena_peak_bytes": st["arena_peak"],
        "cycles": st["cycles"], "rebuilds": st["rebuilds"],
        "submit_lag_max_s": behind, "wall_s": wall,
        "server": {"wpg": wpg, "persistent": persistent, "occ": occ, "kde_policy": kde_policy,
                   "lane_max_pairs": list(lanes), "lane_split_asked": list(lane_split),
                   "lane_max_jobs": list(lane_max_jobs), "lane_prio": list(lane_prio),
                   "lane_max_total": list(lane_max_total),
                   "lane_split": st["lane_split"][:st["lanes"]],
                   "lane_cycles": st["lane_cycles"][:st["lanes"]],
                   "lane_completed": st["lane_completed"][:st["lanes"]]},
    }


serve_leg.pools = {}


def lane_bounds(text):
    """'a,b' -> (a, b): comma-separated integers (lane bounds or split factors)."""
    return tuple(int(x) for x in text.split(",") if x.strip())


def burst_rate(n, steps, seed, wpg=16, persistent=0, occ=8, kde_policy=1, arena_gib=16,
               lanes=(), lane_split=(), lane_max_jobs=(), lane_prio=(), lane_max_total=()):
    """n jobs submitted at once: the completion rate the server sustains."""
    from gpudock_amd import remc, synth
    import numpy as np

    pool = serve_leg.pools.get((seed, steps)) or synth.make_serving_pool(n=256, seed=seed + 1,
                                                                       steps=steps)
    picks = np.random.default_rng(seed + 7).integers(0, len(pool), size=n)
    specs = [remc.JobSpec(*pool[picks[i]], remc.default_options(seed=seed, complex_id=i))
             for i in range(n)]
    srv = remc.Server(arena_bytes=arena_gib << 30, wpg=wpg, persistent=persistent,
                      keep_results=0, occ=occ, kde_policy=kde_policy, lane_max_pairs=lanes,
                      lane_split=lane_split, lane_max_jobs=lane_max_jobs, lane_prio=lane_prio,
                      lane_max_total=lane_max_total)
    tb = time.perf_counter()
    for sp in specs:
        srv.submit(sp)
    srv.drain()
    sb = srv.stats()
    srv.close()
    span = sb["last_done"] - sb["first_submit"]
    out = {"jobs": n, "jobs_per_s": n / span, "pose_evals_per_s": sb["pose_evals_per_s"],
           "wall_s": time.perf_counter() - tb}
    print(f"serve: burst {out}", file=sys.stderr, flush=True)
    return out


def bench_serve(args):
    """The serving workload at --rate for --duration s; then a burst of
    --burst jobs gives the sustainable completion rate, and a Poisson stream
    at 95 % of it for --max-rate-duration s gives p50/p99 at that rate."""
    kw = dict(wpg=args.wpg, persistent=args.persistent, occ=args.occ,
              kde_policy=args.kde_policy, arena_gib=args.arena_gib,
              lanes=lane_bounds(args.lane_pairs), lane_split=lane_bounds(args.lane_split),
              lane_max_jobs=lane_bounds(args.lane_max_jobs),
              lane_prio=lane_bounds(args.lane_prio),
              lane_max_total=lane_bounds(args.lane_max_total))
    leg = serve_leg(args.rate, args.duration, args.serve_steps, args.seed, **kw)
    burst = burst_rate(args.burst, args.serve_steps, args.seed, **kw) if args.burst else None
    at_max = None
    if burst and args.max_rate_duration > 0:
        at_max = serve_leg(0.95 * burst["jobs_per_s"], args.max_rate_duration,
                           args.serve_steps, args.seed, tag="serve@max", **kw)
    out = {
        "metric": "pose energy evals/s (MC replica-steps/s); p50/p99 docking-job latency",
        "value": leg["value"],
        "unit": "pose evals/s",
        "n_gpus": 1,
        "steps": int(leg["completed"]),
        "warmup": 0,
        "ms_per_step": None,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded generator, reference parameter files)",
        "config": {"workload": leg["workload"], "jobs": leg["jobs"],
                   "rate_jobs_per_s": args.rate, "duration_s": args.duration,
                   "parallelism": "single GPU, continuous batching"},
        "serving": leg,
        "latency_s": leg["latency_s"],
        "burst": burst,
        "at_max_sustainable_rate": at_max,
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
    ap.add_argument("--max-rate-duration", type=float, default=60.0,
                    help="serve: seconds of arrivals at 95 %% of the burst completion rate")
    ap.add_argument("--legs", default="batch,serve",
                    help="single: extra legs in the same line (batch: the 256-complex "
                         "batch at --leg-batch-steps with its speed-up over the plain "
                         "path; serve: --leg-serve-duration s of arrivals at --rate); "
                         "'' for none")
    ap.add_argument("--leg-batch-steps", type=int, default=500)
    ap.add_argument("--split", type=int, default=0,
                    help="single: wavefronts per replica (0 = chosen from the replica count "
                         "and tuned, 1, 4, 8, 16)")
    ap.add_argument("--leg-serve-duration", type=float, default=30.0)
    # serving launch shape: 16 waves/group, grid form, register budget for 8
    # waves/SIMD served a 1,000-job burst 3 % faster than 8 waves/group,
    # persistent, 6 waves/SIMD, and halved the p50 latency at 200 jobs/s
    ap.add_argument("--wpg", type=int, default=16,
                    help="serving: wavefronts per workgroup (4, 8 or 16)")
    ap.add_argument("--persistent", type=int, default=0,
                    help="serving: 1 = workgroups pull tasks from a queue, 0 = one per task")
    ap.add_argument("--arena-gib", type=int, default=16)
    ap.add_argument("--occ", type=int, default=8,
                    help="serving: MC kernel register budget (0, 6 or 8 waves per SIMD)")
    ap.add_argument("--kde-policy", type=int, default=1,
                    help="serving: 1 = KDE points read from HBM, 0 = staged in LDS when they fit")
    # serving lanes: a split lane (16 wavefronts per replica) for jobs up to
    # 20,000 pairs per pose while the server holds fewer than 32 jobs (at most 4
    # in the lane), then plain lanes up to 20,000 and 40,000 pairs and one for
    # the rest.  p50 / p99 against one engine: 0.41 / 0.77 s against 0.66-0.68 /
    # 0.78-0.93 s at 200 jobs/s; 0.033 / 0.53 s against 0.40 / 0.54 s at 50 jobs/s
    ap.add_argument("--lane-pairs", default="20000,20001,40000",
                    help="serving: comma-separated upper bounds of ligand atoms x pocket "
                         "points for all lanes but the last (empty = one engine)")
    ap.add_argument("--lane-split", default="16,1,1,1",
                    help="serving: comma-separated wavefronts per replica of each lane "
                         "(4, 8 or 16 = split, KDE points in LDS; 0 = chosen from the resident count)")
    ap.add_argument("--lane-max-jobs", default="4,0,0,0",
                    help="serving: comma-separated jobs each lane holds (queued or resident) "
                         "before it passes new ones to the next lane (0 = no limit)")
    ap.add_argument("--lane-prio", default="",
                    help="serving: comma-separated issue priority (0-3) of each lane's "
                         "wavefronts")
    ap.add_argument("--lane-max-total", default="32,0,0,0",
                    help="serving: comma-separated jobs in the whole server above which "
                         "each lane passes new ones on (0 = no limit)")
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
    # kernel durations come from a separate event-timed run after the
    # throughput run (events around every launch cost the graph replay its
    # edge on short cycles)
    t_cycles = max(5, min(args.steps, 50))
    # every warm-up, timed and event-timed cycle must do a full spe steps
    mc.steps_total = max(mc.steps_total, (args.warmup + args.steps + t_cycles) * spe)
    f_pose, m_kde = flops_per_pose(cx)
    opt = remc.default_options(seed=args.seed, mcc=int(args.mcc))
    job = remc.JobSpec(cx, temps, mc, opt)
    eng = remc.Engine(arena_bytes=job.footprint() + (64 << 20), use_graph=1, split=args.split)
    h = eng.add(job)
    tune = eng.autotune() if not args.no_tune else None
    if args.warmup:
        eng.cycles(args.warmup)
    eng.sync()
    t0 = time.perf_counter()
    eng.cycles(args.steps)
    eng.sync()
    elapsed = time.perf_counter() - t0
    eng.set_timing(True)
    mc0, n0, x0, nx0 = eng.timing()
    eng.cycles(t_cycles)
    eng.sync()
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
    if args.config == "c1" and os.path.exists(tfile) and not args.mcc:
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
        "mc_acceptance": float(res["acs_mc"].sum()) / max(1, n_rep * spe * (args.steps + args.warmup
                                                                           + t_cycles)),
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
            "waves_per_replica": layout["split"],
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
        # default launch shape; t0 as the reference takes it (the MC loop,
        # gpusingle.cu:10,31), t2 the whole call
        mc_full = remc.mcpara(cfg["steps"], spe, list(mc.move_scale))
        _, rlog = remc.run(cx, temps, mc_full, opt)
        full = n_rep * cfg["steps"]
        out["host_path"] = {
            "entry": "gd_run", "mc_steps": cfg["steps"],
            "t0_s": rlog.t0, "t2_s": rlog.t2,
            "pose_evals_per_s_t0": full / rlog.t0, "pose_evals_per_s_t2": full / rlog.t2,
            "note": "t0 = the cycles (the reference's compute time); t2 = the whole call: "
                    "setup, upload (PCIe-inclusive inputs), initial energies, cycles, FP64 "
                    "re-score and read-back of results; not the headline value",
        }
    legs = [x for x in args.legs.split(",") if x]
    if "batch" in legs:
        out["batched"] = batch_leg(args.batch_n, args.leg_batch_steps, args.seed,
                                   tune=not args.no_tune)
    if "serve" in legs:
        out["serving"] = serve_leg(args.rate, args.leg_serve_duration, args.serve_steps,
                                   args.seed, wpg=args.wpg, persistent=args.persistent,
                                   occ=args.occ, kde_policy=args.kde_policy,
                                   arena_gib=args.arena_gib,
                                   lanes=lane_bounds(args.lane_pairs),
                                   lane_split=lane_bounds(args.lane_split),
                                   lane_max_jobs=lane_bounds(args.lane_max_jobs),
                                   lane_prio=lane_bounds(args.lane_prio),
                                   lane_max_total=lane_bounds(args.lane_max_total))
        out["latency_s"] = out["serving"]["latency_s"]
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(cx, temps, list(mc.move_scale), cfg, args.seed,
                                           f_pose)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
