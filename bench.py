"""Benchmark: REMC docking throughput on one MI355X.

A "step" of this benchmark is one Monte Carlo cycle of the whole replica
ensemble: steps_per_exchange (10) MC steps (Move + CalcEnergy + Accept) of
every replica in one fused kernel launch, followed by one temperature-
exchange launch -- the unit the reference's driver loop repeats
(gpusingle.cu:14-20).  Default workload: configs[1] of the benchmark
definition, one synthetic complex with pocket 800 points (14 A sphere),
ligand 48 atoms, 64 ligand x 16 protein conformers x 16 temperatures =
16,384 replicas, 2,000 KDE points, exchange every 10 steps; 500 cycles are
the full 5,000-step run.

Prints one JSON line: pose energy evals/s (= MC replica-steps/s), the
roofline of the MC kernel (FP32 VALU), and the reference-order CPU oracle
timed on this host on a bounded sample of the same complex.
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
    for threads, n_lig, steps in ((1, 8, 40), (cores, 64, 40)):
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500, help="timed MC cycles")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--config", default="c1")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run one device-to-device copy of MIB MiB (calibrates "
                         "the FETCH_SIZE / WRITE_SIZE counters) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU only (--gpus {args.gpus} "
              "requested); the REMC chains shard by whole temperature ladders, "
              "see the sharding notes", file=sys.stderr)
        sys.exit(2)

    from gpudock_amd import remc, synth

    cx, temps, mc, cfg = synth.make_config(args.config, seed=args.seed)
    n_rep = cfg["n_lig"] * cfg["n_prt"] * cfg["n_tmp"]
    spe = cfg["spe"]
    f_pose, m_kde = flops_per_pose(cx)
    opt = remc.default_options(seed=args.seed)
    s = remc.Session(cx, temps, mc, opt)
    s.init()
    if args.warmup:
        s.cycles(args.warmup)
    s.sync()
    s.set_timing(True)
    s.timing()  # drain
    mc0, n0, x0, nx0 = s.timing()
    t0 = time.perf_counter()
    s.cycles(args.steps)
    s.sync()
    elapsed = time.perf_counter() - t0
    mc_ms, mc_n, x_ms, x_n = s.timing()
    mc_ms, mc_n, x_ms, x_n = mc_ms - mc0, mc_n - n0, x_ms - x0, x_n - nx0
    layout = s.layout()
    res, log = s.results()
    s.close()

    poses = n_rep * spe * args.steps
    value = poses / elapsed
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
            "workload": f"{args.config}: one complex, pocket {cfg['pnp']} points "
                        f"({cfg['radius']} A sphere), ligand {cfg['lna']} atoms, "
                        f"{cfg['n_lig']} lig x {cfg['n_prt']} prt conformers x "
                        f"{cfg['n_tmp']} temperatures = {n_rep} replicas, "
                        f"{cfg['pnk']} KDE points, exchange every {spe} steps",
            "replicas": n_rep, "lna": cfg["lna"], "pnp": cfg["pnp"], "pnk": cfg["pnk"],
            "steps_per_exchange": spe, "mc_steps_timed": spe * args.steps,
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
            "traffic": None,
            "flops_per_pose": f_pose,
            "poses_per_launch": n_rep * spe,
            "avg_launch_ms": mc_ms / mc_n if mc_n else None,
            "exchange_avg_ms": x_ms / x_n if x_n else None,
            "lds_bytes": layout["lds_bytes"],
            "kde_in_lds": layout["kde_in_lds"],
        },
        "cpu_baseline": None,
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
