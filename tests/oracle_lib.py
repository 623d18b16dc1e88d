"""ctypes access to the CPU oracle (oracle/liboracle.so) for the tests.

The oracle is what the tests compare against; nothing in the product imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from gpudock_amd import remc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "liboracle.so")

_L = None


def lib():
    global _L
    if _L is None:
        if not os.path.exists(ORACLE_SO):
            subprocess.run(["make", "-s", "-C", ORACLE_DIR, "all"], check=True)
        L = C.CDLL(ORACLE_SO)
        cx6 = [C.POINTER(remc.Ligand), C.POINTER(remc.Protein), C.POINTER(remc.Psp),
               C.POINTER(remc.Kde), C.POINTER(remc.Mcs), C.POINTER(remc.EnePara)]
        FP, IP = remc.FP, remc.IP
        L.orc_pose_ref32.argtypes = cx6 + [remc.ComplexSize, C.c_int, C.c_int, FP, C.c_int, FP]
        L.orc_energy_ref64.argtypes = cx6 + [remc.ComplexSize, C.c_int, C.c_int, FP, C.c_int,
                                             C.POINTER(C.c_double)]
        L.orc_move_ref32.argtypes = [FP, FP, FP, FP, C.c_int, FP, FP, FP, FP, FP]
        L.orc_canon_move.argtypes = [FP, FP, FP, FP, C.c_int, FP, FP, FP, FP, FP]
        L.orc_replay_energy.argtypes = cx6 + [remc.ComplexSize, C.c_int, IP, IP, FP, C.c_int, FP]
        L.orc_replay_run.argtypes = cx6 + [C.POINTER(remc.Temp), C.POINTER(remc.Replica),
                                           C.POINTER(remc.McPara), remc.ComplexSize,
                                           C.POINTER(remc.Options),
                                           C.POINTER(remc.ReplicaResult)]
        L.orc_replay_run_rec.argtypes = L.orc_replay_run.argtypes + [C.c_int, C.c_void_p,
                                                                     C.c_int]
        L.orc_ref_run.argtypes = cx6 + [C.POINTER(remc.Temp), remc.ComplexSize, FP, C.c_int,
                                        C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_uint32,
                                        C.c_int, C.c_int, IP, IP, IP, FP, FP]
        L.orc_ref_run_samples.argtypes = cx6 + [C.POINTER(remc.Temp), remc.ComplexSize, FP,
                                                C.c_int, C.c_int, C.c_int, C.c_int,
                                                C.c_uint32, C.c_uint32, C.c_int, C.c_int, IP,
                                                IP, FP, IP, C.c_int, C.c_int]
        L.orc_accept.argtypes = [C.c_int, C.c_float, C.c_float, C.c_float, C.c_float]
        L.orc_metropolis.argtypes = [C.c_float] * 5 + [C.c_int, C.c_int]
        L.orc_exchange_chain.argtypes = [C.c_int, C.c_int, FP, IP, FP, C.c_int, FP, IP]
        L.orc_schedule.argtypes = [C.c_int, C.c_int, C.c_int, IP, C.c_int]
        L.orc_bayes_pdf.argtypes = [C.c_int, C.c_int, FP, C.c_float, C.c_float, FP]
        L.orc_bayes_total.argtypes = [FP]
        L.orc_bayes_total.restype = C.c_float
        L.orc_philox.argtypes = [C.c_uint32] * 6 + [C.POINTER(C.c_uint32)]
        L.orc_canon_math.argtypes = [C.c_int, C.c_int, FP, FP, FP]
        _L = L
    return _L


def _fp(a):
    return a.ctypes.data_as(remc.FP)


def _ip(a):
    return a.ctypes.data_as(remc.IP)


def pose_ref32(cx, lig_conf, prt_conf, moves, term_order=0):
    moves = np.ascontiguousarray(moves, np.float32).reshape(-1, 6)
    out = np.zeros((len(moves), 10), np.float32)
    for i, m in enumerate(moves):
        m = np.ascontiguousarray(m)
        o = out[i]
        lib().orc_pose_ref32(*cx.args(), cx.size, int(lig_conf[i]), int(prt_conf[i]),
                             _fp(m), term_order, _fp(o))
    return out


def pose_ref64(cx, lig_conf, prt_conf, moves, term_order=0):
    moves = np.ascontiguousarray(moves, np.float32).reshape(-1, 6)
    out = np.zeros((len(moves), 10), np.float64)
    for i, m in enumerate(moves):
        m = np.ascontiguousarray(m)
        o = np.zeros(10, np.float64)
        lib().orc_energy_ref64(*cx.args(), cx.size, int(lig_conf[i]), int(prt_conf[i]),
                               _fp(m), term_order, o.ctypes.data_as(C.POINTER(C.c_double)))
        out[i] = o
    return out


def replay_energy(cx, lig_conf, prt_conf, moves, term_order=0):
    lig_conf = np.ascontiguousarray(lig_conf, np.int32)
    prt_conf = np.ascontiguousarray(prt_conf, np.int32)
    moves = np.ascontiguousarray(moves, np.float32).reshape(-1, 6)
    out = np.zeros((len(moves), 10), np.float32)
    lib().orc_replay_energy(*cx.args(), cx.size, len(moves), _ip(lig_conf), _ip(prt_conf),
                            _fp(moves), term_order, _fp(out))
    return out


def steps_run(mc):
    """MC steps the schedule of McPara mc runs (whole windows of whole
    launches, gpusingle.cu:8-20)."""
    return len(schedule(mc.steps_total, mc.steps_per_dump, mc.steps_per_exchange)) * \
        mc.steps_per_exchange


def replay_run_records(cx, temps, mc, opt, stride):
    """Replay plus the pre-move state of every stride-th step of the run:
    (results, records [n_rep, ceil(steps run / stride)])."""
    size = cx.sized(len(temps))
    res = (remc.ReplicaResult * size.n_rep)()
    n_rec = (steps_run(mc) + stride - 1) // stride
    rec = np.zeros((size.n_rep, n_rec), remc.STEP_RECORD_DTYPE)
    lib().orc_replay_run_rec(*cx.args(), temps, None, C.byref(mc), size, C.byref(opt), res,
                             stride, rec.ctypes.data_as(C.c_void_p), n_rec)
    return remc.results_to_numpy(res), rec


def replay_run(cx, temps, mc, opt):
    size = cx.sized(len(temps))
    res = (remc.ReplicaResult * size.n_rep)()
    lib().orc_replay_run(*cx.args(), temps, None, C.byref(mc), size, C.byref(opt), res)
    return remc.results_to_numpy(res)


def ref_run(cx, temps, scale, n_steps, spe, exchange=1, seed=1, cid=0, term_order=0,
            threads=0, spd=None):
    """Reference-order run; spd (steps per dump window) None = n_steps."""
    size = cx.sized(len(temps))
    n = size.n_rep
    acs = np.zeros(n, np.int32)
    acx = np.zeros(n, np.int32)
    it = np.zeros(n, np.int32)
    e = np.zeros((n, 10), np.float32)
    mv = np.zeros((n, 6), np.float32)
    sc = np.ascontiguousarray(scale, np.float32)
    spd = spd or max(n_steps, 1)
    rc = lib().orc_ref_run(*cx.args(), temps, size, _fp(sc), n_steps, spd, spe, exchange,
                           seed, cid, term_order, threads, _ip(acs), _ip(acx), _ip(it), _fp(e),
                           _fp(mv))
    assert rc == 0
    return dict(acs_mc=acs, acs_exchg=acx, idx_tmp_final=it, energy_final=e, move_final=mv)


def ref_run_samples(cx, temps, scale, n_steps, spe, stride, exchange=0, seed=1, cid=0,
                    term_order=0, threads=0, spd=None):
    """Reference-order run sampled every `stride` steps: (acs_mc [n_rep],
    energies [n_rep, n_samp], temperature index [n_rep, n_samp], acs_exchg [n_rep])."""
    size = cx.sized(len(temps))
    n = size.n_rep
    n_samp = (n_steps + stride - 1) // stride
    acs = np.zeros(n, np.int32)
    acx = np.zeros(n, np.int32)
    es = np.zeros((n, n_samp), np.float32)
    ts = np.zeros((n, n_samp), np.int32)
    sc = np.ascontiguousarray(scale, np.float32)
    spd = spd or max(n_steps, 1)
    rc = lib().orc_ref_run_samples(*cx.args(), temps, size, _fp(sc), n_steps, spd, spe,
                                   exchange, seed, cid, term_order, threads, _ip(acs),
                                   _ip(acx), _fp(es), _ip(ts), stride, n_samp)
    assert rc == 0
    return acs, es, ts, acx


def accept(e_new, e_old, minus_beta, u, canonical):
    """The oracle's Metropolis decision (Accept_d, accept_d.cu:22-23)."""
    return bool(lib().orc_accept(int(canonical), e_new, e_old, minus_beta, u))


def metropolis(e_new, e_old, minus_beta, u, u2, fixed_draws=0, reverse=0):
    """The canonical decision with the reference's draw toggles (IS_RANDOM,
    IS_REVERSE; accept_d.cu:17-27)."""
    return bool(lib().orc_metropolis(e_new, e_old, minus_beta, u, u2, int(fixed_draws),
                                     int(reverse)))


def exchange_chain(minus_beta, temps, ener, mode, u, canonical):
    """The oracle's exchange of one temperature chain with the draws u[pair]:
    (temperature index held by each slot after it, swaps per slot)."""
    mb = np.ascontiguousarray(minus_beta, np.float32)
    t = np.array(temps, np.int32)
    e = np.ascontiguousarray(ener, np.float32)
    uu = np.ascontiguousarray(u, np.float32)
    acs = np.zeros(len(t), np.int32)
    lib().orc_exchange_chain(int(canonical), len(t), _fp(mb), _ip(t), _fp(e), int(mode),
                             _fp(uu), _ip(acs))
    return t, acs


def schedule(steps_total, spd, spe):
    """The launch sequence of the oracle's run loops: exchange mode of each
    launch (launch x runs MC steps [x spe, (x + 1) spe)); None if invalid."""
    n = lib().orc_schedule(steps_total, spd, spe, None, 0)
    if n < 0:
        return None
    modes = np.zeros(max(n, 1), np.int32)
    lib().orc_schedule(steps_total, spd, spe, _ip(modes), n)
    return modes[:n]


def philox(ctr, key):
    out = (C.c_uint32 * 4)()
    lib().orc_philox(*[int(v) for v in ctr], *[int(v) for v in key], out)
    return list(out)


def canon_math(fn, x):
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros_like(x)
    y2 = np.zeros_like(x)
    lib().orc_canon_math(fn, len(x), _fp(x), _fp(y), _fp(y2))
    return (y, y2) if fn == 2 else y


def bayes_pdf(family, x, loc, scale):
    x = np.ascontiguousarray(x, np.float32)
    y = np.zeros_like(x)
    lib().orc_bayes_pdf(family, len(x), _fp(x), loc, scale, _fp(y))
    return y


def bayes_total(norm_file):
    v = np.ascontiguousarray(norm_file, np.float32)
    return lib().orc_bayes_total(_fp(v))
