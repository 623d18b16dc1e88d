"""Parity of the HIP path (through the C ABI) with the oracle.

Per pose: GPU FP32 energies equal the canonical replay bit for bit; every
term is within 1e-5 of the FP64 oracle relative to max(|E64|, 1); its
distance to the reference-order FP32 restatement is within that same bar
plus the restatement's own FP32 rounding |E32 - E64| (the reference order
sums up to ~10^5 pair terms sequentially and is the less accurate of the
two); the GPU FP64 re-score matches the oracle's FP64 within 1e-9 relative.
Per run: with the same Philox stream, every accept and exchange decision
and every final energy equals the replay's (bitwise)."""
import math
import os

import numpy as np
import pytest

import fixtures
import oracle_lib as O
from gpudock_amd import remc, synth

pytestmark = pytest.mark.gpu

REL = 1e-5   # FP32 per-term tolerance against the reference-order oracle


def _poses(rng, n, n_lig, n_prt, tsig=0.7, rsig=1.0):
    lc = rng.integers(0, n_lig, n).astype(np.int32)
    pc = rng.integers(0, n_prt, n).astype(np.int32)
    mv = np.concatenate([rng.normal(0, tsig, (n, 3)), rng.normal(0, rsig, (n, 3))], 1)
    return lc, pc, mv.astype(np.float32)


def tolerance(cx, r64, term_order):
    """Per-term bar: REL relative to max(|E64|, 1), plus the FP32 quantisation
    of the coordinates themselves (one ulp of the largest coordinate
    magnitude, times the normalisation slope the term receives) -- the
    reference's FP32 has exactly that floor too (e.g. its ligand centre is
    the float sum centre + translation, move_d.cu:130-132)."""
    R = float(np.max(np.abs(cx.prt_xyz))) + 10.0
    q = 2.0 ** -23 * R
    a = np.ctypeslib.as_array(cx.enepara.a_para)[:9].astype(np.float64)
    w = np.ctypeslib.as_array(cx.enepara.w)[:9].astype(np.float64)
    slope = np.zeros(10)
    for k in range(9):
        slot = remc.FILE_TO_SLOT[k] if term_order == 0 else k
        par = k
        slope[slot] = abs(a[par])
    wslot = np.zeros(9)
    for k in range(9):
        wslot[remc.FILE_TO_SLOT[k] if term_order == 0 else k] = abs(w[k])
    slope[9] = float(np.sum(wslot * slope[:9]))
    return REL * np.maximum(np.abs(r64), 1.0) + 4.0 * q * slope


def _check_pose_parity(cx, lc, pc, mv, term_order=0):
    e, e64 = remc.score_poses(cx, lc, pc, mv, term_order=term_order)
    rp = O.replay_energy(cx, lc, pc, mv, term_order=term_order)
    np.testing.assert_array_equal(e, rp)               # canonical: bit-exact
    r32 = O.pose_ref32(cx, lc, pc, mv, term_order=term_order)
    r64 = O.pose_ref64(cx, lc, pc, mv, term_order=term_order)
    tol = tolerance(cx, r64, term_order)
    assert np.all(np.abs(e - r64) <= tol)
    assert np.all(np.abs(e - r32) <= tol + np.abs(r32 - r64))
    np.testing.assert_allclose(e64, r64, rtol=1e-9, atol=1e-9)
    return e, r32, r64


def test_g1_on_gpu():
    cx, _ = fixtures.load_1a07c1()
    e, e64 = remc.score_poses(cx, [0], [0], np.zeros((1, 6)))
    np.testing.assert_allclose(e[0], fixtures.g1_slots(), rtol=0, atol=5.001e-5)
    np.testing.assert_allclose(e64[0], fixtures.g1_slots(), rtol=0, atol=1.5e-4)


def test_pose_parity_1a07c1():
    cx, _ = fixtures.load_1a07c1()
    rng = np.random.default_rng(10)
    lc, pc, mv = _poses(rng, 200, 21, 3, tsig=0.5, rsig=0.5)
    _check_pose_parity(cx, lc, pc, mv)
    _check_pose_parity(cx, lc, pc, mv, term_order=remc.TERM_ORDER_SLOT)


def test_pose_parity_c0():
    cx, temps, mc, cfg = synth.make_config("c0")
    rng = np.random.default_rng(11)
    lc, pc, mv = _poses(rng, 300, 4, 2)
    _check_pose_parity(cx, lc, pc, mv)


@pytest.mark.parametrize("lna,pnp,pnk,pos", [
    (1, 1, 0, 0), (1, 64, 5, 1), (63, 65, 64, 3), (64, 63, 1, 11), (65, 129, 130, 2),
    (128, 200, 700, 11), (24, 1000, 0, 0), (7, 300, 300, 0), (64, 150, 600, 1),
    (127, 100, 3000, 0)])
def test_pose_parity_edge_sizes(lna, pnp, pnk, pos):
    cx = synth.make_complex(pnp=pnp, radius=max(6.0, 3.0 * lna ** (1 / 3) + 4),
                            lna=lna, n_lig=2, n_prt=2, pnk=pnk, pos=pos, seed=lna + pnp)
    rng = np.random.default_rng(lna * 1000 + pnp)
    lc, pc, mv = _poses(rng, 40, 2, 2, tsig=0.4, rsig=0.6)
    _check_pose_parity(cx, lc, pc, mv)


def test_pose_parity_c1_shape():
    """The 16,384-replica complex (pnp 800, lna 48, 2000 KDE points), a sample of
    poses across all 64 ligand x 16 protein conformers."""
    cx, temps, mc, cfg = synth.make_config("c1")
    rng = np.random.default_rng(12)
    lc, pc, mv = _poses(rng, 256, 64, 16)
    _check_pose_parity(cx, lc, pc, mv)


def test_pose_parity_c4_shape():
    """The large-pocket complex (pnp 4000, lna 128, 5000 KDE points): KDE does not
    fit LDS beside the pocket and is read from HBM."""
    cx, temps, mc, cfg = synth.make_config("c4", n_lig=8)
    rng = np.random.default_rng(13)
    lc, pc, mv = _poses(rng, 24, 8, 1, tsig=0.3, rsig=0.3)
    _check_pose_parity(cx, lc, pc, mv)


def _compare_runs(g, r):
    for k in ("move_final", "energy_final", "move_best", "energy_best", "step_best",
              "idx_tmp_final", "acs_mc", "acs_exchg", "cmcc_final", "cmcc_best"):
        np.testing.assert_array_equal(g[k], r[k], err_msg=k)


def test_run_matches_replay_c0_full():
    """The small plain-REMC workload in full: 8 T x 4 ligand x 2 protein conformers, 1000 steps,
    exchange every 10: every decision equals the replay's."""
    cx, temps, mc, cfg = synth.make_config("c0")
    opt = remc.default_options(seed=1)
    g, log = remc.run(cx, temps, mc, opt)
    r = O.replay_run(cx, temps, mc, opt)
    _compare_runs(g, r)
    assert g["acs_mc"].sum() > 0 and g["acs_exchg"].sum() > 0
    assert log.ac_mc == g["acs_mc"].sum()
    # clocks as the reference takes them: t0 the MC loop, t1 = t0 without
    # trajectory dumps, t2 the whole call (setup, upload, re-score, copies)
    assert 0.0 < log.t0 <= log.t1 < log.t2
    # best-pose FP64 re-score equals the oracle's FP64 of the same pose
    n_lig, n_tmp = 4, 8
    idx = np.arange(len(g["acs_mc"]))
    lc, pc = idx % n_lig, idx // (n_tmp * n_lig)
    r64 = O.pose_ref64(cx, lc, pc, g["move_best"])
    np.testing.assert_allclose(g["energy_best_fp64"], r64, rtol=1e-9, atol=1e-9)
    assert np.all(np.abs(g["energy_best"] - r64) <= tolerance(cx, r64, 0))


def test_decision_parity_1e4_steps():
    """10^4 MC steps on 8 temperatures with exchange every 10 steps: the GPU
    trajectory (accept counts, exchange counts, final state bits) equals the
    CPU replay's on the same Philox stream, split over several wavefronts per
    replica and with one wavefront per replica."""
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=1, n_prt=1, steps=10000)
    opt = remc.default_options(seed=1)
    g, _ = remc.run(cx, temps, mc, opt)
    r = O.replay_run(cx, temps, mc, opt)
    _compare_runs(g, r)
    assert g["acs_exchg"].sum() > 0
    # the same run with one wavefront per replica (gd_run splits these 8
    # replicas over several wavefronts by itself)
    res, rep = remc.batch_run([remc.JobSpec(cx, temps, mc, opt)], split=1)
    assert rep["layout"]["split"] == 1
    _compare_runs(res[0], r)


@pytest.mark.skipif(os.environ.get("GD_LONG_PARITY") != "1",
                    reason="10^5 steps: about 80 s of CPU replay; set GD_LONG_PARITY=1")
def test_decision_parity_1e5_steps():
    """10^5 MC steps on 8 temperatures with exchange every 10 steps (10^4
    exchange sweeps): every accept and exchange decision and the final state
    bits equal the CPU replay's."""
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=1, n_prt=1, steps=100000)
    opt = remc.default_options(seed=3)
    g, _ = remc.run(cx, temps, mc, opt)
    r = O.replay_run(cx, temps, mc, opt)
    _compare_runs(g, r)
    assert g["acs_exchg"].sum() > 0


@pytest.mark.parametrize("n_tmp,exchange,away,term_order,steps,spe", [
    (1, 1, 0, 0, 57, 10), (5, 1, 0, 0, 200, 7), (4, 0, 1, 0, 100, 10),
    (6, 1, 1, 1, 120, 3), (3, 1, 0, 1, 33, 1)])
def test_run_variants(n_tmp, exchange, away, term_order, steps, spe):
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=2, n_prt=2, n_tmp=n_tmp,
                                           steps=steps, spe=spe)
    opt = remc.default_options(seed=3, complex_id=5, exchange=exchange, away=away,
                               term_order=term_order)
    g, _ = remc.run(cx, temps, mc, opt)
    r = O.replay_run(cx, temps, mc, opt)
    _compare_runs(g, r)


@pytest.mark.parametrize("fixed_draws,reverse", [(1, 0), (0, 1), (1, 1)])
def test_draw_toggles(fixed_draws, reverse):
    """The reference's IS_RANDOM 0 (every draw 20: fixed moves, every step
    accepted, swaps only above probability 20) and IS_REVERSE (the second
    draw against the sign-flipped exponent) toggles (toggle.h:9,24): every
    decision equals the replay's, one wavefront per replica or split."""
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=2, n_prt=2, n_tmp=4, steps=60)
    opt = remc.default_options(seed=8, fixed_draws=fixed_draws, reverse=reverse)
    r = O.replay_run(cx, temps, mc, opt)
    g, _ = remc.run(cx, temps, mc, opt)
    _compare_runs(g, r)
    res, rep = remc.batch_run([remc.JobSpec(cx, temps, mc, opt)], split=1)
    assert rep["layout"]["split"] == 1
    _compare_runs(res[0], r)
    if fixed_draws:
        assert np.all(g["acs_mc"] == 60)


@pytest.mark.parametrize("n_lig,n_prt,n_tmp", [(4, 8, 7), (8, 5, 4), (17, 3, 2)])
def test_exchange_many_chains(n_lig, n_prt, n_tmp):
    """32 chains are exchanged by one workgroup, more by many (the last to
    finish advancing the counters): both forms make the replay's decisions."""
    cx, temps, mc, cfg = synth.make_config("c0", lna=8, pnp=60, pnk=40, n_lig=n_lig,
                                           n_prt=n_prt, n_tmp=n_tmp, steps=80, spe=4)
    opt = remc.default_options(seed=21, complex_id=3)
    g, _ = remc.run(cx, temps, mc, opt)
    r = O.replay_run(cx, temps, mc, opt)
    _compare_runs(g, r)
    assert g["acs_exchg"].sum() > 0


@pytest.mark.parametrize("steps, spd, spe, n_tmp", [
    (100, 50, 10, 4),    # five launches per window: the mode restarts in each window
    (60, 25, 10, 5),     # spe does not divide the window: 3 windows x 3 launches, 90 steps
    (45, 20, 20, 3),     # the last, partial window runs whole: 60 steps
])
def test_reference_schedule_windows(steps, spd, spe, n_tmp):
    """The reference's launch loop (gpusingle.cu:8-20): whole steps_per_dump
    windows of whole launches, the exchange mode restarting at 0 in each;
    every decision equals the replay's, and the MC steps run are those of
    the schedule."""
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=2, n_prt=1, n_tmp=n_tmp, steps=steps,
                                           spe=spe)
    mc.steps_per_dump = spd
    opt = remc.default_options(seed=13, complex_id=2)
    g, _ = remc.run(cx, temps, mc, opt)
    r = O.replay_run(cx, temps, mc, opt)
    _compare_runs(g, r)
    s = remc.Session(cx, temps, mc, opt)
    n_launch = len(O.schedule(steps, spd, spe))
    s.cycles(n_launch + 3)          # cycles past the end do nothing
    s.sync()
    assert s.steps_done() == n_launch * spe
    h, _ = s.results()
    s.close()
    _compare_runs(h, r)


def test_session_equals_run():
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=2, n_prt=1, steps=60)
    opt = remc.default_options(seed=9)
    g, _ = remc.run(cx, temps, mc, opt)
    s = remc.Session(cx, temps, mc, opt)
    s.init()
    s.cycles(2)
    s.cycles(4)
    s.sync()
    h, _ = s.results()
    s.close()
    _compare_runs(g, h)


def test_acceptance_rises_with_temperature():
    """MC acceptance grows with T (README.org:295-346 records the same trend)."""
    cx, temps, mc, cfg = synth.make_config("c0", steps=400)
    opt = remc.default_options(seed=2, exchange=0)
    g, _ = remc.run(cx, temps, mc, opt)
    n_lig, n_prt, n_tmp = 4, 2, 8
    acs = g["acs_mc"].reshape(n_prt, n_tmp, n_lig).sum(axis=(0, 2))
    assert acs[-1] > acs[0]
    assert np.corrcoef(np.arange(n_tmp), acs)[0, 1] > 0.8


@pytest.mark.parametrize("term_order", [0, 1])
def test_bayes_total_run_matches_replay(term_order):
    """IS_BAYE total: every decision equals the replay's; the FP64 re-score
    of each best pose carries the float64 Bayesian total of its FP64
    normalised terms."""
    import test_bayes as TB
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=2, n_prt=2, n_tmp=4, steps=150)
    opt = remc.default_options(seed=6, combine=1, term_order=term_order)
    g, _ = remc.run(cx, temps, mc, opt)
    r = O.replay_run(cx, temps, mc, opt)
    _compare_runs(g, r)
    slot_of_file = [0, 1, 2, 5, 4, 8, 3, 6, 7]
    for e64 in g["energy_best_fp64"]:
        ref = TB.bayes64(e64[slot_of_file])
        assert abs(e64[9] - ref) <= 1e-9 * max(1.0, abs(ref))


def test_g1_and_run_from_input_files():
    """1a07C1 read by the library's own loaders from its SD/PDB/LHM files:
    the published initial energy line on the GPU, then a short REMC run over
    all 21 ligand x 3 protein conformers that equals the replay."""
    from test_loaders import load_1a07c1_files
    cx = load_1a07c1_files()
    e, e64 = remc.score_poses(cx, [0], [0], np.zeros((1, 6)))
    np.testing.assert_allclose(e[0], fixtures.g1_slots(), rtol=0, atol=5.001e-5)
    temps = remc.temperatures(2, 0.044, 0.036)
    mc = remc.mcpara(60, 10, [0.02, 0.02, 0.02, 0.08, 0.08, 0.08])
    opt = remc.default_options(seed=5, mcc=1)
    g, _ = remc.run(cx, temps, mc, opt)
    r = O.replay_run(cx, temps, mc, opt)
    _compare_runs(g, r)


def test_c4_trajectory_poses_against_fp64(tmp_path):
    """The large-pocket stress shape (pocket 4,000, ligand 128, 5,000 KDE points,
    32 temperatures; 8 of the 128 ligand conformers): the kernel's FP32 energies
    of every 100th pose along the trajectories, per term, against the GPU FP64
    re-score of the same poses."""
    from gpudock_amd import traj
    cx, temps, mc, cfg = synth.make_config("c4", n_lig=8, steps=300)
    mc.steps_per_dump = 300
    prefix = str(tmp_path / "c4")
    opt = remc.default_options(seed=3, record_stride=100, record_prefix=prefix)
    remc.run(cx, temps, mc, opt)
    _, rec = traj.read_gdtraj(prefix + "_0000.gdtraj")
    rec = rec.reshape(-1)
    lc = rec["idx_lig"].astype(np.int32)
    pc = rec["idx_prt"].astype(np.int32)
    mv = np.ascontiguousarray(rec["movematrix"], np.float32)
    e32 = rec["energy"][:, :10]
    e, e64 = remc.score_poses(cx, lc, pc, mv)
    np.testing.assert_array_equal(e, e32)      # the same kernel arithmetic re-run
    tol = tolerance(cx, e64, 0)
    assert np.all(np.abs(e32 - e64) <= tol), np.max(np.abs(e32 - e64) / tol)
    assert len(rec) == 8 * 32 * 3


def test_whole_protein_pocket_read_through_caches():
    """A pocket of 7,000 points (168 KB of points, more than LDS holds beside
    the table rows) is read through the caches: poses and a short run equal
    the replay."""
    cx = synth.make_complex(pnp=7000, radius=None, lna=20, n_lig=2, n_prt=2, pnk=300, pos=2,
                            seed=21)
    rng = np.random.default_rng(21)
    lc, pc, mv = _poses(rng, 24, 2, 2, tsig=0.4, rsig=0.6)
    _check_pose_parity(cx, lc, pc, mv)
    temps = remc.temperatures(2, 0.3, 3.0)
    mc = remc.mcpara(30, 10, [0.2, 0.2, 0.2, 0.05, 0.05, 0.05])
    opt = remc.default_options(seed=4, mcc=1)
    s = remc.Session(cx, temps, mc, opt)
    s.init()
    assert s.layout()["lds_bytes"] < 7000 * 24
    s.close()
    g, _ = remc.run(cx, temps, mc, opt)
    r = O.replay_run(cx, temps, mc, opt)
    _compare_runs(g, r)


def test_reference_maximum_complex():
    """The reference's largest complex (gpudocksm-rem-1.2/src/size.h:22-79):
    10,000 protein points (MAXPRO), 100 ligand atoms (MAXLIG), 10,000 KDE
    points (MAXKDE) and 1,000 MCS restraints (MAXPOS) of 100 fields each
    (MAXMCS 256): poses equal the replay bit for bit and FP64 within the
    tolerance; a short run, in gd_run's own layout and with one wavefront per
    replica, equals the replay."""
    cx = synth.make_complex(pnp=10000, radius=None, lna=100, n_lig=2, n_prt=2, pnk=10000,
                            pos=1000, seed=31)
    rng = np.random.default_rng(31)
    lc, pc, mv = _poses(rng, 16, 2, 2, tsig=0.4, rsig=0.6)
    _check_pose_parity(cx, lc, pc, mv)
    temps = remc.temperatures(2, 0.3, 3.0)
    mc = remc.mcpara(20, 10, [0.2, 0.2, 0.2, 0.05, 0.05, 0.05])
    opt = remc.default_options(seed=6, mcc=1)
    r = O.replay_run(cx, temps, mc, opt)
    g, _ = remc.run(cx, temps, mc, opt)
    _compare_runs(g, r)
    res, rep = remc.batch_run([remc.JobSpec(cx, temps, mc, opt)], split=1)
    assert rep["layout"]["split"] == 1
    _compare_runs(res[0], r)


@pytest.mark.parametrize("n_tmp,n_lig,n_prt", [(40, 17, 3), (1, 70, 20)])
def test_reference_maximum_replicas(n_tmp, n_lig, n_prt):
    """The reference's replica limits (gpudocksm-rem-1.2/src/size.h:28-40):
    40 temperatures (MAXTMP) on 2,040 of at most 2,048 replicas (MAXREP), and
    70 ligand x 20 protein conformers (MAXEN2, MAXEN1) at one temperature;
    exchange every 5 steps: every decision equals the replay's."""
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=n_lig, n_prt=n_prt, n_tmp=n_tmp,
                                           steps=30, spe=5)
    opt = remc.default_options(seed=7)
    g, _ = remc.run(cx, temps, mc, opt)
    r = O.replay_run(cx, temps, mc, opt)
    _compare_runs(g, r)
    assert len(g["acs_mc"]) == n_tmp * n_lig * n_prt
    if n_tmp > 1:
        assert g["acs_exchg"].sum() > 0


def test_class2_points_outside_the_hydrophobic_class():
    """Class-2 points (the only ones with a PSP value) whose hydrophobic flag
    is clear sort after the hydrophobic chunks, so the chunks that read PSP
    run past them (the loaders never build such input: class 2 implies the
    hydrophobic class).  Poses and a short run equal the replay."""
    cx = synth.make_complex(pnp=300, radius=12.0, lna=20, n_lig=2, n_prt=2, pnk=200, pos=2,
                            seed=23)
    pc, pj = np.nonzero(cx.prt_c == 2)
    assert len(pc) > 40 and np.all(cx.prt_hpcflag[pc, pj] != 0)
    cx.prt_hpcflag[pc[::2], pj[::2]] = 0          # the struct points at this array
    rng = np.random.default_rng(23)
    lc, pcf, mv = _poses(rng, 64, 2, 2, tsig=0.4, rsig=0.6)
    _check_pose_parity(cx, lc, pcf, mv)
    temps = remc.temperatures(2, 0.3, 3.0)
    mc = remc.mcpara(30, 10, [0.2, 0.2, 0.2, 0.05, 0.05, 0.05])
    opt = remc.default_options(seed=5, mcc=1)
    g, _ = remc.run(cx, temps, mc, opt)
    r = O.replay_run(cx, temps, mc, opt)
    _compare_runs(g, r)


def test_g2_trajectory_vector():
    """Golden vector G2 (gpudocksm-rem-1.2/reference_result.txt:60-80): the
    GPU's energies of the ten poses of the reference's deterministic path on
    1a07C1 equal the canonical replay bit for bit and the reference's printed
    raw terms and total within 1e-6 of max(|E|, 1) each."""
    steps, mv, g2 = fixtures.load_g2()
    cx = fixtures.load_1a07c1_raw()
    z = np.zeros(len(steps), np.int32)
    e, e64 = remc.score_poses(cx, z, z, mv)
    np.testing.assert_array_equal(e, O.replay_energy(cx, z, z, mv))
    assert np.all(np.abs(e - g2) <= 1e-6 * np.maximum(np.abs(g2), 1.0))
    floor = 2.0 * 2.0 ** -23 * float(np.max(np.abs(cx.lig_center)))
    tol = 1e-6 * np.maximum(np.abs(g2), 1.0)
    tol[:, 8] += floor
    assert np.all(np.abs(e64 - g2) <= tol)

