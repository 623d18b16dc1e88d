"""MCC contact scoring (CalcMcc_d, calcmcc_d.cu:148-213) in the replay:
closed forms at the reference pose and a float64 brute force."""
import numpy as np

import oracle_lib as O
from gpudock_amd import remc, synth


def _contacts64(cx, lig_conf, prt_conf, move):
    """Contact matrix [lna, pnp] (d <= pmf0) of a pose, in float64."""
    m = np.asarray(move, np.float64)
    s1, c1 = np.sin(m[3]), np.cos(m[3])
    s2, c2 = np.sin(m[4]), np.cos(m[4])
    s3, c3 = np.sin(m[5]), np.cos(m[5])
    R = np.array([[c1 * c2, c1 * s2 * s3 - c3 * s1, s1 * s3 + c1 * c3 * s2],
                  [c2 * s1, c1 * c3 + s1 * s2 * s3, c3 * s1 * s2 - c1 * s3],
                  [-s2, c2 * s3, c2 * c3]])
    x = cx.lig_xyz[lig_conf].astype(np.float64) @ R.T + m[:3] + cx.lig_center
    p = cx.prt_xyz[prt_conf].astype(np.float64)
    d = np.linalg.norm(x[:, None, :] - p[None, :, :], axis=2)
    pmf0 = np.array(cx.enepara.pmf0)[cx.lig_t][:, cx.prt_t[prt_conf]]
    return d <= pmf0


def _mcc(now, ref):
    tp = float(np.sum(now & ref)); fn = float(np.sum(~now & ref))
    fp = float(np.sum(now & ~ref)); tn = float(np.sum(~now & ~ref))
    den = np.sqrt((tp + fp) * (tp + fn) * (tn + fp) * (tn + fn))
    return (tp * tn - fp * fn) / den if den else 10000.0


def test_mcc_is_one_at_the_reference_pose():
    """Replicas that start from replica 0's ligand and protein conformers at
    movematrix 0 reproduce the reference contacts exactly: MCC = 1."""
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=2, n_prt=2, n_tmp=3, steps=0)
    r = O.replay_run(cx, temps, mc, remc.default_options(mcc=1))
    n_lig, n_tmp = 2, 3
    idx = np.arange(len(r["cmcc_final"]))
    same = (idx % n_lig == 0) & (idx // (n_tmp * n_lig) == 0)
    assert np.all(r["cmcc_final"][same] == 1.0)
    assert np.all(r["cmcc_final"][~same] < 1.0)


def test_mcc_matches_float64_brute_force():
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=2, n_prt=2, n_tmp=2, steps=30,
                                           spe=5)
    opt = remc.default_options(seed=3, mcc=1)
    res, rec = O.replay_run_records(cx, temps, mc, opt, 1)
    ref = _contacts64(cx, 0, 0, np.zeros(6))
    n_lig, n_tmp = 2, 2
    checked = 0
    for r in range(rec.shape[0]):
        lc, pc = r % n_lig, r // (n_tmp * n_lig)
        for k in (0, 7, 29):
            now = _contacts64(cx, lc, pc, rec["movematrix"][r, k])
            assert abs(rec["energy"][r, k, 10] - _mcc(now, ref)) < 1e-6
            checked += 1
    assert checked == rec.shape[0] * 3
    # without the option the field stays 0
    res0, rec0 = O.replay_run_records(cx, temps, mc, remc.default_options(seed=3), 1)
    assert np.all(rec0["energy"][..., 10] == 0.0) and np.all(res0["cmcc_final"] == 0.0)
    np.testing.assert_array_equal(res0["energy_final"], res["energy_final"])
