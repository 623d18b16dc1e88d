"""GPU per-step records: every record equals the replay's, in the files
gd_run writes per steps_per_dump window and through gd_session_records."""
import numpy as np
import pytest

import oracle_lib as O
from gpudock_amd import remc, synth, traj

pytestmark = pytest.mark.gpu


def _cfg(steps, spe, spd, n_tmp=4):
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=2, n_prt=2, n_tmp=n_tmp, steps=steps,
                                           spe=spe)
    mc.steps_per_dump = spd
    return cx, temps, mc


@pytest.mark.parametrize("stride,steps,spe,spd", [(1, 120, 10, 50), (5, 100, 5, 25),
                                                   (2, 37, 1, 8)])
def test_record_files_equal_replay(tmp_path, stride, steps, spe, spd):
    cx, temps, mc = _cfg(steps, spe, spd)
    prefix = str(tmp_path / "a")
    opt = remc.default_options(seed=11, record_stride=stride, record_prefix=prefix)
    g, _ = remc.run(cx, temps, mc, opt)
    r, rec = O.replay_run_records(cx, temps, mc, remc.default_options(seed=11), stride)
    for k in ("move_final", "energy_final", "acs_mc", "acs_exchg", "idx_tmp_final"):
        np.testing.assert_array_equal(g[k], r[k], err_msg=k)
    n_files = (steps + spd - 1) // spd
    got = []
    for w in range(n_files):
        h, part = traj.read_gdtraj(f"{prefix}_{w:04d}.gdtraj")
        assert h["first_step"] == w * spd and h["stride"] == stride
        got.append(part)
    allrec = np.concatenate(got, axis=1)
    assert allrec.shape == rec.shape
    np.testing.assert_array_equal(allrec, rec)


def test_session_records_window():
    cx, temps, mc = _cfg(60, 10, 20)
    opt = remc.default_options(seed=5, record_stride=1)
    s = remc.Session(cx, temps, mc, opt)
    s.cycles(3)          # steps 0..29: the second window holds steps 20..29
    first, part = s.records()
    s.close()
    _, rec = O.replay_run_records(cx, temps, mc, remc.default_options(seed=5), 1)
    assert first == 20 and part.shape[1] == 10
    np.testing.assert_array_equal(part, rec[:, 20:30])


def test_record_window_must_align():
    cx, temps, mc = _cfg(60, 10, 25)
    with pytest.raises(remc.GDError):
        remc.run(cx, temps, mc, remc.default_options(record_stride=1))


@pytest.mark.parametrize("n_prt,away", [(1, 0), (2, 0), (2, 1)])
def test_mcc_equals_replay(tmp_path, n_prt, away):
    """MCC per pose (one ballot per chunk against the reference mask on the
    GPU, per-pair counts on the CPU): final/best values and every record."""
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=2, n_prt=n_prt, n_tmp=4, steps=60,
                                           spe=10)
    mc.steps_per_dump = 20
    prefix = str(tmp_path / "m")
    opt = remc.default_options(seed=21, mcc=1, away=away, record_stride=1,
                               record_prefix=prefix)
    g, _ = remc.run(cx, temps, mc, opt)
    r, rec = O.replay_run_records(cx, temps, mc,
                                  remc.default_options(seed=21, mcc=1, away=away), 1)
    for k in ("cmcc_final", "cmcc_best", "energy_final", "move_final", "acs_mc"):
        np.testing.assert_array_equal(g[k], r[k], err_msg=k)
    allrec = np.concatenate([traj.read_gdtraj(f"{prefix}_{w:04d}.gdtraj")[1]
                             for w in range(3)], axis=1)
    np.testing.assert_array_equal(allrec, rec)
    assert np.any(rec["energy"][..., 10] != 0.0)
