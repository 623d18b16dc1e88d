"""Per-step trajectory records (RecordLigand_d, util_d.cu:72-94) on the CPU
side: the replay's records against closed-form consistency rules, and the
.gdtraj reader."""
import numpy as np
import pytest

import oracle_lib as O
from gpudock_amd import remc, synth, traj


def _small(steps=60, spe=10, n_tmp=4):
    return synth.make_config("c0", n_lig=2, n_prt=1, n_tmp=n_tmp, steps=steps, spe=spe)


def test_replay_records_are_the_pre_move_states():
    """Stride 1: the record of step s is the state before step s; the state
    changes between consecutive records exactly when a move was accepted, so
    the number of changes (plus the change into the final state) equals the
    acceptance count; temperatures change only at exchange boundaries."""
    cx, temps, mc, cfg = _small()
    opt = remc.default_options(seed=4)
    res, rec = O.replay_run_records(cx, temps, mc, opt, 1)
    n_rep, n = rec.shape
    assert n == mc.steps_total
    assert np.all(rec["step"] == np.arange(n)[None, :])
    assert np.all(rec["idx_rep"] == np.arange(n_rep)[:, None])
    # step 0 holds the initial energies: a zero-step replay gives them
    mc0 = remc.mcpara(0, cfg["spe"], list(mc.move_scale))
    init = O.replay_run(cx, temps, mc0, opt)
    np.testing.assert_array_equal(rec["energy"][:, 0, :10], init["energy_final"])
    moved = np.any(rec["movematrix"][:, 1:] != rec["movematrix"][:, :-1], axis=2).sum(axis=1)
    moved += np.any(res["move_final"] != rec["movematrix"][:, -1], axis=1)
    np.testing.assert_array_equal(moved, res["acs_mc"])
    change = np.nonzero(np.any(rec["idx_tmp"][:, 1:] != rec["idx_tmp"][:, :-1], axis=0))[0] + 1
    assert np.all(change % cfg["spe"] == 0)
    assert np.all(rec["energy"][..., 10] == 0.0)


def test_replay_records_stride_subsamples():
    cx, temps, mc, cfg = _small()
    opt = remc.default_options(seed=4)
    _, r1 = O.replay_run_records(cx, temps, mc, opt, 1)
    _, r3 = O.replay_run_records(cx, temps, mc, opt, 3)
    np.testing.assert_array_equal(r3, r1[:, ::3])


def test_gdtraj_roundtrip_and_checks(tmp_path):
    rec = np.zeros((3, 5), remc.STEP_RECORD_DTYPE)
    rec["step"] = np.arange(5)[None, :] * 2 + 10
    rec["energy"][..., 9] = np.random.default_rng(0).normal(size=(3, 5))
    p = tmp_path / "t_0000.gdtraj"
    traj.write_gdtraj(p, 10, 2, rec)
    h, back = traj.read_gdtraj(p)
    assert h == {"n_rep": 3, "n_step": 5, "first_step": 10, "stride": 2}
    np.testing.assert_array_equal(back, rec)
    assert remc.STEP_RECORD_DTYPE.itemsize == 88
    import ctypes
    assert ctypes.sizeof(remc.StepRecord) == 88
    bad = tmp_path / "bad.gdtraj"
    bad.write_bytes(b"NOTTRAJ!" + bytes(32))
    with pytest.raises(ValueError):
        traj.read_gdtraj(bad)
    short = tmp_path / "short.gdtraj"
    short.write_bytes(p.read_bytes()[:-8])
    with pytest.raises(ValueError):
        traj.read_gdtraj(short)
