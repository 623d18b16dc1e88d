"""Host preparation, parameter parsing and the C ABI surface (no GPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import fixtures
from gpudock_amd import remc, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "gd_remc.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(gd_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = remc.lib()
    names = _header_functions()
    assert "gd_run" in names and "gd_score_poses" in names
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_version_string():
    assert b"gfx950" in remc.lib().gd_version()


def test_enepara_tables_match_reference_prep():
    """gd_load_enepara + gd_load_weight + gd_load_norpara + gd_optimize_enepara
    on the reference's parameter files give the tables the reference's own
    loadEnePara/OptimizeEnepara produced (bitwise)."""
    _, r = fixtures.load_1a07c1()
    ep0, ep = synth.enepara_from_params()
    for name, key in (("p1a", "ep_p1a"), ("p2a", "ep_p2a"), ("pmf0", "ep_pmf0"),
                      ("pmf1", "ep_pmf1"), ("hdb0", "ep_hdb0"), ("hdb1", "ep_hdb1"),
                      ("hdb2", "ep_hdb2")):
        got = np.ctypeslib.as_array(getattr(ep, name)).ravel()
        want = r[key]
        # hdb1/hdb2 of absent pairs are 1/0: compare with inf/nan equal
        np.testing.assert_array_equal(got, want, err_msg=name)
    sc = [ep.lj0, ep.lj1, ep.el0, ep.el1, ep.a1, ep.b1, ep.kde2, ep.kde3]
    np.testing.assert_array_equal(np.float32(sc), r["ep_scalars"])
    for name, key in (("ele", "ep_ele"), ("hpp", "ep_hpp"), ("hpl0", "ep_hpl0"),
                      ("hpl1", "ep_hpl1"), ("hpl2", "ep_hpl2"), ("a_para", "ep_a"),
                      ("b_para", "ep_b")):
        np.testing.assert_array_equal(np.ctypeslib.as_array(getattr(ep, name)), r[key],
                                      err_msg=name)
    # weights: the 9 file values (the reference's 10th slot is never used)
    np.testing.assert_array_equal(np.ctypeslib.as_array(ep.w)[:9], r["ep_w"][:9])


def test_set_temperature_matches_reference():
    _, r = fixtures.load_1a07c1()
    t = remc.temperatures(1, 0.044, 0.036)
    assert np.float32(t[0].minus_beta) == r["temp_minus_beta"][0]


@pytest.mark.parametrize("n_tmp", [1, 2, 3, 8, 16, 32, 64])
@pytest.mark.parametrize("lo, hi", [("0.3", "3.0"), ("0.044", "0.036"), ("0.02", "0.036")])
def test_set_temperature_ladders_match_reference(n_tmp, lo, hi):
    """Every minus_beta of the geometric ladder equals the reference build's
    SetTemperature (util.C:490-532) bit for bit, for the ladder lengths the
    configurations use and beyond (tests/golden/make_1a07c1.py writes them)."""
    lad = np.load(os.path.join(fixtures.GOLDEN, "set_temperature_ladders.npz"))
    want = lad[f"n{n_tmp}_{lo}_{hi}"]
    t = remc.temperatures(n_tmp, float(lo), float(hi))
    got = np.array([x.minus_beta for x in t], np.float32)
    np.testing.assert_array_equal(got, want)
    assert [x.order for x in t] == list(range(n_tmp))


def test_set_replica_layout():
    size = remc.ComplexSize(3, 2, 4, 24, 5, 10, 0, 0)
    arr = (remc.Replica * 24)()
    remc.check(remc.lib().gd_set_replica(arr, size))
    for p in range(2):
        for t in range(4):
            for l in range(3):
                f = 4 * 3 * p + 3 * t + l
                assert (arr[f].idx_rep, arr[f].idx_prt, arr[f].idx_tmp, arr[f].idx_lig) == (f, p, t, l)


def test_replica_table_must_hold_whole_chains():
    """The exchange inverts each chain's temperature indices
    (exchangereplicas_d.cu:34-38): a caller's replica table is accepted when
    every chain slot holds its own conformers and the chain's temperatures
    are a permutation, and refused otherwise (before anything reaches a GPU)."""
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=2, n_prt=2, n_tmp=3, steps=10)
    size = cx.sized(3)

    def slot(p, t, l):
        return 3 * 2 * p + 2 * t + l

    rep = remc.replica_table(size)
    assert remc.JobSpec(cx, temps, mc, replica=rep).footprint() > 0
    # a permutation of the temperatures of chain (p 1, l 0) is a valid state
    rep[slot(1, 0, 0)].idx_tmp, rep[slot(1, 2, 0)].idx_tmp = 2, 0
    assert remc.JobSpec(cx, temps, mc, replica=rep).footprint() > 0
    # the same temperature twice in chain (p 0, l 1)
    rep = remc.replica_table(size)
    rep[slot(0, 1, 1)].idx_tmp = 0
    with pytest.raises(RuntimeError, match="appears twice"):
        remc.JobSpec(cx, temps, mc, replica=rep).footprint()
    # a slot of chain (p 0, l 1) holding protein conformer 1
    rep = remc.replica_table(size)
    rep[slot(0, 2, 1)].idx_prt = 1
    with pytest.raises(RuntimeError, match="belongs to the chain"):
        remc.JobSpec(cx, temps, mc, replica=rep).footprint()


def test_bad_inputs_are_rejected_without_a_gpu():
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=1, n_prt=1, n_tmp=2, steps=10)
    size = cx.sized(2)
    bad = remc.ComplexSize(size.n_lig, size.n_prt, size.n_tmp, size.n_rep, 0, size.pnp,
                           size.pnk, size.pos)
    rc = remc.lib().gd_run(*cx.args(), temps, None, C.byref(mc), None, bad, None, None)
    assert rc == -1
    assert b"lna" in remc.lib().gd_last_error()
    cx.lig_t[0] = 40
    rc = remc.lib().gd_run(*cx.args(), temps, None, C.byref(mc), None, size, None, None)
    assert rc == -1 and b"type" in remc.lib().gd_last_error()


def test_kde_point_limit():
    """KDE points of the ligand's atom types are indexed in 16 bits (the
    reference holds at most MAXKDE = 10,000, size.h:73): more is an error
    from the host image, before any device work."""
    temps = remc.temperatures(1, 0.3, 0.3)
    mc = remc.mcpara(10, 10, [0.1] * 6)
    cx = synth.make_complex(pnp=100, radius=10.0, lna=8, n_lig=1, n_prt=1, pnk=200000, pos=0,
                            seed=3)
    assert np.isin(cx.kde_t, cx.lig_t).sum() > 65535
    with pytest.raises(remc.GDError, match="KDE points"):
        remc.JobSpec(cx, temps, mc).footprint()
    cx = synth.make_complex(pnp=100, radius=10.0, lna=8, n_lig=1, n_prt=1, pnk=10000, pos=0,
                            seed=3)
    assert remc.JobSpec(cx, temps, mc).footprint() > 0


def test_valu_probe_arguments():
    ns, tf, n = C.c_double(), C.c_double(), C.c_int()
    rc = remc.lib().gd_valu_probe(0, 0, C.byref(ns), C.byref(tf), C.byref(n))
    assert rc == -1 and b"waves_per_simd" in remc.lib().gd_last_error()
    if remc.lib().gd_device_count() == 0:
        rc = remc.lib().gd_valu_probe(0, 8, C.byref(ns), C.byref(tf), C.byref(n))
        assert rc == -4


def test_valu_probe_kinds_arguments():
    ns = (C.c_double * remc.VALU_KINDS)()
    rc = remc.lib().gd_valu_probe_kinds(0, 8, ns, remc.VALU_KINDS - 1)
    assert rc == -1 and b"n_kinds" in remc.lib().gd_last_error()
    rc = remc.lib().gd_valu_probe_kinds(0, 9, ns, remc.VALU_KINDS)
    assert rc == -1
    names = [remc.lib().gd_valu_kind_name(k) for k in range(remc.VALU_KINDS)]
    assert all(names) and len(set(names)) == remc.VALU_KINDS
    assert remc.lib().gd_valu_kind_name(remc.VALU_KINDS) is None
    assert remc.lib().gd_valu_kind_name(-1) is None
    if remc.lib().gd_device_count() == 0:
        assert remc.lib().gd_valu_probe_kinds(0, 8, ns, remc.VALU_KINDS) == -4


@pytest.mark.gpu
def test_valu_probe_kinds_on_gpu():
    """Issue time per instruction kind; only plausibility is checked (every
    kind between 0.2 and 50 ns per wavefront instruction per SIMD)."""
    rates = remc.valu_probe_kinds(8)
    assert len(rates) == remc.VALU_KINDS
    assert all(0.2 < v < 50.0 for v in rates.values()), rates


@pytest.mark.gpu
def test_valu_probe_on_gpu():
    """The FP32 FMA issue ceiling: about one wavefront instruction per 1.1 ns
    per SIMD on MI355X (1,024 SIMDs); only its plausibility is checked."""
    ns, tflops, n_simd = remc.valu_probe(8)
    assert n_simd >= 4 and 0.2 < ns < 20.0
    assert abs(tflops - n_simd * 128 / ns / 1e3) < 1e-6 * tflops


def test_load_enepara_missing_file():
    ep0 = remc.EnePara0()
    rc = remc.lib().gd_load_enepara(b"/nonexistent/gpudocksm.ff", C.byref(ep0))
    assert rc == -5


@pytest.mark.skipif(remc.lib().gd_device_count() > 0, reason="a GPU is visible")
def test_no_gpu_is_an_error_not_a_fallback():
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=1, n_prt=1, n_tmp=2, steps=10)
    with pytest.raises(remc.GDError, match="no HIP device"):
        remc.run(cx, temps, mc)
    with pytest.raises(remc.GDError, match="no HIP device"):
        remc.score_poses(cx, [0], [0], np.zeros((1, 6)))


def test_server_lane_settings_checked_before_device_work():
    """Lane bounds must be positive and increasing, at most four lanes, and
    a lane's split factor one of 0, 1, 4, 8, 16: refused before any engine
    touches a device (so this runs without a GPU)."""
    for kw, msg in [(dict(lane_max_pairs=(5000, 3000)), "increasing"),
                    (dict(lane_max_pairs=(0,)), "increasing"),
                    (dict(lane_max_pairs=(1, 2, 3, 4)), "lanes"),
                    (dict(lane_max_pairs=(1000,), lane_split=(3, 0)), "lane_split")]:
        with pytest.raises(remc.GDError, match=msg):
            remc.Server(**kw)
    import bench
    assert bench.lane_bounds("20000, 40000") == (20000, 40000) and bench.lane_bounds("") == ()


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU; --gpus N > 1 exits with a message."""
    import subprocess
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
