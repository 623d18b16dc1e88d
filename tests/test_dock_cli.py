"""gd_dock, the command-line front end (the reference's dock, dock.C:14-130).

CPU: argument checks and input errors exit with status 1 and the reference's
messages (util.C:152-166).  GPU: a run on the 1a07C1 files writes, per
replica, exactly what remc.run gives for the same inputs and options.
"""
import csv
import os
import subprocess

import numpy as np
import pytest

from gpudock_amd import remc

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
GOLDEN = os.path.join(HERE, "golden")
PARAMS = os.path.join(ROOT, "gpudock_amd", "params")
DOCK = os.path.join(ROOT, "gpudock_amd", "gd_dock")

FILES = ["-p", os.path.join(GOLDEN, "1a07C.pdb"), "-l", os.path.join(GOLDEN, "1a07C1.sdf"),
         "-s", os.path.join(GOLDEN, "1a07C1-0.8.ff"), "-id", "1a07C1",
         "-para", os.path.join(PARAMS, "gpudocksm.ff"),
         "-opt_fn", os.path.join(PARAMS, "08ff_opt"),
         "-nor_a", os.path.join(PARAMS, "08_nor_a"), "-nor_b", os.path.join(PARAMS, "08_nor_b")]


def _run(args):
    return subprocess.run([DOCK] + args, capture_output=True, text=True, timeout=300)


def test_cli_requires_inputs():
    r = _run([])
    assert r.returncode == 1 and "Provide target protein structure" in r.stdout
    r = _run(["-p", "x.pdb"])
    assert r.returncode == 1 and "Provide compound library in SD format" in r.stdout
    r = _run(["-p", "x.pdb", "-l", "y.sdf"])
    assert r.returncode == 1 and "Provide LHM potentials" in r.stdout
    r = _run(["-h"])
    assert r.returncode == 0 and "-floor_temp" in r.stdout


def test_cli_reports_unreadable_inputs(tmp_path):
    args = list(FILES)
    args[args.index("-l") + 1] = str(tmp_path / "missing.sdf")
    r = _run(args)
    assert r.returncode == 1 and "inputs: cannot open" in r.stderr
    args = list(FILES)
    args[args.index("-para") + 1] = str(tmp_path / "missing.ff")
    r = _run(args)
    assert r.returncode == 1 and "parameters: cannot open" in r.stderr


@pytest.mark.skipif(remc.lib().gd_device_count() > 0, reason="a GPU is present")
def test_cli_without_gpu_fails_loudly():
    """Inputs load; the run itself refuses to fall back to the CPU."""
    r = _run(FILES + ["-ns", "10", "-nc", "5"])
    assert r.returncode == 1 and r.stderr.startswith("run:")


@pytest.mark.gpu
def test_cli_run_equals_library(tmp_path):
    out = tmp_path / "res.csv"
    r = _run(FILES + ["-ns", "60", "-nc", "10", "-nt", "3", "-floor_temp", "0.044",
                      "-ceiling_temp", "0.036", "-t", "0.02", "-r", "0.08", "-exchange", "1",
                      "-seed", "7", "-csv", str(out)])
    assert r.returncode == 0, r.stderr
    assert "AR of MC" in r.stdout and "MC sweeps per second" in r.stdout
    assert "compute time" in r.stdout and "call time" in r.stdout
    rows = list(csv.DictReader(open(out)))

    from test_loaders import load_1a07c1_files
    cx = load_1a07c1_files()
    temps = remc.temperatures(3, 0.044, 0.036)
    mc = remc.mcpara(60, 10, [0.02, 0.02, 0.02, 0.08, 0.08, 0.08], steps_per_dump=1000)
    opt = remc.default_options(seed=7, exchange=1, away=1, mcc=1)
    g, _ = remc.run(cx, temps, mc, opt)
    assert len(rows) == len(g["acs_mc"]) == 21 * 3 * 3
    names = ["vdw", "ele", "pmf", "psp", "hdb", "hpc", "kde", "lhm", "dst", "total"]
    for i, row in enumerate(rows):
        assert int(row["acs_mc"]) == g["acs_mc"][i]
        assert int(row["step_best"]) == g["step_best"][i]
        mv = np.array([float(row[f"move_best{k}"]) for k in range(6)], np.float32)
        assert np.array_equal(mv, g["move_best"][i])
        eb = np.array([float(row["best_" + n]) for n in names], np.float32)
        assert np.array_equal(eb, g["energy_best"][i])
        e64 = np.array([float(row["best64_" + n]) for n in names])
        assert np.array_equal(e64, g["energy_best_fp64"][i])
