"""Long-run statistics: the GPU chains against the reference algorithm.

The GPU run (canonical FP32) and the reference-order CPU run (libm expf,
IEEE division and square root, the reference's loop order) use independent
Philox seeds, so their trajectories are independent samples.  Per
temperature, the mean energy over the second half of each chain and each
chain's acceptance rate must agree between the two runs (paired t-test on the
16 chains of each side, one per protein conformer, paired by conformer).
"""
import numpy as np
import pytest
from scipy import stats

import oracle_lib as O
from gpudock_amd import remc, synth, traj

pytestmark = pytest.mark.gpu

N_STEPS = 4000
STRIDE = 50


def test_energy_and_acceptance_statistics_match_reference(tmp_path):
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=1, n_prt=16, steps=N_STEPS)
    n_tmp, n_prt = cfg["n_tmp"], 16
    mc.steps_per_dump = N_STEPS
    prefix = str(tmp_path / "t")
    opt = remc.default_options(seed=101, exchange=0, away=0, record_stride=STRIDE,
                               record_prefix=prefix)
    g, _ = remc.run(cx, temps, mc, opt)
    _, rec = traj.read_gdtraj(prefix + "_0000.gdtraj")
    e_gpu = rec["energy"][:, :, 9]                 # [n_rep, n_samp] total energy
    t_gpu = rec["idx_tmp"]

    acs_ref, e_ref, t_ref = O.ref_run_samples(cx, temps, list(mc.move_scale), N_STEPS,
                                              cfg["spe"], STRIDE, exchange=0, seed=202,
                                              threads=16)
    assert np.array_equal(t_gpu[:, 0], t_ref[:, 0])
    half = e_gpu.shape[1] // 2
    # replica r = n_tmp * n_lig * p + n_lig * t + l (one ligand conformer)
    tmp_of = t_gpu[:, 0]
    for t in range(n_tmp):
        rows = np.nonzero(tmp_of == t)[0]
        assert len(rows) == n_prt
        m_gpu = e_gpu[rows, half:].astype(np.float64).mean(axis=1)
        m_ref = e_ref[rows, half:].astype(np.float64).mean(axis=1)
        p_e = stats.ttest_rel(m_gpu, m_ref).pvalue
        a_gpu = g["acs_mc"][rows] / N_STEPS
        a_ref = acs_ref[rows] / N_STEPS
        p_a = stats.ttest_rel(a_gpu, a_ref).pvalue
        print(f"T{t}: energy {m_gpu.mean():.4f} / {m_ref.mean():.4f} (p {p_e:.3f}), "
              f"acceptance {a_gpu.mean():.4f} / {a_ref.mean():.4f} (p {p_a:.3f})")
        assert p_e > 1e-3, (t, m_gpu.mean(), m_ref.mean(), p_e)
        assert p_a > 1e-3, (t, a_gpu.mean(), a_ref.mean(), p_a)
    # acceptance rises with temperature on both sides
    a_gpu_t = [g["acs_mc"][tmp_of == t].mean() for t in range(n_tmp)]
    a_ref_t = [acs_ref[tmp_of == t].mean() for t in range(n_tmp)]
    assert a_gpu_t[-1] > a_gpu_t[0] and a_ref_t[-1] > a_ref_t[0]
