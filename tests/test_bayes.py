"""Bayesian total energy (IS_BAYE; gpudocksm-bayes-1.0 calcenergy_d.cu:315-367,
util_d.cu:220-303, distribution.h:21-77).  No reference value exists for
these pdfs: they are pinned by their closed forms (scipy.stats), by a
float64 restatement of the total, and by the MC run staying finite."""
import math

import numpy as np
import pytest
from scipy import stats

import oracle_lib as O
from gpudock_amd import remc, synth

HIGH = [(0.05286683, 0.29916642), (-0.3653806, 0.00681102), (-0.0221777, 0.11847532),
        (-1.0041231, 0.1151713), (0.60986808, 0.30974999), (-0.8246522, 0.05218529),
        (-0.2346999, 0.09871713), (-1.0112055, 0.09882933), (-0.3377100, 0.09282947)]
LOW = [(0.15818486, 0.29032471), (-0.3647824, 0.00534225), (0.04257942, 0.08801551),
       (-1.0019512, 0.10312611), (0.76967708, 0.11224029), (-0.7509618, 0.08757299),
       (-0.2022064, 0.07854212), (-1.004605, 0.02967548), (-0.2387023, 0.12134983)]
FAMILY = [0, 1, 2, 3, 0, 2, 2, 3, 2]   # file order vdw ele pmf hpc hdb dst psp kde lhm


def pdf64(fam, x, loc, scale):
    """float64 pdf with the canonical floor: 1e-8 where the pdf is zero or
    its exponential factor is below e^-86 (where the FP32 exp returns 0)."""
    with np.errstate(all="ignore"):
        return _pdf64(fam, x, loc, scale)


def _pdf64(fam, x, loc, scale):
    z = (np.asarray(x, np.float64) - loc) / scale
    arg = np.zeros_like(z)
    if fam == 0:
        v = stats.norm.pdf(z) / scale
        arg = -0.5 * z * z
    elif fam == 1:
        v = stats.cauchy.pdf(z) / scale
    elif fam == 2:
        v = stats.logistic.pdf(z) / scale
        arg = -np.abs(z)
    else:
        zz = np.maximum(z, 1e-300)
        v = np.where(z > 0, stats.invgauss.pdf(zz, mu=1.0) / scale, 1e-8)
        arg = np.where(z > 0, -(zz - 1.0) ** 2 / (2.0 * zz), 0.0)
    return np.where((v > 0) & np.isfinite(v) & (arg >= -86.0), v, 1e-8)


@pytest.mark.parametrize("fam,loc,scale", [(0, 0.05, 0.3), (1, -0.36, 0.0068),
                                           (2, -0.02, 0.118), (3, -1.0, 0.115)])
def test_pdfs_match_closed_forms(fam, loc, scale):
    rng = np.random.default_rng(fam)
    x = (loc + scale * rng.uniform(-8, 8, 20000)).astype(np.float32)
    if fam == 3:
        x = (loc + scale * rng.uniform(0.02, 8, 20000)).astype(np.float32)
    got = O.bayes_pdf(fam, x, loc, scale)
    ref = pdf64(fam, x.astype(np.float64), loc, scale)
    big = ref > 1e-6
    np.testing.assert_allclose(got[big], ref[big], rtol=5e-6)


def test_pdf_special_values():
    # peak values: 1/(s sqrt(2 pi)), 1/(pi s), 1/(4 s), Wald at nx = 1: 1/(s sqrt(2 pi))
    s = np.float32(0.25)
    assert abs(O.bayes_pdf(0, [0.5], 0.5, s)[0] - 1 / (0.25 * math.sqrt(2 * math.pi))) < 1e-6
    assert abs(O.bayes_pdf(1, [0.5], 0.5, s)[0] - 1 / (math.pi * 0.25)) < 1e-6
    assert abs(O.bayes_pdf(2, [0.5], 0.5, s)[0] - 1.0) < 1e-6
    assert abs(O.bayes_pdf(3, [0.75], 0.5, s)[0] - 1 / (0.25 * math.sqrt(2 * math.pi))) < 1e-5
    # floors: Wald at or below its location, underflowed normal tail; the
    # logistic stays finite far out on both sides (the reference's e^-z
    # overflows for z < -88)
    assert O.bayes_pdf(3, [0.5, 0.4], 0.5, s).tolist() == [np.float32(1e-8)] * 2
    assert O.bayes_pdf(0, [100.0], 0.0, s)[0] == np.float32(1e-8)
    far = O.bayes_pdf(2, [-40.0, 40.0], 0.0, s)
    assert np.all(np.isfinite(far)) and far[0] == far[1]


def _f32(p):
    return tuple(float(np.float32(v)) for v in p)


def bayes64(nf):
    """float64 total with the parameters as the FP32 constants they are"""
    ph = sum(math.log10(pdf64(FAMILY[k], nf[k], *_f32(HIGH[k]))) for k in range(9))
    pl = sum(math.log10(pdf64(FAMILY[k], nf[k], *_f32(LOW[k]))) for k in range(9))
    return pl - ph


def test_total_matches_float64():
    rng = np.random.default_rng(7)
    for _ in range(300):
        nf = np.array([rng.normal(l, 2 * s) for l, s in HIGH], np.float32)
        nf[3] = abs(nf[3] - HIGH[3][0]) + HIGH[3][0] + 1e-3     # Wald terms above loc
        nf[7] = abs(nf[7] - HIGH[7][0]) + HIGH[7][0] + 1e-3
        got = O.bayes_total(nf)
        ref = bayes64(nf.astype(np.float64))
        assert abs(got - ref) <= 2e-5 * max(1.0, abs(ref)), (nf, got, ref)


def test_replay_with_bayes_total():
    """The MC run with the Bayesian total stays finite, its totals are the
    Bayesian combination of its own normalised terms, and it differs from
    the weighted-sum run."""
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=2, n_prt=1, n_tmp=4, steps=100)
    rb = O.replay_run(cx, temps, mc, remc.default_options(seed=2, combine=1))
    rl = O.replay_run(cx, temps, mc, remc.default_options(seed=2, combine=0))
    assert np.all(np.isfinite(rb["energy_final"]))
    slot_of_file = [0, 1, 2, 5, 4, 8, 3, 6, 7]
    for e in rb["energy_final"]:
        assert O.bayes_total(e[slot_of_file]) == e[9]
    assert not np.array_equal(rb["move_final"], rl["move_final"])
