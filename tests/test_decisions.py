"""CPU checks that pin the oracle's Monte Carlo decisions and launch schedule
to the reference without retyping the oracle's own code: the schedule is
compared with launch sequences walked by hand through gpusingle.cu:8-20, the
Metropolis and exchange rules with their closed forms
(accept_d.cu:22-23, exchangereplicas_d.cu:43-66), and the rotation with a
matrix product built here from its three factors (move_d.cu:97-105).  Each
test is chosen so that a plausible mistake (a flipped sign of beta, a mode
that never restarts, a transposed composition, a swap in the wrong
direction) fails it."""
import math

import numpy as np
import pytest

import oracle_lib as O
from gpudock_amd import remc, synth


# ------------------------------------------------------------ launch schedule
# gpusingle.cu:8-20 walked by hand:
#   for (s1 = 0; s1 < steps_total; s1 += steps_per_dump)
#     for (s2 = 0; s2 < steps_per_dump; s2 += steps_per_exchange)
#       MonteCarlo_d; ExchangeReplicas_d (mode_t = (s2 / spe) % 2)
@pytest.mark.parametrize("total, spd, spe, modes", [
    # windows at s1 = 0, 10, 20; s2 = 0, 3, 6, 9 in each: 12 launches (36 steps)
    (25, 10, 3, [0, 1, 0, 1] * 3),
    # five launches per window: the mode restarts at 0 after an odd count
    (100, 50, 10, [0, 1, 0, 1, 0, 0, 1, 0, 1, 0]),
    # a partial last window still runs whole: 2 windows x 2 launches
    (15, 10, 5, [0, 1, 0, 1]),
    # steps_per_dump below steps_per_exchange: one launch per window, mode 0
    (20, 4, 10, [0, 0, 0, 0, 0]),
    # one window: the mode alternates over the whole run
    (60, 60, 10, [0, 1, 0, 1, 0, 1]),
    (0, 10, 5, []),
])
def test_schedule_follows_the_reference_loops(total, spd, spe, modes):
    assert list(O.schedule(total, spd, spe)) == modes


def test_schedule_rejects_zero_step_sizes():
    assert O.schedule(10, 0, 5) is None
    assert O.schedule(10, 5, 0) is None


def test_window_restart_changes_the_run():
    """With five launches per window the restarted mode pairs different
    temperatures than one alternation over the run would: the two replays
    part ways (so the GPU comparison with the replay sees the schedule)."""
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=2, n_prt=1, n_tmp=4, steps=100,
                                           pnp=120, lna=12, pnk=100)
    opt = remc.default_options(seed=13, complex_id=2)
    mc.steps_per_dump = 50
    a = O.replay_run(cx, temps, mc, opt)
    mc.steps_per_dump = 100
    b = O.replay_run(cx, temps, mc, opt)
    assert not np.array_equal(a["acs_exchg"], b["acs_exchg"]) or \
        not np.array_equal(a["idx_tmp_final"], b["idx_tmp_final"])


def test_library_rejects_a_zero_dump_window():
    """The library takes steps_per_dump from McPara as the reference does;
    a window of 0 steps (an endless loop in gpusingle.cu:8) is an error."""
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=1, n_prt=1, n_tmp=2, steps=20)
    mc.steps_per_dump = 0
    with pytest.raises(RuntimeError, match="steps_per_dump"):
        remc.JobSpec(cx, temps, mc).footprint()


# ----------------------------------------------------------------- Metropolis
def _f32(x):
    return float(np.float32(x))


@pytest.mark.parametrize("canonical", [0, 1])
def test_metropolis_closed_form(canonical):
    """accept iff u < exp((E_new - E_old) * minus_beta), minus_beta = -1/T:
    draws just below and just above exp(-dE/T) (float64) land on either side,
    whatever the sign of dE."""
    rng = np.random.default_rng(11)
    n_checked = 0
    for _ in range(3000):
        T = rng.uniform(0.05, 5.0)
        mb = _f32(-1.0 / T)
        e_old = _f32(rng.uniform(-20.0, 20.0))
        e_new = _f32(e_old + rng.normal(0.0, 3.0 * T))
        de = np.float32(e_new) - np.float32(e_old)
        p = math.exp(float(de) * mb)
        if p >= 1.0:
            # downhill (or level) moves: every draw below 1 is accepted
            assert O.accept(e_new, e_old, mb, _f32(1.0 - 2.0 ** -24), canonical)
            continue
        if p < 2.0 ** -20:
            continue
        assert O.accept(e_new, e_old, mb, _f32(p * (1.0 - 1e-4)), canonical)
        assert not O.accept(e_new, e_old, mb, _f32(p * (1.0 + 1e-4)), canonical)
        n_checked += 1
    assert n_checked > 1000


@pytest.mark.parametrize("canonical", [0, 1])
def test_metropolis_edges(canonical):
    # dE < 0: accepted at the largest draw u = 1
    assert O.accept(-5.0, -4.0, -2.0, 1.0, canonical)
    # dE = 0: exp(0) = 1 and the test is strict, so u = 1 is refused
    # (curand_uniform returns (0, 1], accept_d.cu:23)
    assert not O.accept(-4.0, -4.0, -2.0, 1.0, canonical)
    assert O.accept(-4.0, -4.0, -2.0, _f32(1.0 - 2.0 ** -24), canonical)
    # uphill by 5 T at u = 0.5: refused (a flipped sign of beta accepts it)
    assert not O.accept(1.0, 0.0, -5.0, 0.5, canonical)
    # far uphill at the smallest draw 2^-24: exp(-200) = 0, refused
    assert not O.accept(200.0, 0.0, -1.0, 2.0 ** -24, canonical)


def test_reversed_metropolis_closed_form():
    """IS_REVERSE (accept_d.cu:24-27): the second draw decides, against
    exp(-(E_new - E_old) * minus_beta); the first draw plays no part."""
    rng = np.random.default_rng(12)
    n_checked = 0
    for _ in range(3000):
        T = rng.uniform(0.05, 5.0)
        mb = _f32(-1.0 / T)
        e_old = _f32(rng.uniform(-20.0, 20.0))
        e_new = _f32(e_old + rng.normal(0.0, 3.0 * T))
        de = np.float32(e_new) - np.float32(e_old)
        p = math.exp(-float(de) * mb)
        for u in (_f32(2.0 ** -24), _f32(1.0)):
            if p >= 1.0:
                assert O.metropolis(e_new, e_old, mb, u, _f32(1.0 - 2.0 ** -24), reverse=1)
                continue
            if p < 2.0 ** -20:
                continue
            assert O.metropolis(e_new, e_old, mb, u, _f32(p * (1.0 - 1e-4)), reverse=1)
            assert not O.metropolis(e_new, e_old, mb, u, _f32(p * (1.0 + 1e-4)), reverse=1)
            n_checked += 1
    assert n_checked > 1000
    # the reversed test takes uphill moves the plain one refuses, and back
    assert O.metropolis(1.0, 0.0, -5.0, 1.0, 0.5, reverse=1)
    assert not O.metropolis(1.0, 0.0, -5.0, 0.5, 1.0)
    assert not O.metropolis(-1.0, 0.0, -5.0, 1.0, 0.5, reverse=1)


def test_fixed_draws_accept_everything():
    """IS_RANDOM 0 (accept_d.cu:19-20): every step is accepted, whatever the
    energies and draws; the move perturbation is 2 x 20 - 1 = 39 (move_d.cu:81)."""
    for e_new, e_old in ((1e6, -1e6), (0.0, 0.0), (-3.0, 5.0)):
        for u in (2.0 ** -24, 1.0):
            assert O.metropolis(e_new, e_old, -1.0, u, u, fixed_draws=1)
            assert O.metropolis(e_new, e_old, -1.0, u, u, fixed_draws=1, reverse=1)


def test_fixed_draws_replay_moves_deterministically():
    """The replay with IS_RANDOM 0: every step accepted, each movematrix the
    float32 sum of its start and step x (39 x move_scale), and a temperature
    pair swaps only when exp(delta) > 20 (exchangereplicas_d.cu:61)."""
    cx, temps, mc, cfg = synth.make_config("c0", lna=8, pnp=60, pnk=40, n_lig=2, n_prt=1,
                                           n_tmp=4, steps=30, spe=10)
    opt = remc.default_options(seed=4, fixed_draws=1)
    r = O.replay_run(cx, temps, mc, opt)
    n_rep = 2 * 4
    assert np.all(r["acs_mc"] == 30)
    scale = np.asarray(mc.move_scale, np.float32)
    step = scale * np.float32(39.0)
    m = np.zeros(6, np.float32)
    for _ in range(30):
        m = step + m
    for i in range(n_rep):
        np.testing.assert_array_equal(r["move_final"][i], m)
    # every pose of a replica is the same on every step, so the exchanges
    # follow from the energies alone: swap iff exp((mb1 - mb2)(E2 - E1)) > 20
    r2 = O.replay_run(cx, temps, mc, remc.default_options(seed=99, fixed_draws=1))
    np.testing.assert_array_equal(r["idx_tmp_final"], r2["idx_tmp_final"])
    np.testing.assert_array_equal(r["acs_exchg"], r2["acs_exchg"])


# ------------------------------------------------------------------ exchange
def _always(n):
    """Draws that every pair with equal energies passes: exp(0) = 1 > 0.5."""
    return np.full(max(n // 2, 1), 0.5, np.float32)


@pytest.mark.parametrize("n, mode, after", [
    (2, 0, [1, 0]),
    (2, 1, [0, 1]),            # mode 1 with two temperatures has no pair
    (3, 0, [1, 0, 2]),
    (3, 1, [0, 2, 1]),
    (4, 0, [1, 0, 3, 2]),
    (4, 1, [0, 2, 1, 3]),
    (5, 0, [1, 0, 3, 2, 4]),
    (5, 1, [0, 2, 1, 4, 3]),
    (6, 1, [0, 2, 1, 4, 3, 5]),
])
@pytest.mark.parametrize("canonical", [0, 1])
def test_exchange_pairs_by_mode(n, mode, after, canonical):
    """exchangereplicas_d.cu:43-46: mode 0 swaps temperatures (0,1), (2,3), ...;
    mode 1 swaps (1,2), (3,4), ... -- listed here by hand."""
    mb = -1.0 / np.linspace(0.3, 3.0, n)
    temps, acs = O.exchange_chain(mb, list(range(n)), np.zeros(n), mode, _always(n), canonical)
    assert list(temps) == after
    assert list(acs) == [int(a != b) for a, b in zip(after, range(n))]


@pytest.mark.parametrize("canonical", [0, 1])
def test_exchange_follows_temperatures_not_slots(canonical):
    """Slots hold temperatures (2, 0, 1): mode 0 pairs temperatures 0 and 1,
    held by slots 1 and 2, whose temperatures trade places."""
    temps, acs = O.exchange_chain(-1.0 / np.array([0.5, 1.0, 2.0]), [2, 0, 1], np.zeros(3), 0,
                                  _always(3), canonical)
    assert list(temps) == [2, 1, 0]
    assert list(acs) == [0, 1, 1]


@pytest.mark.parametrize("canonical", [0, 1])
def test_exchange_lower_energy_goes_to_lower_temperature(canonical):
    """Whenever the colder slot holds the higher energy, delta > 0 and the swap
    happens at every draw, so afterwards the lower energy is the colder."""
    rng = np.random.default_rng(5)
    for _ in range(500):
        T0, T1 = sorted(rng.uniform(0.1, 4.0, 2))
        if T1 / T0 < 1.001:
            continue
        e_cold, e_hot = sorted(rng.uniform(-30.0, 30.0, 2))[::-1]
        if e_cold - e_hot < 1e-3:
            continue
        temps, _ = O.exchange_chain([-1.0 / T0, -1.0 / T1], [0, 1], [e_cold, e_hot], 0,
                                    [1.0], canonical)
        assert list(temps) == [1, 0]


@pytest.mark.parametrize("canonical", [0, 1])
@pytest.mark.parametrize("e0, e1", [(-3.5, -3.0), (-3.0, -3.5), (0.0, 0.4), (2.0, 0.5)])
def test_two_temperature_swap_frequency(canonical, e0, e1):
    """Over 10^5 draws of the (0, 1] grid the swap rate of temperatures
    T0 = 0.5 (slot 0, energy e0) and T1 = 0.8 (slot 1, e1) is
    min(1, exp((1/T0 - 1/T1)(e0 - e1))) within 5 binomial standard errors."""
    T0, T1 = 0.5, 0.8
    p = min(1.0, math.exp((1.0 / T0 - 1.0 / T1) * (e0 - e1)))
    n = 100000
    rng = np.random.default_rng(17)
    draws = ((rng.integers(0, 1 << 24, n) + 1) * 2.0 ** -24).astype(np.float32)
    mb = np.array([-1.0 / T0, -1.0 / T1], np.float32)
    ener = np.array([e0, e1], np.float32)
    t = np.zeros(2, np.int32)
    acs = np.zeros(2, np.int32)
    fn = O.lib().orc_exchange_chain
    u = np.zeros(1, np.float32)
    swaps = 0
    for k in range(n):
        t[:] = (0, 1)
        u[0] = draws[k]
        fn(canonical, 2, O._fp(mb), O._ip(t), O._fp(ener), 0, O._fp(u), O._ip(acs))
        swaps += int(t[0] == 1)
    sigma = math.sqrt(max(p * (1.0 - p), 1e-12) / n)
    assert abs(swaps / n - p) <= 5.0 * sigma + 1e-9, (swaps / n, p)


# ------------------------------------------------------------ cold chains
def _ladder(minus_betas):
    t = (remc.Temp * len(minus_betas))()
    for i, mb in enumerate(minus_betas):
        t[i].t = -1.0 / mb
        t[i].minus_beta = mb
        t[i].order = i
    return t


def test_energy_never_rises_near_zero_temperature():
    """At T = 1e-9 every uphill move has exp(-dE/T) = 0 below any draw, so the
    energy before each step never rises, in the canonical replay (every
    recorded step) and in the reference-order run (sampled steps)."""
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=2, n_prt=1, n_tmp=2, steps=80,
                                           pnp=120, lna=12, pnk=100)
    cold = _ladder([-1e9, -1e9])
    opt = remc.default_options(seed=3, exchange=0)
    res, rec = O.replay_run_records(cx, cold, mc, opt, 1)
    e = rec["energy"][:, :, 9]
    assert np.all(np.diff(e, axis=1) <= 0.0)
    assert np.all(res["energy_final"][:, 9] <= e[:, 0])
    assert res["acs_mc"].sum() > 0          # downhill moves were taken
    acs, es, ts, acx = O.ref_run_samples(cx, cold, list(mc.move_scale), 80, 10, 1)
    assert np.all(np.diff(es, axis=1) <= 0.0)
    assert acs.sum() > 0


# ------------------------------------------------------------------- rotation
def _rz(a):
    c, s = math.cos(a), math.sin(a)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def _ry(b):
    c, s = math.cos(b), math.sin(b)
    return np.array([[c, 0.0, s], [0.0, 1.0, 0.0], [-s, 0.0, c]])


def _rx(g):
    c, s = math.cos(g), math.sin(g)
    return np.array([[1.0, 0.0, 0.0], [0.0, c, -s], [0.0, s, c]])


@pytest.mark.parametrize("fn", ["orc_move_ref32", "orc_canon_move"])
def test_rotation_is_rz_ry_rx_composed(fn):
    """x' = Rz(m3) Ry(m4) Rx(m5) x + t + c for angles all non-zero
    (move_d.cu:97-105, 120-127), in the reference-order move (libm sinf/cosf)
    and the canonical one the GPU runs: the matrix is built here as the
    product of the three elementary rotations; Rx Ry Rz or any other order
    misses by far more than FP32 rounding."""
    rng = np.random.default_rng(23)
    n = 16
    x = rng.normal(0.0, 3.0, (3, n)).astype(np.float32)
    c = rng.normal(0.0, 10.0, 3).astype(np.float32)
    move = getattr(O.lib(), fn)
    for _ in range(50):
        m = np.concatenate([rng.normal(0.0, 1.0, 3), rng.uniform(-math.pi, math.pi, 3)])
        m = m.astype(np.float32)
        assert np.all(np.abs(m[3:]) > 1e-3)
        out = np.zeros((3, n), np.float32)
        cen = np.zeros(3, np.float32)
        move(O._fp(m), O._fp(x[0]), O._fp(x[1]), O._fp(x[2]), n, O._fp(c), O._fp(out[0]),
             O._fp(out[1]), O._fp(out[2]), O._fp(cen))
        R = _rz(m[3]) @ _ry(m[4]) @ _rx(m[5])
        want = R @ x.astype(np.float64) + (m[:3].astype(np.float64) + c)[:, None]
        np.testing.assert_allclose(out, want, atol=2e-5 * (1.0 + np.abs(want).max()))
        np.testing.assert_allclose(cen, c.astype(np.float64) + m[:3], atol=1e-5)
        wrong = _rx(m[5]) @ _ry(m[4]) @ _rz(m[3])
        bad = wrong @ x.astype(np.float64) + (m[:3].astype(np.float64) + c)[:, None]
        assert np.abs(bad - want).max() > 1e-2
