"""CPU checks that pin the oracle: the reference's golden vector, Philox
known answers, the canonical math against float64, closed forms and
invariants.  No GPU needed."""
import math
import os

import numpy as np
import pytest

import fixtures
import oracle_lib as O
from gpudock_amd import remc, synth

GOLDEN = fixtures.GOLDEN


# ---------------------------------------------------------------- golden G1
def test_g1_reference_order_fp32():
    """Initial energy line of SimilarityMatirxGen/doc/instructions.txt:159."""
    cx, _ = fixtures.load_1a07c1()
    e = O.pose_ref32(cx, [0], [0], np.zeros((1, 6)))[0]
    np.testing.assert_allclose(e, fixtures.g1_slots(), rtol=0, atol=5.001e-5)


def test_g1_fp64_and_replay():
    cx, _ = fixtures.load_1a07c1()
    z = np.zeros((1, 6))
    e64 = O.pose_ref64(cx, [0], [0], z)[0]
    rp = O.replay_energy(cx, [0], [0], z)[0]
    g = fixtures.g1_slots()
    np.testing.assert_allclose(e64, g, rtol=0, atol=1.5e-4)
    np.testing.assert_allclose(rp, g, rtol=0, atol=5.001e-5)


def test_g1_pins_term_order():
    """The slot-indexed normalisation of rem-1.2 does not give G1: the
    published line is file order (a dropped/mis-indexed term fails here)."""
    cx, _ = fixtures.load_1a07c1()
    e = O.pose_ref32(cx, [0], [0], np.zeros((1, 6)), term_order=remc.TERM_ORDER_SLOT)[0]
    assert abs(e[9] - fixtures.G1_TOTAL) > 1.0


# ---------------------------------------------------------------- golden G2
# gpudocksm-rem-1.2/reference_result.txt:60-80: raw terms of replica 0 of
# 1a07C1 along ten poses of a deterministic path with all three rotation
# angles non-zero.  The reference printed FP32 values; they agree with the
# FP64 oracle to 3e-7 of max(|E|, 1) except the distance term, whose FP32
# value carries the rounding of centre + translation (one ulp of the ~43 A
# centre coordinates, 3.8e-6).
G2_REL = 1e-6


def _g2_dst_floor(cx):
    return 2.0 * 2.0 ** -23 * float(np.max(np.abs(cx.lig_center)))


def test_g2_fp64_oracle():
    """Move (rotation about the ligand centre, ZYX), all nine raw terms, the
    LHM restraints of the full 1a07C1.ff, and the file-order weighted total,
    in float64, at every step of G2."""
    steps, mv, g2 = fixtures.load_g2()
    cx = fixtures.load_1a07c1_raw()
    z = np.zeros(len(steps), np.int32)
    e = O.pose_ref64(cx, z, z, mv)
    tol = G2_REL * np.maximum(np.abs(g2), 1.0)
    tol[:, 8] += _g2_dst_floor(cx)
    assert np.all(np.abs(e - g2) <= tol), np.abs(e - g2) / np.maximum(np.abs(g2), 1.0)


def test_g2_canonical_replay():
    """The canonical FP32 arithmetic the GPU runs (bitwise, tested on the GPU)
    reproduces G2 within 1e-6 of max(|E|, 1) on every term."""
    steps, mv, g2 = fixtures.load_g2()
    cx = fixtures.load_1a07c1_raw()
    z = np.zeros(len(steps), np.int32)
    e = O.replay_energy(cx, z, z, mv)
    assert np.all(np.abs(e - g2) <= G2_REL * np.maximum(np.abs(g2), 1.0))


def test_g2_reference_order_fp32():
    """The reference-order restatement: within 1e-5 of max(|E|, 1) plus its
    own sequential-sum rounding |E32 - E64| (its PMF sum is the least
    accurate, 2e-5 relative)."""
    steps, mv, g2 = fixtures.load_g2()
    cx = fixtures.load_1a07c1_raw()
    z = np.zeros(len(steps), np.int32)
    e32 = O.pose_ref32(cx, z, z, mv)
    e64 = O.pose_ref64(cx, z, z, mv)
    tol = 1e-5 * np.maximum(np.abs(g2), 1.0) + np.abs(e32 - e64)
    tol[:, 8] += _g2_dst_floor(cx)
    assert np.all(np.abs(e32 - g2) <= tol)


def test_g2_pins_lhm_file_and_term_order():
    """G2's LHM term comes from 1a07C1.ff (1a07C1-0.8.ff misses it by 0.06),
    and its total weights the terms in file order (slot order misses by 0.5)."""
    steps, mv, g2 = fixtures.load_g2()
    z = np.zeros(len(steps), np.int32)
    e = O.replay_energy(fixtures.load_1a07c1_raw("1a07C1-0.8.ff"), z, z, mv)
    assert np.abs(e[:, 7] - g2[:, 7]).max() > 0.03
    e = O.replay_energy(fixtures.load_1a07c1_raw(), z, z, mv, term_order=remc.TERM_ORDER_SLOT)
    assert np.abs(e[:, 9] - g2[:, 9]).min() > 0.3


def test_reference_prepared_start_pose():
    """InitLigCoord (util.C:229-247): coord_new = coord_orig + centre; the
    oracle's move at movematrix 0 must give exactly that pose."""
    _, r = fixtures.load_1a07c1()
    lna = int(r["size"][4])
    xyz = r["lig_xyz"].reshape(-1, lna, 3)[0]
    c = r["lig_center"]
    m = np.zeros(6, np.float32)
    out = np.zeros((3, lna), np.float32)
    cen = np.zeros(3, np.float32)
    ox, oy, oz = [np.ascontiguousarray(xyz[:, i]) for i in range(3)]
    O.lib().orc_move_ref32(O._fp(m), O._fp(ox), O._fp(oy), O._fp(oz), lna,
                           O._fp(np.ascontiguousarray(c)), O._fp(out[0]), O._fp(out[1]),
                           O._fp(out[2]), O._fp(cen))
    np.testing.assert_array_equal(out.T, r["lig0_new_xyz"].reshape(lna, 3))


# ----------------------------------------------------------------- Philox KAT
def _kat_rows():
    rows = []
    with open(os.path.join(GOLDEN, "philox4x32_10_kat.txt")) as f:
        for line in f:
            v = [int(t, 16) for t in line.split()]
            rows.append((v[:4], v[4:6], v[6:10]))
    return rows


def test_philox_kat():
    rows = _kat_rows()
    assert len(rows) >= 9
    for ctr, key, want in rows:
        assert O.philox(ctr, key) == want, (ctr, key)


def test_philox_published_vectors():
    """Random123's published kat_vectors for philox4x32_10 (zero, all-ones,
    pi digits) -- the first three lines of the fixture."""
    rows = _kat_rows()
    assert rows[0][2] == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert rows[1][2] == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert rows[2][2] == [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_u01_mapping():
    """(w >> 8 + 1) * 2^-24: in (0, 1], both ends exact."""
    words = np.array([0, 255, 256, 0xFFFFFFFF, 0x80000000], dtype=np.uint32)
    u = O.canon_math(3, words.view(np.float32))
    assert u[0] == 2.0 ** -24 and u[1] == 2.0 ** -24
    assert u[2] == 2.0 * 2.0 ** -24
    assert u[3] == 1.0
    assert u[4] == 0.5 + 2.0 ** -24


# ------------------------------------------------------------ canonical math
def _ulp_err(got, want):
    want32 = want.astype(np.float32)
    spacing = np.spacing(np.abs(want32)).astype(np.float64)
    return np.abs(got.astype(np.float64) - want) / spacing


def test_canonical_exp_accuracy():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.uniform(-86, 88, 200000), rng.uniform(-2, 2, 100000),
                        [0.0, -0.0, 1.0, -1.0, 88.0, -86.0]]).astype(np.float32)
    y = O.canon_math(0, x)
    assert np.max(_ulp_err(y, np.exp(x.astype(np.float64)))) < 2.0
    assert O.canon_math(0, np.array([89.0, np.inf], np.float32)).tolist() == [np.inf, np.inf]
    assert O.canon_math(0, np.array([-87.0, -np.inf], np.float32)).tolist() == [0.0, 0.0]
    assert np.isnan(O.canon_math(0, np.array([np.nan], np.float32))[0])


def test_canonical_exp_hot_accuracy():
    """The energy-loop exp (degree-5 fit): within 3 ulp on [-20, 20] (the PMF
    range) with a mean under 1 ulp, 5 ulp down to -86 (where e^x < 1e-17),
    e^-86 (its value at -86) below."""
    rng = np.random.default_rng(6)
    x = np.concatenate([rng.uniform(-20, 20, 300000), rng.uniform(-1, 1, 100000),
                        [0.0, 18.0]]).astype(np.float32)
    err = _ulp_err(O.canon_math(6, x), np.exp(x.astype(np.float64)))
    assert np.max(err) < 3.0 and np.mean(err) < 1.0
    x = rng.uniform(-86, -20, 300000).astype(np.float32)
    assert np.max(_ulp_err(O.canon_math(6, x), np.exp(x.astype(np.float64)))) < 5.0
    lo = O.canon_math(6, np.array([-86.0], np.float32))[0]
    assert 0.0 < lo < 1e-37
    assert O.canon_math(6, np.array([-86.5, -1000.0], np.float32)).tolist() == [lo, lo]


def test_canonical_log_accuracy():
    rng = np.random.default_rng(1)
    x = np.concatenate([np.exp(rng.uniform(-80, 80, 200000)), rng.uniform(0.5, 2, 100000),
                        [1.0, 2.0, 0.5, 1e-40]]).astype(np.float32)
    y = O.canon_math(1, x)
    want = np.log(x.astype(np.float64))
    err = np.abs(y - want) / np.maximum(np.spacing(np.abs(want.astype(np.float32))), 1e-45)
    assert np.max(err[np.abs(want) > 1e-3]) < 2.0
    assert np.max(np.abs(y - want)[np.abs(want) <= 1e-3]) < 1e-9
    assert O.canon_math(1, np.array([1.0], np.float32))[0] == 0.0
    assert O.canon_math(1, np.array([0.0], np.float32))[0] == -np.inf


def test_canonical_sincos_accuracy():
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(-20, 20, 200000), rng.uniform(-1, 1, 50000),
                        [0.0, math.pi / 2, -math.pi]]).astype(np.float32)
    s, c = O.canon_math(2, x)
    xd = x.astype(np.float64)
    assert np.max(np.abs(s - np.sin(xd))) < 2.5e-7
    assert np.max(np.abs(c - np.cos(xd))) < 2.5e-7
    assert O.canon_math(2, np.zeros(1, np.float32))[0][0] == 0.0
    assert O.canon_math(2, np.zeros(1, np.float32))[1][0] == 1.0


def test_canonical_rsqrt_rcp_accuracy():
    rng = np.random.default_rng(5)
    x = np.exp(rng.uniform(-60, 60, 300000)).astype(np.float32)
    x = np.concatenate([x, np.array([1.0, 2.0, 4.0, 0.25, 81.0], np.float32)])
    xd = x.astype(np.float64)
    assert np.max(_ulp_err(O.canon_math(4, x), 1.0 / np.sqrt(xd))) < 2.0
    assert np.max(_ulp_err(O.canon_math(5, x), 1.0 / xd)) < 2.0


def test_canonical_rsqrt_exhaustive_over_two_binades():
    """Every float in [1, 4): the seed and the step depend only on the
    mantissa and the exponent's parity, so this covers every positive
    normal input up to the power-of-two scaling."""
    x = np.arange(0x3F800000, 0x40800000, dtype=np.uint32).view(np.float32)
    err = _ulp_err(O.canon_math(4, x), 1.0 / np.sqrt(x.astype(np.float64)))
    assert err.max() < 1.2
    assert np.mean(err > 1.0) < 0.002
    # scaling by 4 scales the result by exactly 1/2
    xs = x[::4099]
    assert np.array_equal(O.canon_math(4, xs * np.float32(4.0)) * np.float32(2.0),
                          O.canon_math(4, xs))


# ---------------------------------------------------------- closed forms
def _single_pair_complex(d, lig_t=3, prt_t=5, cls=2, hpcflag=1, q=0.4, ele=-0.3,
                         psp_val=-0.7, kde_at=None, mcs_at=None):
    ep0, ep = synth.enepara_from_params()
    kde_xyz = np.zeros((0, 3)) if kde_at is None else np.array([kde_at])
    kde_t = np.zeros(0, int) if kde_at is None else np.array([lig_t])
    if mcs_at is None:
        mcs = np.zeros((0, 3, 3))
        tcc = np.zeros(0)
    else:
        mcs = np.full((1, 3, 3), remc.MCS_INVALID)
        mcs[0, 1] = mcs_at
        tcc = np.array([0.8])
    psp = np.zeros((24, 4))
    psp[lig_t, 2] = psp_val
    return remc.Complex(
        lig_xyz=np.zeros((1, 1, 3)), lig_t=[lig_t], lig_c=[q], lig_n=[0],
        lig_center=[0, 0, 0], prt_xyz=np.array([[[d, 0.0, 0.0]]]), prt_t=[[prt_t]],
        prt_c=[[cls]], prt_ele=[[ele]], prt_seq3r=[[2]], prt_hpcflag=[[hpcflag]],
        prt_hpp=[[0.6]], pocket_center=[0.5, 0.0, 0.0], psp=psp, kde_xyz=kde_xyz,
        kde_t=kde_t, mcs_xyz=mcs, mcs_tcc=tcc, enepara=ep), ep0, ep


@pytest.mark.parametrize("d", [2.1, 3.7, 6.0, 8.9, 11.0])
def test_single_pair_by_hand(d):
    """One ligand atom, one point: every raw term from the parameter-file
    values (EnePara0, before OptimizeEnepara) in float64 by hand."""
    lt, pt = 3, 5
    cx, ep0, ep = _single_pair_complex(d, kde_at=[0.3, 0.4, 0.0], mcs_at=[1.0, 2.0, 2.0])
    vdw0, eps = ep0.vdw[pt][lt][0], ep0.vdw[pt][lt][1]
    lj = list(ep0.lj)
    p1 = 2 * eps * (lj[2] * vdw0) ** 9 / d ** 9
    p2 = 3 * eps * (lj[2] * vdw0) ** 6 / d ** 6
    vdw = (p1 - p2) / (p1 * lj[0] * (1 + lj[1] * d * d) + 1)
    el0, el1 = ep0.el
    s = el1 * d
    g = el0 + (4 - 3 * el0) * s * s + (2 * el0 - 3) * s ** 3 if s < 1 else 1 / s
    ele = 0.4 * -0.3 * g
    r0, e0 = ep0.pmf[pt][lt][0], ep0.pmf[pt][lt][1]
    pmf = e0 / (1 + math.exp((-0.5 * d + 6) * (d - r0)))
    psp = -0.7 if d <= r0 else 0.0
    h0, hs = ep0.hdb[pt][lt][0], ep0.hdb[pt][lt][1]
    hdb = (-1 / (hs * math.sqrt(2 * math.pi)) * math.exp(-0.5 * ((d - h0) / hs) ** 2)
           if h0 > 0.1 else 0.0)
    u = (d / 9) ** 2
    hpc1 = 0.6 * (1 - 0.5 * (7 * u - 9 * u ** 2 + 5 * u ** 3 - u ** 4)) if d <= 9 else 0.0
    m0, m1 = ep0.hpl[lt][0], ep0.hpl[lt][1]
    hpc = 0.5 * ((hpc1 - m0) / m1) ** 2 - math.log(1 / (m1 * math.sqrt(2 * math.pi)))
    h = ep0.kde
    kde = math.exp(-0.5 * 0.25 / h ** 2) / (h * math.sqrt(2 * math.pi)) ** 3
    lhm = math.log(0.8 * 3.0)
    dst = 0.5
    raw = dict(vdw=vdw, ele=ele, pmf=pmf, hpc=hpc, hdb=hdb, dst=dst, psp=psp, kde=kde, lhm=lhm)
    a, b, w = list(ep0.a_para), list(ep0.b_para), list(ep0.w)
    want = np.zeros(10)
    for k, name in enumerate(remc.FILE_ORDER):
        want[remc.FILE_TO_SLOT[k]] = a[k] * raw[name] + b[k]
    want[9] = sum(w[k] * want[remc.FILE_TO_SLOT[k]] for k in range(9))
    e32 = O.pose_ref32(cx, [0], [0], np.zeros((1, 6)))[0]
    e64 = O.pose_ref64(cx, [0], [0], np.zeros((1, 6)))[0]
    rp = O.replay_energy(cx, [0], [0], np.zeros((1, 6)))[0]
    np.testing.assert_allclose(e64, want, rtol=2e-6, atol=1e-6)
    np.testing.assert_allclose(e32, want, rtol=3e-5, atol=3e-5)
    np.testing.assert_allclose(rp, want, rtol=3e-5, atol=3e-5)


def test_single_pair_class_gates():
    """psp only for class 2 inside the PMF radius; hpc only for flagged
    points within 9 A."""
    d = 2.0
    cx2, _, _ = _single_pair_complex(d, cls=2, hpcflag=1)
    cx3, _, _ = _single_pair_complex(d, cls=3, hpcflag=0)
    e2 = O.pose_ref64(cx2, [0], [0], np.zeros((1, 6)))[0]
    e3 = O.pose_ref64(cx3, [0], [0], np.zeros((1, 6)))[0]
    ep = cx2.enepara
    assert e2[3] != e3[3]            # psp slot differs
    assert e3[3] == pytest.approx(ep.b_para[6])   # psp raw 0 -> b
    assert e2[5] != e3[5]            # hpc slot differs


# ------------------------------------------------------------- invariants
def test_move_invariants():
    cx, temps, mc, cfg = synth.make_config("c0")
    rng = np.random.default_rng(3)
    lna = cx.lig_xyz.shape[1]
    ox, oy, oz = [np.ascontiguousarray(cx.lig_xyz[0, :, i]) for i in range(3)]
    for _ in range(20):
        m = np.concatenate([rng.normal(0, 2, 3), rng.uniform(-7, 7, 3)]).astype(np.float32)
        out = np.zeros((3, lna), np.float32)
        cen = np.zeros(3, np.float32)
        O.lib().orc_move_ref32(O._fp(m), O._fp(ox), O._fp(oy), O._fp(oz), lna,
                               O._fp(cx.lig_center), O._fp(out[0]), O._fp(out[1]),
                               O._fp(out[2]), O._fp(cen))
        new = out.T.astype(np.float64)
        old = cx.lig_xyz[0].astype(np.float64)
        dn = np.linalg.norm(new[:, None] - new[None], axis=2)
        do = np.linalg.norm(old[:, None] - old[None], axis=2)
        np.testing.assert_allclose(dn, do, atol=2e-5)
        np.testing.assert_allclose(new.mean(0), cen, atol=2e-5)
        # rotation recovered from the point sets is proper
        u, _, vt = np.linalg.svd((new - new.mean(0)).T @ (old - old.mean(0)))
        assert np.linalg.det(u @ vt) > 0


def test_rotation_convention_zyx():
    """R = Rz(a) Ry(b) Rx(g) (move_d.cu:97-105): a pure a-rotation turns x
    into (cos a, sin a, 0)."""
    m = np.array([0, 0, 0, 0.3, 0, 0], np.float32)
    one = np.array([1.0], np.float32)
    zero = np.array([0.0], np.float32)
    out = np.zeros((3, 1), np.float32)
    cen = np.zeros(3, np.float32)
    O.lib().orc_move_ref32(O._fp(m), O._fp(one), O._fp(zero), O._fp(zero), 1,
                           O._fp(np.zeros(3, np.float32)), O._fp(out[0]), O._fp(out[1]),
                           O._fp(out[2]), O._fp(cen))
    np.testing.assert_allclose(out[:, 0], [math.cos(0.3), math.sin(0.3), 0], atol=1e-7)
    m = np.array([0, 0, 0, 0, 0.3, 0], np.float32)
    O.lib().orc_move_ref32(O._fp(m), O._fp(one), O._fp(zero), O._fp(zero), 1,
                           O._fp(np.zeros(3, np.float32)), O._fp(out[0]), O._fp(out[1]),
                           O._fp(out[2]), O._fp(cen))
    np.testing.assert_allclose(out[:, 0], [math.cos(0.3), 0, -math.sin(0.3)], atol=1e-7)


def test_three_implementations_agree_synthetic():
    """Reference-order FP32, canonical replay and FP64 on random poses: both
    FP32 forms within 1e-5 of FP64 relative to the term scale."""
    cx, temps, mc, cfg = synth.make_config("c0")
    rng = np.random.default_rng(4)
    n = 64
    lc = rng.integers(0, 4, n)
    pc = rng.integers(0, 2, n)
    mv = np.concatenate([rng.normal(0, 0.7, (n, 3)), rng.normal(0, 1.0, (n, 3))], 1)
    a = O.pose_ref32(cx, lc, pc, mv)
    b = O.pose_ref64(cx, lc, pc, mv)
    c = O.replay_energy(cx, lc, pc, mv)
    scale = np.maximum(np.abs(b), 1.0)
    assert np.max(np.abs(a - b) / scale) < 1e-5
    assert np.max(np.abs(c - b) / scale) < 1e-5


def test_replay_deterministic():
    cx, temps, mc, cfg = synth.make_config("c0", n_lig=2, n_prt=1, n_tmp=4, steps=30)
    opt = remc.default_options(seed=7)
    r1 = O.replay_run(cx, temps, mc, opt)
    r2 = O.replay_run(cx, temps, mc, opt)
    for k in r1:
        np.testing.assert_array_equal(r1[k], r2[k])
    assert r1["acs_mc"].sum() > 0


def test_temperature_ladder_geometric():
    t = remc.temperatures(8, 0.3, 3.0)
    T = np.array([x.t for x in t])
    np.testing.assert_allclose(T[0], 0.3, rtol=1e-7)
    np.testing.assert_allclose(T[-1], 3.0, rtol=1e-5)
    ratios = T[1:] / T[:-1]
    np.testing.assert_allclose(ratios, ratios[0], rtol=1e-6)
    np.testing.assert_allclose([x.minus_beta for x in t], -1.0 / T, rtol=1e-6)
