"""Seeded synthetic complexes for the benchmark configurations.

A complex is built the way the reference's host preparation leaves it
(util.C:199-558): centroid-relative ligand conformers, per-point protein
arrays with ele / hpp / class flags, a PSP table, type-tagged KDE points and
MCS position restraints, and the EnePara tables built by gd_optimize_enepara
from the reference's own parameter files (gpudock_amd/params).

Choices the configurations leave open, fixed here:
  * point types: the first 24 of the 30 effective-point types; residue code
    d uniform over 20; class 0 for CA, 1 for PP, 2 or 3 (even odds) otherwise;
    residue number uniform over pnp/3 residues (PSP index).
  * "charges N(0, 0.3)": drawn for the pocket points (the per-point ele
    factor) and for the ligand atoms.
  * ligand: 1.5 A-bond random walk with no non-bonded pair closer than 2.5 A;
    further conformers by random torsions about random bonds under the same
    clash rule; atom types uniform over the 24 ligand types.
  * the pocket keeps every point at least 3 A from every ligand atom of every
    conformer at the start pose (LJ has 1/d^9 and no guard).
  * KDE points uniform in the pocket sphere with ligand types uniform.
  * 11 MCS restraints (as the 1a07C1 complex has), tcc U(0.5, 1), each ligand
    atom present with probability 0.7 at its start position + N(0, 1 A).
  * PSP values -U(0, 1) on 30 % of (type, residue) entries, else 0.
  * move half-widths: translation sigma -> uniform half-width in A, rotation
    degrees -> radians (the reference's move is uniform, move_d.cu:79-84).
"""
import math
import os

import numpy as np
from scipy.spatial import cKDTree

from . import remc

PARAMS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "params")

CONFIGS = {
    # plain REMC: one small complex, 64 replicas, 1,000 steps
    "c0": dict(pnp=300, radius=10.0, lna=24, n_lig=4, n_prt=2, pnk=300, n_tmp=8,
               t_lo=0.3, t_hi=3.0, steps=1000, spe=10, trans=0.2, rot_deg=3.0),
    # single large complex: 16,384 replicas, 5,000 steps
    "c1": dict(pnp=800, radius=14.0, lna=48, n_lig=64, n_prt=16, pnk=2000, n_tmp=16,
               t_lo=0.3, t_hi=3.0, steps=5000, spe=10, trans=0.2, rot_deg=3.0),
    # large-pocket stress: 4,000-point pocket, 128-atom ligand, 4,096 replicas
    "c4": dict(pnp=4000, radius=25.0, lna=128, n_lig=128, n_prt=1, pnk=5000, n_tmp=32,
               t_lo=0.3, t_hi=3.0, steps=2000, spe=10, trans=0.2, rot_deg=3.0),
}


def enepara_from_params():
    ep0 = remc.load_enepara0(os.path.join(PARAMS, "gpudocksm.ff"),
                             os.path.join(PARAMS, "08ff_opt"),
                             os.path.join(PARAMS, "08_nor_a"),
                             os.path.join(PARAMS, "08_nor_b"))
    return ep0, remc.optimize_enepara(ep0)


def _random_walk(rng, lna, bond=1.5, clash=2.5):
    for _ in range(1000):
        x = np.zeros((lna, 3))
        ok = True
        for i in range(1, lna):
            for _ in range(200):
                v = rng.normal(size=3)
                v *= bond / np.linalg.norm(v)
                cand = x[i - 1] + v
                if i < 2 or np.min(np.linalg.norm(x[: i - 1] - cand, axis=1)) >= clash:
                    x[i] = cand
                    break
            else:
                ok = False
                break
        if ok:
            return x
    raise RuntimeError("random walk failed")


def _min_nonbonded(x):
    d = np.linalg.norm(x[:, None, :] - x[None, :, :], axis=2)
    n = len(x)
    mask = np.abs(np.arange(n)[:, None] - np.arange(n)[None, :]) >= 2
    return d[mask].min() if mask.any() else np.inf


def _torsion_conformer(rng, x, clash=2.5, n_tors=3):
    n = len(x)
    if n < 4:
        return x.copy()
    for _ in range(1000):
        y = x.copy()
        for _ in range(n_tors):
            k = int(rng.integers(1, n - 2))
            axis = y[k] - y[k - 1]
            axis /= np.linalg.norm(axis)
            ang = rng.uniform(-math.pi, math.pi)
            c, s = math.cos(ang), math.sin(ang)
            K = np.array([[0, -axis[2], axis[1]], [axis[2], 0, -axis[0]],
                          [-axis[1], axis[0], 0]])
            R = np.eye(3) + s * K + (1 - c) * (K @ K)
            y[k + 1:] = (y[k + 1:] - y[k]) @ R.T + y[k]
        if _min_nonbonded(y) >= clash:
            return y
    return x.copy()


def _in_sphere(rng, n, radius):
    v = rng.normal(size=(n, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    r = radius * rng.uniform(size=n) ** (1.0 / 3.0)
    return v * r[:, None]


def make_complex(pnp, radius, lna, n_lig, n_prt, pnk, seed=1, pos=11, enepara=None,
                 center=(0.0, 0.0, 0.0), **_unused):
    """Build one synthetic prepared complex (remc.Complex).  radius None:
    pocket_radius(pnp), or more when the ligand ensemble needs it."""
    rng = np.random.default_rng(seed)
    if enepara is None:
        ep0, ep = enepara_from_params()
    else:
        ep0, ep = enepara
    hpp_tab = np.array(ep0.hpp, dtype=np.float32)
    center = np.asarray(center, dtype=np.float64)

    # ligand conformers, centroid-relative
    base = _random_walk(rng, lna)
    confs = [base] + [_torsion_conformer(rng, base) for _ in range(n_lig - 1)]
    confs = np.stack([c - c.mean(axis=0) for c in confs])
    lig_t = rng.integers(0, 24, size=lna)
    lig_c = rng.normal(0.0, 0.3, size=lna)

    # pocket: uniform in the sphere, >= 3 A from every start-pose ligand atom
    allatoms = confs.reshape(-1, 3)
    if radius is None:
        # pocket_radius(pnp), widened so that the ligand ensemble (plus the
        # 3 A exclusion) leaves room for every pocket point
        r_lig = float(np.max(np.linalg.norm(allatoms, axis=1)))
        radius = max(pocket_radius(pnp), r_lig + 5.0)
    tree = cKDTree(allatoms)
    pts = np.zeros((0, 3))
    for _ in range(200):
        if len(pts) >= pnp:
            break
        cand = _in_sphere(rng, 4 * pnp, radius)
        d, _ = tree.query(cand, k=1)
        pts = np.concatenate([pts, cand[d >= 3.0]])
    else:
        raise RuntimeError(f"pocket of {pnp} points does not fit a {radius:.1f} A sphere "
                           "around the ligand")
    pts = pts[:pnp]
    prt_t = rng.integers(0, 24, size=pnp)
    d_code = rng.integers(0, 20, size=pnp)
    cls = np.where(prt_t == 0, 0, np.where(prt_t == 1, 1, 2 + rng.integers(0, 2, size=pnp)))
    n_res = max(1, pnp // 3)
    seq = rng.integers(0, n_res, size=pnp)
    ele = rng.normal(0.0, 0.3, size=pnp)
    hpcflag = ((cls == 0) & (d_code == 12)) | (cls == 2)
    hpp = hpp_tab[d_code]
    prt_xyz = np.stack([pts + (rng.normal(0.0, 0.5, size=pts.shape) if i else 0.0)
                        for i in range(n_prt)]) + center
    tile = lambda a: np.tile(a, (n_prt, 1))

    psp = np.where(rng.uniform(size=(24, n_res)) < 0.3, -rng.uniform(size=(24, n_res)), 0.0)

    kde_xyz = _in_sphere(rng, pnk, radius) + center
    kde_t = rng.integers(0, 24, size=pnk)

    n_field = lna + 2
    mcs = np.full((pos, n_field, 3), remc.MCS_INVALID, dtype=np.float32)
    for m in range(pos):
        present = rng.uniform(size=lna) < 0.7
        if not present.any():
            present[int(rng.integers(0, lna))] = True
        tgt = confs[0] + center + rng.normal(0.0, 1.0, size=(lna, 3))
        for a in np.nonzero(present)[0]:
            mcs[m, a + 1] = tgt[a]
    tcc = rng.uniform(0.5, 1.0, size=pos)

    return remc.Complex(
        lig_xyz=confs, lig_t=lig_t, lig_c=lig_c, lig_n=np.arange(lna),
        lig_center=center, prt_xyz=prt_xyz, prt_t=tile(prt_t), prt_c=tile(cls),
        prt_ele=tile(ele), prt_seq3r=tile(seq), prt_hpcflag=tile(hpcflag.astype(np.int32)),
        prt_hpp=tile(hpp), pocket_center=center, psp=psp, kde_xyz=kde_xyz, kde_t=kde_t,
        mcs_xyz=mcs, mcs_tcc=tcc, enepara=ep)


def make_config(name, seed=1, **over):
    """(complex, temps, mcpara, cfg) for a named configuration; keyword
    overrides change any size (used for reduced parity cases)."""
    cfg = dict(CONFIGS[name])
    cfg.update(over)
    cx = make_complex(seed=seed, **cfg)
    temps = remc.temperatures(cfg["n_tmp"], cfg["t_lo"], cfg["t_hi"])
    scale = [cfg["trans"]] * 3 + [math.radians(cfg["rot_deg"])] * 3
    mc = remc.mcpara(cfg["steps"], cfg["spe"], scale)
    return cx, temps, mc, cfg


def subset(cx, n_lig, n_prt):
    """The first n_lig ligand and n_prt protein conformers of a complex."""
    return remc.Complex(
        lig_xyz=cx.lig_xyz[:n_lig], lig_t=cx.lig_t, lig_c=cx.lig_c, lig_n=cx.lig_n,
        lig_center=cx.lig_center, prt_xyz=cx.prt_xyz[:n_prt], prt_t=cx.prt_t[:n_prt],
        prt_c=cx.prt_c[:n_prt], prt_ele=cx.prt_ele[:n_prt], prt_seq3r=cx.prt_seq3r[:n_prt],
        prt_hpcflag=cx.prt_hpcflag[:n_prt], prt_hpp=cx.prt_hpp[:n_prt],
        pocket_center=cx.pocket_center, psp=cx.psp, kde_xyz=cx.kde_xyz, kde_t=cx.kde_t,
        mcs_xyz=cx.mcs_xyz, mcs_tcc=cx.mcs_tcc, enepara=cx.enepara)


def pocket_radius(pnp, density=0.07):
    """Sphere radius holding pnp points at the density of the single-complex
    configurations (300 points in 10 A, 800 in 14 A: about 0.07 per A^3)."""
    return (3.0 * pnp / (4.0 * math.pi * density)) ** (1.0 / 3.0)


def make_batch_jobs(n=256, seed=1, steps=2000, spe=10):
    """Batched uneven workload: n complexes with ligand atoms U[12, 96],
    pocket points U[200, 1500], ligand conformers U[4, 64], 8 or 16
    temperatures (0.3-3.0), one protein conformer, steps MC steps, exchange
    every spe steps.  Left open by the definition and fixed here: KDE points
    = pocket points; pocket radius from pocket_radius(), widened around
    large ligands (make_complex with radius None); move half-widths
    0.2 A / 3 deg as in the single-complex runs; Philox key (seed, index)."""
    rng = np.random.default_rng(seed)
    ep = enepara_from_params()
    scale = [0.2] * 3 + [math.radians(3.0)] * 3
    jobs = []
    for i in range(n):
        lna = int(rng.integers(12, 97))
        pnp = int(rng.integers(200, 1501))
        n_lig = int(rng.integers(4, 65))
        n_tmp = int(rng.choice([8, 16]))
        cx = make_complex(pnp=pnp, radius=None, lna=lna, n_lig=n_lig, n_prt=1,
                          pnk=pnp, seed=seed * 100003 + i, enepara=ep)
        temps = remc.temperatures(n_tmp, 0.3, 3.0)
        mc = remc.mcpara(steps, spe, scale)
        opt = remc.default_options(seed=seed, complex_id=i)
        jobs.append(remc.JobSpec(cx, temps, mc, opt))
    return jobs


def make_serving_pool(n=256, seed=2, steps=1000, spe=10, sigma=0.5):
    """Complexes of the serving workload: ligand atoms lognormal (median 32,
    clipped to [10, 120]), pocket points lognormal (median 600, clipped to
    [150, 2000]), 16 ligand conformers x 8 temperatures, one protein
    conformer, steps MC steps.  Left open and fixed here: lognormal sigma
    0.5, KDE points = pocket points, pocket radius as in make_batch_jobs.
    Returns [(complex, temps, mcpara)]; each served job draws one and gets
    its own Philox key."""
    rng = np.random.default_rng(seed)
    ep = enepara_from_params()
    scale = [0.2] * 3 + [math.radians(3.0)] * 3
    temps = remc.temperatures(8, 0.3, 3.0)
    pool = []
    for i in range(n):
        lna = int(np.clip(round(rng.lognormal(math.log(32.0), sigma)), 10, 120))
        pnp = int(np.clip(round(rng.lognormal(math.log(600.0), sigma)), 150, 2000))
        cx = make_complex(pnp=pnp, radius=None, lna=lna, n_lig=16, n_prt=1,
                          pnk=pnp, seed=seed * 100003 + i, enepara=ep)
        pool.append((cx, temps, remc.mcpara(steps, spe, scale)))
    return pool
