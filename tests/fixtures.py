"""Loaders for the committed golden fixtures."""
import ctypes as C
import os

import numpy as np

from gpudock_amd import remc

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# SimilarityMatirxGen/doc/instructions.txt:157-159: initial energy state of
# replica 0 (ligand conformer 0, protein model 0, native pose) for 1a07C1.
# The printed values are the normalised terms in parameter-file order
# (vdw ele pmf hpc hdb dst psp kde lhm) followed by the weighted total.
G1_FILE_ORDER = [0.6434, -0.0369, -0.2079, 0.2210, 0.8521, -1.0402, -53.2772, -0.1593,
                 17.0834]
G1_TOTAL = 21.8245


def g1_slots():
    """G1 in slot order (vdw ele pmf psp hdb hpc kde lhm dst total)."""
    out = [0.0] * 10
    for k, v in enumerate(G1_FILE_ORDER):
        out[remc.FILE_TO_SLOT[k]] = v
    out[9] = G1_TOTAL
    return np.array(out)


def enepara_from_arrays(r):
    ep = remc.EnePara()
    for name, key in (("p1a", "ep_p1a"), ("p2a", "ep_p2a"), ("pmf0", "ep_pmf0"),
                      ("pmf1", "ep_pmf1"), ("hdb0", "ep_hdb0"), ("hdb1", "ep_hdb1"),
                      ("hdb2", "ep_hdb2")):
        C.memmove(C.addressof(getattr(ep, name)),
                  np.ascontiguousarray(r[key], np.float32).ctypes.data, 24 * 30 * 4)
    sc = r["ep_scalars"]
    ep.lj0, ep.lj1, ep.el0, ep.el1, ep.a1, ep.b1, ep.kde2, ep.kde3 = [float(v) for v in sc]
    for name, key, n in (("ele", "ep_ele", 50), ("hpp", "ep_hpp", 20), ("hpl0", "ep_hpl0", 24),
                         ("hpl1", "ep_hpl1", 24), ("hpl2", "ep_hpl2", 24), ("w", "ep_w", 10),
                         ("a_para", "ep_a", 10), ("b_para", "ep_b", 10)):
        C.memmove(C.addressof(getattr(ep, name)),
                  np.ascontiguousarray(r[key], np.float32).ctypes.data, n * 4)
    return ep


def load_1a07c1():
    """The 1a07C1 complex as prepared by the reference's own loaders."""
    r = dict(np.load(os.path.join(GOLDEN, "1a07c1_prepared.npz")))
    n_lig, n_prt, n_tmp, n_rep, lna, pnp, pnk, pos = (int(v) for v in r["size"])
    c = r["prt_c"].reshape(n_prt, pnp)
    seq = r["prt_seq3r"].reshape(n_prt, pnp)
    n_res = int(max(seq[c == 2].max() + 1, r["psp_r"].max() + 1))
    psp = np.zeros((24, n_res), np.float32)
    psp[r["psp_t"], r["psp_r"]] = r["psp_v"]
    prt_xyz = np.stack([r["prt_x"], r["prt_y"], r["prt_z"]], -1).reshape(n_prt, pnp, 3)
    cx = remc.Complex(
        lig_xyz=r["lig_xyz"].reshape(n_lig, lna, 3), lig_t=r["lig_t"], lig_c=r["lig_c"],
        lig_n=r["lig_n"], lig_center=r["lig_center"], prt_xyz=prt_xyz,
        prt_t=r["prt_t"].reshape(n_prt, pnp), prt_c=c,
        prt_ele=r["prt_ele"].reshape(n_prt, pnp), prt_seq3r=np.where(c == 2, seq, 0),
        prt_hpcflag=r["prt_c0_and_d12_or_c2"].reshape(n_prt, pnp),
        prt_hpp=r["prt_hpp"].reshape(n_prt, pnp), pocket_center=r["pocket_center"],
        psp=psp, kde_xyz=np.stack([r["kde_x"], r["kde_y"], r["kde_z"]], -1),
        kde_t=r["kde_t"], mcs_xyz=r["mcs_xyz"], mcs_tcc=r["mcs_tcc"],
        enepara=enepara_from_arrays(r))
    return cx, r


def load_g2():
    """Golden vector G2 (gpudocksm-rem-1.2/reference_result.txt:60-80):
    (steps [10], movematrices [10, 6] float32, raw terms in slot order with
    the total [10, 10] float64)."""
    a = np.loadtxt(os.path.join(GOLDEN, "g2_trajectory.txt"), comments="#")
    return a[:, 0].astype(int), a[:, 1:7].astype(np.float32), a[:, 7:17]


def load_1a07c1_raw(lhm="1a07C1.ff"):
    """The 1a07C1 complex read from its files by the library's loaders, with
    the rem-1.2 weight file and unit normalisation (a = 1, b = 0), so that
    the energy terms come out raw: the setting of G2."""
    params = os.path.join(os.path.dirname(GOLDEN), "..", "gpudock_amd", "params")
    ep0 = remc.load_enepara0(os.path.join(params, "gpudocksm.ff"),
                             os.path.join(GOLDEN, "rem12_08ff_opt"))
    for k in range(10):
        ep0.a_para[k] = 1.0
        ep0.b_para[k] = 0.0
    return remc.load_complex(os.path.join(GOLDEN, "1a07C1.sdf"),
                             os.path.join(GOLDEN, "1a07C.pdb"),
                             os.path.join(GOLDEN, lhm), "1a07C1", ep0)
