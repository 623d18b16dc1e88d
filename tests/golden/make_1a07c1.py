"""Write tests/golden/1a07c1_prepared.npz from the reference's own loaders.

Runs oracle/_ref/ref_dump (built by `make -C oracle ref` from the reference
sources gpudocksm-rem-1.2/src/{load,data,util,quicksort}.C and rmsd.f) on
the 1a07C1 complex and parameter files that the reference's release
instructions use for their verification line
(SimilarityMatirxGen/doc/instructions.txt:136-159), and stores the prepared
arrays (the inputs of Run()) as a fixture.  The PSP table is stored sparse.

usage: python tests/golden/make_1a07c1.py [REFERENCE_ROOT]
"""
import os
import struct
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def read_dump(path):
    out = {}
    with open(path, "rb") as f:
        assert f.read(7) == b"GDREF1\n"
        while True:
            name = b""
            c = f.read(1)
            if not c:
                break
            while c != b"\0":
                name += c
                c = f.read(1)
            t = f.read(1).decode()
            (n,) = struct.unpack("<Q", f.read(8))
            dt = {"f": np.float32, "i": np.int32, "d": np.float64}[t]
            out[name.decode()] = np.frombuffer(f.read(n * np.dtype(dt).itemsize), dtype=dt).copy()
    return out


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    d = os.path.join(ref, "SimilarityMatirxGen", "data")
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", f"REF={ref}"],
                   check=True)
    tool = os.path.join(ROOT, "oracle", "_ref", "ref_dump")
    tmp = os.path.join(ROOT, "oracle", "_ref", "1a07c1.bin")
    subprocess.run([tool, os.path.join(d, "1a07C1", "1a07C1.sdf"),
                    os.path.join(d, "1a07C1", "1a07C.pdb"),
                    os.path.join(d, "1a07C1", "1a07C1-0.8.ff"),
                    os.path.join(d, "parameters", "paras"),
                    os.path.join(d, "parameters", "08ff_opt"),
                    os.path.join(d, "parameters", "08_nor_a"),
                    os.path.join(d, "parameters", "08_nor_b"),
                    "1a07C1", "1", "0.044", "0.036", tmp], check=True)
    r = read_dump(tmp)
    n_lig, n_prt, n_tmp, n_rep, lna, pnp, pnk, pos = (int(v) for v in r["size"])
    psp = r["psp"].reshape(24, 10000)
    nz = np.nonzero(psp)
    out = dict(size=r["size"])
    for k, v in r.items():
        if k not in ("psp", "size"):
            out[k] = v
    out["psp_t"] = nz[0].astype(np.int32)
    out["psp_r"] = nz[1].astype(np.int32)
    out["psp_v"] = psp[nz].astype(np.float32)
    # MCS fields: keep only the first lna + 2 (the loader indexes atom number + 1)
    mcs = r["mcs_xyz"].reshape(pos, 256, 3)[:, : lna + 2, :]
    out["mcs_xyz"] = np.ascontiguousarray(mcs)
    dst = os.path.join(HERE, "1a07c1_prepared.npz")
    np.savez_compressed(dst, **out)
    print("wrote", dst, os.path.getsize(dst), "bytes;", f"n_lig={n_lig} n_prt={n_prt} "
          f"lna={lna} pnp={pnp} pnk={pnk} pos={pos}")

    # SetTemperature (util.C:490-532) ladders of the reference build: the
    # minus_beta of every temperature for several ladder lengths and ranges
    ladders = {}
    for n_tmp in LADDER_SIZES:
        for lo, hi in LADDER_RANGES:
            subprocess.run([tool, os.path.join(d, "1a07C1", "1a07C1.sdf"),
                            os.path.join(d, "1a07C1", "1a07C.pdb"),
                            os.path.join(d, "1a07C1", "1a07C1-0.8.ff"),
                            os.path.join(d, "parameters", "paras"),
                            os.path.join(d, "parameters", "08ff_opt"),
                            os.path.join(d, "parameters", "08_nor_a"),
                            os.path.join(d, "parameters", "08_nor_b"),
                            "1a07C1", str(n_tmp), lo, hi, tmp], check=True,
                           stdout=subprocess.DEVNULL)
            ladders[ladder_key(n_tmp, lo, hi)] = read_dump(tmp)["temp_minus_beta"]
    dst = os.path.join(HERE, "set_temperature_ladders.npz")
    np.savez_compressed(dst, **ladders)
    print("wrote", dst, len(ladders), "ladders")


LADDER_SIZES = (1, 2, 3, 8, 16, 32, 64)
# the benchmark range, the reference's own command line (ceiling below floor)
# and a narrow range
LADDER_RANGES = (("0.3", "3.0"), ("0.044", "0.036"), ("0.02", "0.036"))


def ladder_key(n_tmp, lo, hi):
    return f"n{n_tmp}_{lo}_{hi}"


if __name__ == "__main__":
    main()
