"""Write tests/golden/philox4x32_10_kat.txt from Random123's philox.h.

Builds oracle/_ref/philox_kat (make -C oracle ref) from the header as
vendored in the reference (docking/ising/src/Random123/philox.h) and stores
its output: one line per case, "c0 c1 c2 c3 k0 k1 r0 r1 r2 r3" in hex.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))

if __name__ == "__main__":
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref", f"REF={ref}"],
                   check=True)
    out = subprocess.run([os.path.join(ROOT, "oracle", "_ref", "philox_kat")], check=True,
                         capture_output=True, text=True).stdout
    with open(os.path.join(HERE, "philox4x32_10_kat.txt"), "w") as f:
        f.write(out)
    print(out)
