"""The reference-side binding printed in INTEGRATION.md compiles against the
reference's own headers (gpudocksm-rem-1.2/src/dock.h, toggle.h), so its
static_asserts on the EnePara, Temp and Replica layouts run, and it links
against libgdremc.so.  Needs the reference tree (read as headers only); it
is skipped where that tree is absent (the GPU box has none)."""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/gpudocksm-rem-1.2/src"
LIB_DIR = os.path.join(ROOT, "gpudock_amd")


def _shim_source():
    txt = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"```c\+\+\n(// run_mi355x\.C.*?)```", txt, re.S)
    assert m, "INTEGRATION.md lost its run_mi355x.C block"
    return m.group(1)


@pytest.mark.skipif(not os.path.isfile(os.path.join(REF_SRC, "dock.h")),
                    reason="reference headers (gpudocksm-rem-1.2/src/dock.h) not present")
@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_integration_shim_compiles_and_links(tmp_path):
    src = tmp_path / "run_mi355x.C"
    src.write_text(_shim_source())
    lib = os.path.join(LIB_DIR, "libgdremc.so")
    assert os.path.exists(lib), "build the library first (gpudock_amd/build.py)"
    out = tmp_path / "librun_mi355x.so"
    cmd = ["g++", "-std=c++11", "-shared", "-fPIC", "-Wall", "-Wno-format-truncation",
           "-Werror", f"-I{REF_SRC}", f"-I{os.path.join(ROOT, 'include')}", str(src),
           "-o", str(out), f"-L{LIB_DIR}", "-lgdremc", "-Wl,--no-undefined",
           f"-Wl,-rpath,{LIB_DIR}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    nm = subprocess.run(["nm", "-D", "--defined-only", str(out)], capture_output=True,
                        text=True).stdout
    assert re.search(r"\bT _Z3RunPK6LigandPK7Protein", nm), "Run() with the reference's signature"
