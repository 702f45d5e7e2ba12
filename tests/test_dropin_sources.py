"""The C++ drop-in claim: the reference's own BS0-BS8 driver sources compile unchanged against this
library's libp headers (which replace the reference's include/ and OCCA).  Runs where the reference
sources are present; compiles with -fsyntax-only, so nothing is built from them."""
import glob
import os
import subprocess

import pytest

REF = "/root/reference/BS"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference sources not present here")
@pytest.mark.parametrize("bs", [f"BS{i}" for i in range(9)])
def test_reference_driver_compiles_against_libp(bs):
    d = os.path.join(REF, bs)
    files = sorted(glob.glob(os.path.join(d, "*.cpp")) + glob.glob(os.path.join(d, "src", "*.cpp")))
    assert files
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-fopenmp", '-DSTREAM_DIR="/nonexistent"', f"-I{d}",
           f"-I{os.path.join(ROOT, 'streamparanumal_amd', 'libp', 'include')}", f"-I{os.path.join(ROOT, 'include')}",
           *files]
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-3000:]
