"""The reference's compressed arrays (zfp::array1..4, include/zfp/array*.hpp)
with the GPU block codecs of include/zfp/hipcodec.hpp, against the same
arrays with their default codecs (tests/arrays/hip_arrays.cpp, built by
oracle/Makefile against the reference's headers and library): identical
compressed data byte for byte and identical values -- set() and get() block
by block and in a batch scope, element reads (cache fills), element writes
and cache flushes (write-back) -- for float and double at rates 5, 8, 16 and
32, with partial blocks; and the time of set() and get() of a 256^3 array.
"""
import os
import subprocess

import pytest

from zfptest import ROOT

pytestmark = pytest.mark.gpu
BIN = os.path.join(ROOT, "oracle", "_ref", "hip_arrays")


@pytest.mark.skipif(not os.path.exists(BIN), reason="oracle/_ref/hip_arrays not built (needs the reference tree)")
def test_reference_arrays_with_gpu_codec():
    print("running", BIN, flush=True)
    r = subprocess.run([BIN, "--time"], capture_output=True, text=True, timeout=240)
    print(r.stdout[-3000:], r.stderr[-2000:], flush=True)
    lines = r.stdout.splitlines()
    assert r.returncode == 0
    assert not [ln for ln in lines if ln.startswith("FAIL")]
    assert sum(ln.startswith("ok ") for ln in lines) == 384  # 48 cases x 8 checks
