"""The hybrid4 and hybrid8<dims> block-index layouts of the reference's
variable-rate compressed arrays (include/zfp/index.hpp), as exported and
imported by include/zfp_hip.h, against the reference's own index classes
compiled from its header (oracle/_ref/libhybrid4_ref.so): the same records
and words byte for byte, the same offsets back, the same refusals.
Host-only conversions: no GPU needed."""
import ctypes
import os

import numpy as np
import pytest

import zfp_amd as Z
from zfptest import ROOT

REF = os.path.join(ROOT, "oracle", "_ref", "libhybrid4_ref.so")
needs_ref = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref/libhybrid4_ref.so not built")
U64P = ctypes.POINTER(ctypes.c_uint64)


def _ref(sizes):
    lib = ctypes.CDLL(REF)
    lib.ref_hybrid4.restype = ctypes.c_int
    lib.ref_hybrid4.argtypes = [U64P, ctypes.c_size_t, ctypes.c_void_p, U64P]
    n = len(sizes)
    s = np.asarray(sizes, dtype=np.uint64)
    rec = np.zeros(((n + 3) // 4) * 12, dtype=np.uint8)
    off = np.zeros(n + 1, dtype=np.uint64)
    rc = lib.ref_hybrid4(s.ctypes.data_as(U64P), n, rec.ctypes.data, off.ctypes.data_as(U64P))
    return rc, rec, off


def _ours(offsets):
    n = offsets.size - 1
    rec = np.zeros(((n + 3) // 4) * 12, dtype=np.uint8)
    rc = Z.hip_lib().zfp_hip_hybrid4_from_offsets(offsets.ctypes.data_as(U64P), n, rec.ctypes.data)
    return rc, rec


SEQS = {
    "one": [7],
    "four": [100, 2000, 3, 16658],
    "ragged": [12] * 5 + [0, 0, 1] + [9000] * 3,
    "zeros": [0] * 9,
    "max": [16658] * 13,
}


@needs_ref
@pytest.mark.parametrize("name", list(SEQS) + ["random"])
def test_records_equal_the_reference(name):
    rng = np.random.default_rng(3)
    sizes = SEQS[name] if name != "random" else list(rng.integers(0, 16659, size=1001))
    rc, rec, off = _ref(sizes)
    assert rc == 0
    # offsets as the reference's index reports them
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(np.asarray(sizes, dtype=np.uint64))]))
    rc, ours = _ours(off)
    assert rc == 0
    assert np.array_equal(ours, rec)
    # and back
    back = np.zeros(len(sizes) + 1, dtype=np.uint64)
    rc = Z.hip_lib().zfp_hip_offsets_from_hybrid4(rec.ctypes.data, len(sizes), int(off[-1]), back.ctypes.data_as(U64P))
    assert rc == 0 and np.array_equal(back, off)


@needs_ref
def test_block_too_long_is_refused_by_both():
    sizes = [10, 16659, 3]
    rc, _, _ = _ref(sizes)
    assert rc == 1
    off = np.concatenate([[0], np.cumsum(np.asarray(sizes, dtype=np.uint64))])
    assert _ours(off)[0] == Z.ZFP_HIP_EINVAL


def test_offsets_relative_to_the_first_block():
    """Absolute stream positions (after a header) give the same records as
    offsets counted from the first block."""
    sizes = np.array([5, 70, 300, 1, 44, 9], dtype=np.uint64)
    rel = np.concatenate([[0], np.cumsum(sizes)]).astype(np.uint64)
    a = _ours(rel)
    b = _ours(rel + np.uint64(148))
    assert a[0] == b[0] == 0 and np.array_equal(a[1], b[1])


def test_decreasing_records_are_refused():
    rec = np.zeros(12, dtype=np.uint8)
    r = rec.view(np.uint32)
    r[0] = 0
    rec[4:12].view(np.uint16)[:] = [0, 50, 20, 60]
    out = np.zeros(5, dtype=np.uint64)
    assert Z.hip_lib().zfp_hip_offsets_from_hybrid4(rec.ctypes.data, 4, 80, out.ctypes.data_as(U64P)) == Z.ZFP_HIP_EINVAL


# ---------------------------------------------------------------------------
# hybrid8<dims>
# ---------------------------------------------------------------------------

def _ref8(dims, sizes):
    lib = ctypes.CDLL(REF)
    lib.ref_hybrid8.argtypes = [ctypes.c_uint, U64P, ctypes.c_size_t, U64P, U64P]
    n = len(sizes)
    s = np.asarray(sizes, dtype=np.uint64)
    words = np.zeros(2 * ((n + 7) // 8), dtype=np.uint64)
    off = np.zeros(n + 1, dtype=np.uint64)
    rc = lib.ref_hybrid8(dims, s.ctypes.data_as(U64P), n, words.ctypes.data_as(U64P), off.ctypes.data_as(U64P))
    return rc, words, off


@needs_ref
@pytest.mark.parametrize("dims", [1, 2, 3, 4])
@pytest.mark.parametrize("shape", ["one", "ragged", "random", "max"])
def test_hybrid8_words_equal_the_reference(dims, shape):
    limit = 1 << (8 + 2 * (dims - 1))     # block sizes below this
    rng = np.random.default_rng(dims)
    sizes = {"one": [limit - 1], "ragged": [0, 3, limit - 1, 7, 0, 0, 255, 256 % limit, 1, 2, 9],
             "random": list(rng.integers(0, limit, size=1003)), "max": [limit - 1] * 17}[shape]
    rc, ref_words, off = _ref8(dims, sizes)
    assert rc == 0
    assert np.array_equal(off, np.concatenate([[0], np.cumsum(np.asarray(sizes, dtype=np.uint64))]))
    words = np.zeros_like(ref_words)
    assert Z.hip_lib().zfp_hip_hybrid8_from_offsets(dims, off.ctypes.data_as(U64P), len(sizes),
                                                    words.ctypes.data_as(U64P)) == 0
    assert np.array_equal(words, ref_words)
    back = np.zeros_like(off)
    assert Z.hip_lib().zfp_hip_offsets_from_hybrid8(dims, ref_words.ctypes.data_as(U64P), len(sizes), int(off[-1]),
                                                    back.ctypes.data_as(U64P)) == 0
    assert np.array_equal(back, off)


@needs_ref
@pytest.mark.parametrize("dims", [1, 2, 3, 4])
def test_hybrid8_block_too_long_is_refused_by_both(dims):
    sizes = [5, 1 << (8 + 2 * (dims - 1)), 5]
    rc, _, _ = _ref8(dims, sizes)
    assert rc == 1
    off = np.concatenate([[0], np.cumsum(np.asarray(sizes, dtype=np.uint64))])
    words = np.zeros(2, dtype=np.uint64)
    assert Z.hip_lib().zfp_hip_hybrid8_from_offsets(dims, off.ctypes.data_as(U64P), 3,
                                                    words.ctypes.data_as(U64P)) == Z.ZFP_HIP_EINVAL
