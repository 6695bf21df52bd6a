"""Parallel decompression of variable-rate streams through a block index in
the reference's hybrid4 and hybrid8 layouts (include/zfp/index.hpp).

* A stream written by the reference library itself (serial zfp_compress,
  oracle/_ref/libzfp_ref.so), with the hybrid4 or hybrid8 index the
  reference's own classes build from that stream's block sizes, is decoded
  on the GPU with the index imported: the result equals serial
  decompression bit for bit.
* A stream written here exports its index as hybrid4 records equal to the
  reference's for the same block sizes; imported back after the library's
  index registry is cleared, it decodes the same.
"""
import ctypes
import os

import numpy as np
import pytest
import torch

import zfp_amd as Z
from zfptest import (ROOT, T_DOUBLE, T_FLOAT, api_compress, have_ref, oracle_compress, oracle_decompress,
                     ref_api, smooth_field, stream_params_accuracy, stream_params_precision)

pytestmark = pytest.mark.gpu
REF4 = os.path.join(ROOT, "oracle", "_ref", "libhybrid4_ref.so")
needs_ref = pytest.mark.skipif(not (have_ref() and os.path.exists(REF4)), reason="reference builds not present")
U64P = ctypes.POINTER(ctypes.c_uint64)
SZ4 = ctypes.c_size_t * 4
PD4 = ctypes.c_ssize_t * 4


def _ref_hybrid4(sizes):
    lib = ctypes.CDLL(REF4)
    lib.ref_hybrid4.argtypes = [U64P, ctypes.c_size_t, ctypes.c_void_p, U64P]
    n = len(sizes)
    s = np.asarray(sizes, dtype=np.uint64)
    rec = np.zeros(((n + 3) // 4) * 12, dtype=np.uint8)
    off = np.zeros(n + 1, dtype=np.uint64)
    assert lib.ref_hybrid4(s.ctypes.data_as(U64P), n, rec.ctypes.data, off.ctypes.data_as(U64P)) == 0
    return rec, off


def _ref_hybrid8(dims, sizes):
    lib = ctypes.CDLL(REF4)
    lib.ref_hybrid8.argtypes = [ctypes.c_uint, U64P, ctypes.c_size_t, U64P, U64P]
    n = len(sizes)
    s = np.asarray(sizes, dtype=np.uint64)
    words = np.zeros(2 * ((n + 7) // 8), dtype=np.uint64)
    off = np.zeros(n + 1, dtype=np.uint64)
    assert lib.ref_hybrid8(dims, s.ctypes.data_as(U64P), n, words.ctypes.data_as(U64P), off.ctypes.data_as(U64P)) == 0
    return words, off


def _decode(ztype, n, p, d_words, out):
    nn = SZ4(*(list(n) + [0] * (4 - len(n))))
    bits = ctypes.c_uint64()
    rc = Z.hip_lib().zfp_hip_decode(ztype, nn, PD4(0, 0, 0, 0), ctypes.c_void_p(out.data_ptr()), *p,
                                    ctypes.c_void_p(d_words.data_ptr()), 0, ctypes.byref(bits), None)
    return rc, bits.value


CASES = [(T_DOUBLE, [40, 36, 28], stream_params_accuracy(1e-6)), (T_FLOAT, [257, 131], stream_params_precision(18)),
         (T_DOUBLE, [20, 12, 16, 8], stream_params_precision(24))]


@needs_ref
@pytest.mark.parametrize("layout", ["hybrid4", "hybrid8"])
@pytest.mark.parametrize("k", range(len(CASES)))
def test_reference_stream_with_reference_index(k, layout):
    ztype, n, p = CASES[k]
    a = smooth_field(n, ztype, seed=40 + k)
    words, size = api_compress(ref_api(), a, 0, n, None, ztype, p)      # serial zfp, the reference library
    worst, bits, per_block = oracle_compress(a, 0, n, None, ztype, p, block_bits=True)
    sizes = [int(b) for b in per_block]
    rec, off = _ref_hybrid4(sizes) if layout == "hybrid4" else _ref_hybrid8(len(n), sizes)
    assert int(off[-1]) == bits
    # the decoder's buffer contract: room for the worst-case stream
    buf = np.zeros(worst.size, dtype=np.uint64)
    buf[:words.size] = words
    d_words = torch.from_numpy(buf.view(np.int64)).to("cuda:0")
    Z.hip_lib().zfp_hip_index_clear()
    if layout == "hybrid4":
        rc = Z.hip_lib().zfp_hip_index_import_hybrid4(ctypes.c_void_p(d_words.data_ptr()), 0, rec.ctypes.data,
                                                      len(per_block), int(off[-1]))
    else:
        rc = Z.hip_lib().zfp_hip_index_import_hybrid8(ctypes.c_void_p(d_words.data_ptr()), len(n), 0,
                                                      rec.ctypes.data_as(U64P), len(per_block), int(off[-1]))
    assert rc == 0
    assert Z.hip_lib().zfp_hip_index_size(ctypes.c_void_p(d_words.data_ptr())) == len(per_block) + 1
    ref = np.zeros_like(a)
    oracle_decompress(words, ref, 0, n, None, ztype, p)
    out = torch.zeros(a.size, dtype=torch.from_numpy(a).dtype, device="cuda:0")
    rc, dbits = _decode(ztype, n, p, d_words, out)
    assert rc == 0 and dbits == bits
    assert np.array_equal(out.cpu().numpy().view(np.uint8), ref.view(np.uint8))
    Z.hip_lib().zfp_hip_index_clear()


@needs_ref
def test_own_index_exported_as_hybrid4():
    ztype, n, p = CASES[0]
    a = smooth_field(n, ztype, seed=77)
    exp_words, bits, per_block = oracle_compress(a, 0, n, None, ztype, p, block_bits=True)
    d_words = torch.zeros(exp_words.size + 8, dtype=torch.int64, device="cuda:0")
    d_a = torch.from_numpy(a).to("cuda:0")
    nn = SZ4(*(list(n) + [0] * (4 - len(n))))
    ebits = ctypes.c_uint64()
    Z.hip_lib().zfp_hip_index_clear()
    assert Z.hip_lib().zfp_hip_encode(ztype, nn, PD4(0, 0, 0, 0), ctypes.c_void_p(d_a.data_ptr()), *p,
                                      ctypes.c_void_p(d_words.data_ptr()), 0, ctypes.byref(ebits), None) == 0
    assert ebits.value == bits
    nblocks = len(per_block)
    rec = np.zeros(((nblocks + 3) // 4) * 12, dtype=np.uint8)
    start, end = ctypes.c_uint64(), ctypes.c_uint64()
    assert Z.hip_lib().zfp_hip_index_export_hybrid4(ctypes.c_void_p(d_words.data_ptr()), rec.ctypes.data,
                                                    (nblocks + 3) // 4, ctypes.byref(start), ctypes.byref(end)) == 0
    ref_rec, ref_off = _ref_hybrid4([int(b) for b in per_block])
    assert start.value == 0 and end.value == int(ref_off[-1])
    assert np.array_equal(rec, ref_rec)
    Z.hip_lib().zfp_hip_index_clear()
    assert Z.hip_lib().zfp_hip_index_import_hybrid4(ctypes.c_void_p(d_words.data_ptr()), 0, rec.ctypes.data,
                                                    nblocks, end.value) == 0
    ref = np.zeros_like(a)
    oracle_decompress(exp_words, ref, 0, n, None, ztype, p)
    out = torch.zeros_like(d_a)
    rc, dbits = _decode(ztype, n, p, d_words, out)
    assert rc == 0 and dbits == bits
    assert np.array_equal(out.cpu().numpy().view(np.uint8), ref.view(np.uint8))
    Z.hip_lib().zfp_hip_index_clear()


def test_repeated_encodes_into_one_buffer():
    """Variable-rate fields of different sizes compressed one after another
    into one stream buffer: each time the recorded index holds that field's
    offsets (the registry hands a replaced index's buffer to the next encode)
    and decoding through it equals serial decompression."""
    ztype, p = T_DOUBLE, stream_params_accuracy(1e-5)
    shapes = [[36, 28, 20], [9, 10, 11], [44, 40, 24], [9, 10, 11], [36, 28, 20]]
    worst = max(oracle_compress(smooth_field(n, ztype, seed=1), 0, n, None, ztype, p)[0].size for n in shapes)
    d_words = torch.zeros(worst + 8, dtype=torch.int64, device="cuda:0")
    Z.hip_lib().zfp_hip_index_clear()
    for k, n in enumerate(shapes):
        a = smooth_field(n, ztype, seed=90 + k)
        exp_words, bits, per_block = oracle_compress(a, 0, n, None, ztype, p, block_bits=True)
        d_a = torch.from_numpy(a).to("cuda:0")
        nn = SZ4(*(list(n) + [0] * (4 - len(n))))
        ebits = ctypes.c_uint64()
        assert Z.hip_lib().zfp_hip_encode(ztype, nn, PD4(0, 0, 0, 0), ctypes.c_void_p(d_a.data_ptr()), *p,
                                          ctypes.c_void_p(d_words.data_ptr()), 0, ctypes.byref(ebits), None) == 0
        assert ebits.value == bits
        got = d_words.cpu().numpy().view(np.uint64)[:(bits + 63) // 64]
        assert np.array_equal(got, exp_words[:got.size])
        nb = len(per_block)
        assert Z.hip_lib().zfp_hip_index_size(ctypes.c_void_p(d_words.data_ptr())) == nb + 1
        off = np.zeros(nb + 1, dtype=np.uint64)
        assert Z.hip_lib().zfp_hip_index_export(ctypes.c_void_p(d_words.data_ptr()), off.ctypes.data_as(U64P),
                                                nb + 1) == 0
        assert np.array_equal(off, np.concatenate([[0], np.cumsum(per_block, dtype=np.uint64)]))
        ref = np.zeros_like(a)
        oracle_decompress(exp_words, ref, 0, n, None, ztype, p)
        out = torch.zeros_like(d_a)
        rc, dbits = _decode(ztype, n, p, d_words, out)
        assert rc == 0 and dbits == bits, f"shape {n}"
        assert np.array_equal(out.cpu().numpy().view(np.uint8), ref.view(np.uint8)), f"shape {n}"
    Z.hip_lib().zfp_hip_index_clear()
