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
from zfptest import (ROOT, T_DOUBLE, T_FLOAT, api_compress, have_ref, header_prefix, oracle_compress,
                     oracle_decompress, ref_api, smooth_field, stream_params_accuracy, stream_params_precision)

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
                                    ctypes.c_void_p(d_words.data_ptr()), d_words.numel(), 0, ctypes.byref(bits),
                                    None)
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


def _encode(ztype, n, p, d_a, d_words):
    nn = SZ4(*(list(n) + [0] * (4 - len(n))))
    ebits = ctypes.c_uint64()
    assert Z.hip_lib().zfp_hip_encode(ztype, nn, PD4(0, 0, 0, 0), ctypes.c_void_p(d_a.data_ptr()), *p,
                                      ctypes.c_void_p(d_words.data_ptr()), 0, ctypes.byref(ebits), None) == 0
    return ebits.value


@pytest.mark.parametrize("mode", ["prec16", "acc1e-3"])
def test_stale_index_after_buffer_overwrite(mode):
    """Field A is compressed into buffer X (its block index is recorded under
    X's address); X is then overwritten with serial zfp's stream of a
    different field B of the same shape and mode.  Decoding X must give
    serial zfp's decompression of B: the recorded index does not describe
    X's contents any more, every block's length is checked against it, and
    the mismatch sends the decode to a parse of the stream."""
    ztype, n = T_FLOAT, [61, 47]
    p = stream_params_precision(16) if mode == "prec16" else stream_params_accuracy(1e-3)
    a = smooth_field(n, ztype, seed=501)
    b = smooth_field(n, ztype, seed=502, noise=0.3)
    wb, bits_b = oracle_compress(b, 0, n, None, ztype, p)
    cap = max(wb.size, oracle_compress(a, 0, n, None, ztype, p)[0].size)
    d_words = torch.zeros(cap, dtype=torch.int64, device="cuda:0")
    Z.hip_lib().zfp_hip_index_clear()
    _encode(ztype, n, p, torch.from_numpy(a).to("cuda:0"), d_words)
    assert Z.hip_lib().zfp_hip_index_size(ctypes.c_void_p(d_words.data_ptr())) > 0
    d_words.copy_(torch.from_numpy(np.concatenate([wb, np.zeros(cap - wb.size, np.uint64)]).view(np.int64)))
    ref = np.zeros_like(b)
    oracle_decompress(wb, ref, 0, n, None, ztype, p)
    out = torch.zeros(b.size, dtype=torch.float32, device="cuda:0")
    rc, dbits = _decode(ztype, n, p, d_words, out)
    assert rc == 0 and dbits == bits_b
    assert np.array_equal(out.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    # the stale entry is gone; decoding again (now by a parse) agrees
    assert Z.hip_lib().zfp_hip_index_size(ctypes.c_void_p(d_words.data_ptr())) == 0
    out.zero_()
    rc, dbits = _decode(ztype, n, p, d_words, out)
    assert rc == 0 and dbits == bits_b
    assert np.array_equal(out.cpu().numpy().view(np.uint32), ref.view(np.uint32))
    Z.hip_lib().zfp_hip_index_clear()


def test_stale_index_through_zfp_compress():
    """The same through zfp_compress / zfp_decompress with zfp_exec_hip on a
    device stream buffer (the index is keyed by the bitstream's begin)."""
    from hiputil import Run
    ztype, n, p = T_DOUBLE, [18, 14, 10], stream_params_precision(16)
    a = smooth_field(n, ztype, seed=511)
    b = smooth_field(n, ztype, seed=512, noise=0.2)
    Z.hip_lib().zfp_hip_index_clear()
    run = Run(a, 0, n, [0, 0, 0], ztype, p)
    try:
        assert run.compress() > 0
        wb, bits_b = oracle_compress(b, 0, n, None, ztype, p)
        nwb = (bits_b + 63) // 64
        run.dbuf.zero_()
        run.dbuf[:nwb].copy_(torch.from_numpy(wb[:nwb].view(np.int64)))
        ref = np.zeros_like(b)
        oracle_decompress(wb, ref, 0, n, None, ztype, p)
        consumed, out = run.decompress(np.zeros_like(b))
        assert consumed == nwb * 8
        assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))
    finally:
        run.close()
        Z.hip_lib().zfp_hip_index_clear()


def test_zfpy_roundtrips_of_different_streams_in_a_loop():
    """zfpy.compress_numpy(A) then decompress_numpy(stream of B), repeatedly:
    NumPy hands freed host buffers out again, so a decompression buffer often
    sits where an earlier compression wrote a stream of the same shape and
    mode.  Every result equals serial zfp's decompression of its own stream."""
    from zfp_amd import zfpy
    shape = (40, 52)
    Z.hip_lib().zfp_hip_index_clear()
    fields = [smooth_field(list(reversed(shape)), T_DOUBLE, seed=530 + k, noise=0.05 * k).reshape(shape)
              for k in range(4)]
    streams = [zfpy.compress_numpy(f, precision=16) for f in fields]
    for it in range(6):
        for k, f in enumerate(fields):
            zfpy.compress_numpy(fields[(k + 1 + it) % 4], precision=16)   # registers an index somewhere
            got = zfpy.decompress_numpy(streams[k])
            ref = np.zeros(f.size)
            words = np.frombuffer(streams[k] + b"\0" * (-len(streams[k]) % 8), dtype=np.uint64)
            _, hbits = header_prefix(T_DOUBLE, list(reversed(shape)), stream_params_precision(16))
            oracle_decompress(words, ref, 0, list(reversed(shape)), None, T_DOUBLE, stream_params_precision(16),
                              offset=hbits)
            assert np.array_equal(got.reshape(-1).view(np.uint64), ref.view(np.uint64)), (it, k)
    Z.hip_lib().zfp_hip_index_clear()


def test_exact_length_buffer():
    """The decoder reads no word past nwords: a stream held in exactly its own
    ceil(bits / 64) words decodes like one in a worst-case buffer."""
    ztype, n, p = T_FLOAT, [33, 29], stream_params_accuracy(1e-4)
    a = smooth_field(n, ztype, seed=540)
    wa, bits = oracle_compress(a, 0, n, None, ztype, p)
    nw = (bits + 63) // 64
    d_words = torch.from_numpy(wa[:nw].copy().view(np.int64)).to("cuda:0")
    ref = np.zeros_like(a)
    oracle_decompress(wa, ref, 0, n, None, ztype, p)
    out = torch.zeros(a.size, dtype=torch.float32, device="cuda:0")
    Z.hip_lib().zfp_hip_index_clear()
    rc, dbits = _decode(ztype, n, p, d_words, out)
    assert rc == 0 and dbits == bits
    assert np.array_equal(out.cpu().numpy().view(np.uint32), ref.view(np.uint32))


def test_index_registry_is_bounded():
    """Indexes recorded under many host stream buffers (zfp_compress into host
    memory): the registry keeps the 64 most recently used and frees the rest;
    an evicted stream still decodes (by a parse).  Device buffers are kept in
    far larger numbers (a batch of 1024 arrays compressed one call each keeps
    all 1024)."""
    from hiputil import Run
    ztype, n, p = T_FLOAT, [16, 12], stream_params_precision(12)
    a = smooth_field(n, ztype, seed=550)
    wa, bits = oracle_compress(a, 0, n, None, ztype, p)
    Z.hip_lib().zfp_hip_index_clear()
    runs = [Run(a, 0, n, [0, 0], ztype, p, device_stream=False) for _ in range(70)]
    try:
        for r in runs:
            assert r.compress() == (bits + 63) // 64 * 8
        keys = [ctypes.c_void_p(r.hbuf.ctypes.data) for r in runs]
        sizes = [Z.hip_lib().zfp_hip_index_size(k) for k in keys]
        assert sum(1 for s in sizes if s) == 64
        assert all(s == 0 for s in sizes[:6]) and all(s > 0 for s in sizes[6:])
        ref = np.zeros_like(a)
        oracle_decompress(wa, ref, 0, n, None, ztype, p)
        for r in (runs[0], runs[-1]):
            consumed, out = r.decompress(np.zeros_like(a))
            assert consumed == (bits + 63) // 64 * 8
            assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))
    finally:
        for r in runs:
            r.close()
    # device buffers: 200 kept
    d_a = torch.from_numpy(a).to("cuda:0")
    bufs = [torch.zeros(wa.size, dtype=torch.int64, device="cuda:0") for _ in range(200)]
    for w in bufs:
        assert _encode(ztype, n, p, d_a, w) == bits
    assert all(Z.hip_lib().zfp_hip_index_size(ctypes.c_void_p(w.data_ptr())) for w in bufs)
    Z.hip_lib().zfp_hip_index_clear()
