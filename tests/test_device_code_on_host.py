"""The codec's per-block device source (zfp_amd/csrc/codec_device.h and
codec_block.h), compiled for the host by tests/hostcodec, against the oracle.

This runs the very encode_block/decode_block templates the gfx950 kernels
instantiate (gather with padding, block-floating-point cast, lifting,
reorder, 32x32 bit-plane transposes, plane coder, bit I/O), one
block after another, so that logic errors show up without a GPU.  The GPU
parity tests remain the proof for the device build itself."""
import ctypes
import os
import zlib

import numpy as np
import pytest

from zfptest import (ROOT, T_DOUBLE, T_FLOAT, T_INT32, T_INT64, field_args, make_layout, oracle_compress,
                     oracle_decompress, pdp, smooth_field, stream_params_accuracy, stream_params_precision,
                     stream_params_rate, szp, u64p)

LIB = os.path.join(ROOT, "tests", "hostcodec", "libhostcodec.so")
pytestmark = pytest.mark.skipif(not os.path.exists(LIB), reason="tests/hostcodec not built")

_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = ctypes.CDLL(LIB)
        _lib.hostcodec_compress.restype = ctypes.c_uint64
        _lib.hostcodec_compress.argtypes = [ctypes.c_int, szp, pdp, ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint,
                                            ctypes.c_uint, ctypes.c_int, u64p, ctypes.c_uint64]
        _lib.hostcodec_decompress.restype = ctypes.c_uint64
        _lib.hostcodec_decompress.argtypes = [ctypes.c_int, szp, pdp, ctypes.c_void_p, ctypes.c_uint,
                                              ctypes.c_uint, ctypes.c_uint, ctypes.c_int, u64p, ctypes.c_uint64,
                                              ctypes.c_uint64]
    return _lib


TYPES = {"i32": T_INT32, "i64": T_INT64, "f32": T_FLOAT, "f64": T_DOUBLE}
SHAPES = {"1d": [103], "2d": [13, 11], "3d": [9, 10, 7], "4d": [5, 6, 7, 5], "3d_full": [16, 16, 8],
          # even extents: 16-byte rows along the unit-stride axis next to partial blocks
          "2d_edge": [10, 6], "3d_edge": [6, 12, 10]}
MODES = {
    "rate8": lambda t, d: stream_params_rate(8, t, d),
    "rate21": lambda t, d: stream_params_rate(21.5, t, d),
    "rate5a": lambda t, d: stream_params_rate(5, t, d, True),
    "prec16": lambda t, d: stream_params_precision(16),
    "prec28": lambda t, d: stream_params_precision(28),
    "acc1e-3": lambda t, d: stream_params_accuracy(1e-3),
    "acc1e-7": lambda t, d: stream_params_accuracy(1e-7),
    "expert": lambda t, d: (100, 600, 20, -30),
    "reversible": lambda t, d: (1, 16658, 64, -1075),
    "rev_expert": lambda t, d: (300, 2500, 40, -1075),
}


def compare(a, n, t, p, layout="as_is", offset=0):
    base, off, s = make_layout(a, n, layout)
    ew, eb = oracle_compress(base, off, n, s, t, p, offset=offset)
    hw = np.zeros_like(ew)
    nn, ss = field_args(n, s)
    ptr = base.ctypes.data + off * base.itemsize
    hb = lib().hostcodec_compress(t, nn, ss, ctypes.c_void_p(ptr), *p, hw.ctypes.data_as(u64p), offset)
    assert hb == eb
    assert np.array_equal(hw, ew)
    ref = np.zeros_like(base)
    ob = oracle_decompress(ew, ref, off, n, s, t, p, offset=offset)
    out = np.zeros_like(base)
    db = lib().hostcodec_decompress(t, nn, ss, ctypes.c_void_p(out.ctypes.data + off * base.itemsize), *p,
                                    ew.ctypes.data_as(u64p), ew.size, offset)
    # the bits decompression consumes, which equal those compression wrote
    # except for reversible streams with minbits > 1 (all-zero blocks are
    # written as one bit and read as minbits bits)
    assert db == ob
    if p[3] >= -1074 or p[0] <= 1:
        assert db == eb
    assert np.array_equal(out.view(np.uint8), ref.view(np.uint8))


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("tname", list(TYPES))
@pytest.mark.parametrize("shape", list(SHAPES))
def test_block_code_matches_oracle(shape, tname, mode):
    n = SHAPES[shape]
    t = TYPES[tname]
    a = smooth_field(n, t, seed=zlib.crc32(f"{shape}/{tname}".encode()) & 0xffff)
    compare(a, n, t, MODES[mode](t, len(n)))
    if len(n) > 1:
        compare(a, n, t, MODES[mode](t, len(n)), layout="permuted_neg", offset=77)


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("tname", ["f32", "f64"])
def test_special_values(tname, mode):
    from test_oracle_vs_reference import _special
    t = TYPES[tname]
    a = _special(t)
    compare(a, [4, 4, a.size // 16], t, MODES[mode](t, 3))


def _split_field(n, t, seed):
    """A smooth field with a constant region and a few zero blocks: constant
    and near-constant blocks code planes below 32 even at low rates, smooth
    blocks stop above it, so both passes of the split decode run."""
    a = smooth_field(n, t, seed=seed)
    v = a.reshape(n[::-1])
    v[..., :4] = v.flat[0]
    v[..., -2:] = 0
    return a


@pytest.mark.parametrize("rate", [4, 8, 16])
@pytest.mark.parametrize("tname", ["i64", "f64"])
@pytest.mark.parametrize("shape", ["1d", "2d", "3d", "3d_full"])
def test_split_decode_matches_full(shape, tname, rate):
    """The two-pass fixed-rate decode (planes 63..32 of every block, then a
    full decode of the blocks that go below plane 32) reproduces the oracle's
    reconstruction, with blocks in both passes."""
    n = SHAPES[shape]
    t = TYPES[tname]
    a = _split_field(n, t, seed=zlib.crc32(f"split/{shape}/{tname}".encode()) & 0xffff)
    p = stream_params_rate(rate, t, len(n), True)
    for layout, offset in (("as_is", 0), ("permuted_neg", 77)):
        if layout != "as_is" and len(n) == 1:
            continue
        base, off, s = make_layout(a, n, layout)
        ew, eb = oracle_compress(base, off, n, s, t, p, offset=offset)
        ref = np.zeros_like(base)
        oracle_decompress(ew, ref, off, n, s, t, p, offset=offset)
        nn, ss = field_args(n, s)
        out = np.zeros_like(base)
        nredo = ctypes.c_uint64(0)
        f = lib().hostcodec_decompress_split
        f.restype = ctypes.c_uint64
        f.argtypes = [ctypes.c_int, szp, pdp, ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_int, u64p,
                      ctypes.c_uint64, ctypes.c_uint64, ctypes.POINTER(ctypes.c_uint64)]
        bits = f(t, nn, ss, ctypes.c_void_p(out.ctypes.data + off * base.itemsize), p[1], p[2], p[3],
                 ew.ctypes.data_as(u64p), ew.size, offset, ctypes.byref(nredo))
        assert bits == eb
        nblocks = int(np.prod([(x + 3) // 4 for x in n]))
        assert nredo.value < nblocks
        if rate << (2 * len(n)) >= 128:
            assert nredo.value > 0
        assert np.array_equal(out.view(np.uint8), ref.view(np.uint8))


SLOT_RATES = [1, 2, 4, 8, 16, 32, 48]


@pytest.mark.parametrize("rate", SLOT_RATES)
@pytest.mark.parametrize("tname", list(TYPES))
@pytest.mark.parametrize("shape", ["1d", "2d", "3d", "3d_full"])
def test_slot_encoder_matches_oracle(shape, tname, rate):
    """The LDS-slot fixed-rate encoder (straight-line plane coder, write-
    through slot writer) writes serial zfp's stream, including blocks with
    special values and the permuted negative-stride layout."""
    n = SHAPES[shape]
    t = TYPES[tname]
    dims = len(n)
    bits_per_block = rate << (2 * dims)
    if bits_per_block % 64 or (t in (T_FLOAT, T_DOUBLE) and bits_per_block < (9 if t == T_FLOAT else 12)):
        pytest.skip("not a word-aligned rate")
    p = stream_params_rate(rate, t, dims)
    assert p[1] % 64 == 0
    a = smooth_field(n, t, seed=zlib.crc32(f"slot/{shape}/{tname}".encode()) & 0xffff,
                     special=0.02 if t in (T_FLOAT, T_DOUBLE) else 0.0)
    f = lib().hostcodec_compress_slot
    f.restype = ctypes.c_uint64
    f.argtypes = [ctypes.c_int, szp, pdp, ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_int, u64p,
                  ctypes.c_uint64]
    for layout, offset in (("as_is", 0), ("permuted_neg", 128)):
        if layout != "as_is" and dims == 1:
            continue
        base, off, s = make_layout(a, n, layout)
        ew, eb = oracle_compress(base, off, n, s, t, p, offset=offset)
        hw = np.zeros_like(ew)
        hw[: offset // 64] = ew[: offset // 64]
        nn, ss = field_args(n, s)
        hb = f(t, nn, ss, ctypes.c_void_p(base.ctypes.data + off * base.itemsize), p[1], p[2], p[3],
               hw.ctypes.data_as(u64p), offset)
        assert hb == eb
        assert np.array_equal(hw, ew)


@pytest.mark.parametrize("rate", [1, 4, 8, 16, 21.5, 32])
@pytest.mark.parametrize("tname", list(TYPES))
@pytest.mark.parametrize("shape", list(SHAPES))
def test_slot_reader_decoder_matches_oracle(shape, tname, rate):
    """The decoder of fixed-rate blocks staged in LDS (bit-index reads from
    the slot, no refill window) reconstructs serial zfp's values, at word-
    aligned and unaligned block offsets."""
    n = SHAPES[shape]
    t = TYPES[tname]
    p = stream_params_rate(rate, t, len(n))
    a = smooth_field(n, t, seed=zlib.crc32(f"sread/{shape}/{tname}".encode()) & 0xffff,
                     special=0.02 if t in (T_FLOAT, T_DOUBLE) else 0.0)
    f = lib().hostcodec_decompress_slot
    f.restype = ctypes.c_uint64
    f.argtypes = [ctypes.c_int, szp, pdp, ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_int, u64p,
                  ctypes.c_uint64, ctypes.c_uint64]
    for layout, offset in (("as_is", 0), ("permuted_neg", 77)):
        if layout != "as_is" and len(n) == 1:
            continue
        base, off, s = make_layout(a, n, layout)
        ew, eb = oracle_compress(base, off, n, s, t, p, offset=offset)
        ref = np.zeros_like(base)
        oracle_decompress(ew, ref, off, n, s, t, p, offset=offset)
        out = np.zeros_like(base)
        nn, ss = field_args(n, s)
        bits = f(t, nn, ss, ctypes.c_void_p(out.ctypes.data + off * base.itemsize), p[1], p[2], p[3],
                 ew.ctypes.data_as(u64p), ew.size, offset)
        assert bits == eb
        assert np.array_equal(out.view(np.uint8), ref.view(np.uint8))


CHAIN_MODES = ["prec16", "prec28", "acc1e-3", "acc1e-7", "expert", "reversible", "rev_expert"]
CHAIN_SHAPES = {"1d": [3001], "2d": [61, 67], "3d": [21, 18, 23], "4d": [9, 10, 7, 11]}


@pytest.mark.parametrize("mode", CHAIN_MODES)
@pytest.mark.parametrize("tname", list(TYPES))
@pytest.mark.parametrize("shape", list(CHAIN_SHAPES))
def test_chain_offsets_match_serial_parse(shape, tname, mode):
    """The parallel recovery of block offsets (chain_lane, the body of
    k_chain, run lane by lane with the launcher's rounds) gives exactly the
    block starts of serial zfp's stream, for segments much shorter than the
    distance at which chains started at arbitrary bits meet the true one (so
    that many rounds are needed), at bit offsets 0 and 77, with arbitrary
    words after the stream, and refuses a stream cut short."""
    n = CHAIN_SHAPES[shape]
    t = TYPES[tname]
    p = MODES[mode](t, len(n))
    a = smooth_field(n, t, seed=zlib.crc32(f"chain/{shape}/{tname}/{mode}".encode()) & 0xffff,
                     special=0.02 if t in (T_FLOAT, T_DOUBLE) else 0.0)
    f = lib().hostcodec_chain_offsets
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_int, szp, pdp, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_int, u64p,
                  ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint64, u64p]
    nn, ss = field_args(n, None)
    rng = np.random.default_rng(zlib.crc32(f"{shape}{tname}{mode}".encode()))
    for offset in (0, 77):
        words, bits, bb = oracle_compress(a, 0, n, None, t, p, offset=offset, block_bits=True)
        truth = np.concatenate([[0], np.cumsum(bb.astype(np.uint64))]).astype(np.uint64)
        assert truth[-1] == bits
        used = (offset + bits + 63) // 64
        tail = rng.integers(0, 2 ** 63, size=40, dtype=np.int64).view(np.uint64)
        tail[5:15] = 0  # a run of zero words, as in a zeroed buffer
        buf = np.concatenate([words[:used], tail])
        if offset + bits < used * 64:  # the rest of the last word is the stream's padding
            buf[used - 1] &= np.uint64((1 << ((offset + bits) & 63)) - 1) if (offset + bits) & 63 else ~np.uint64(0)
        for L in (128, 512, 4096):
            out = np.zeros(len(bb) + 1, dtype=np.uint64)
            rounds = f(t, nn, ss, p[0], p[1], p[2], p[3], buf.ctypes.data_as(u64p), buf.size, offset, L,
                       out.ctypes.data_as(u64p))
            assert rounds >= 1
            assert np.array_equal(out, truth), f"L={L} offset={offset}: first difference at block " \
                f"{int(np.argmax(out != truth))}"
        # fewer readable bits than the blocks need: refused
        if bits > 4096:
            out = np.zeros(len(bb) + 1, dtype=np.uint64)
            short = buf[: (offset + bits // 2) // 64]
            assert f(t, nn, ss, p[0], p[1], p[2], p[3], short.ctypes.data_as(u64p), short.size, offset, 512,
                     out.ctypes.data_as(u64p)) == -1
