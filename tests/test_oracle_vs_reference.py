"""The oracle against the reference library itself (oracle/_ref/libzfp_ref.so,
built from the reference sources by oracle/Makefile): bit-exact streams and
reconstructions of serial zfp_compress / zfp_decompress on seeded smooth
fields, special-value blocks, every mode (reversible included), the four
memory layouts, header-prefixed streams, and the reference's OpenMP policy,
whose streams must equal the serial ones (docs/source/faq.rst Q23)."""
import zlib

import numpy as np
import pytest

from zfptest import (T_DOUBLE, T_FLOAT, T_INT32, T_INT64, ZFP_MIN_EXP, api_compress, api_decompress, have_ref,
                     make_layout, oracle_decompress, oracle_stream, read_layout, ref_api, smooth_field,
                     stream_params_accuracy, stream_params_precision, stream_params_rate)

pytestmark = pytest.mark.skipif(not have_ref(), reason="oracle/_ref (reference build) not present")

TYPES = {"i32": T_INT32, "i64": T_INT64, "f32": T_FLOAT, "f64": T_DOUBLE}
SHAPES = {"1d": [37], "2d": [9, 7], "3d": [7, 6, 5], "4d": [5, 4, 6, 3]}
MODES = {
    "rate7": lambda t, d: stream_params_rate(7, t, d),
    "rate24a": lambda t, d: stream_params_rate(24, t, d, align=True),
    "prec12": lambda t, d: stream_params_precision(12),
    "prec64": lambda t, d: stream_params_precision(64),
    "acc1e-2": lambda t, d: stream_params_accuracy(1e-2),
    "acc1e-9": lambda t, d: stream_params_accuracy(1e-9),
    "expert": lambda t, d: (40, 300, 30, -20),
    "reversible": lambda t, d: (1, 16658, 64, ZFP_MIN_EXP - 1),
}


def compare(a, n, ztype, params, layout="as_is", header=0):
    api = ref_api()
    base, off, s = make_layout(a, n, layout)
    ref_words, ref_size = api_compress(api, base, off, n, s, ztype, params, header)
    words, size = oracle_stream(base, off, n, s, ztype, params, header)
    assert size == ref_size
    assert np.array_equal(words, ref_words)
    ref_out = np.zeros_like(base)
    assert api_decompress(api, ref_words, ref_out, off, n, s, ztype, params, header) == ref_size
    out = np.zeros_like(base)
    hbits = 0
    if header:
        from zfptest import header_prefix
        _, hbits = header_prefix(ztype, n, params, header)
    oracle_decompress(words, out, off, n, s, ztype, params, offset=hbits)
    assert np.array_equal(read_layout(out, off, n, s).view(np.uint8),
                          read_layout(ref_out, off, n, s).view(np.uint8))


@pytest.mark.parametrize("tname", list(TYPES))
def test_reversible_with_minbits(tname):
    """Reversible mode with minbits > 1: an all-zero float block is written
    as one bit without padding but read as minbits bits
    (src/template/revencodef.c:64-68 against revdecodef.c:24-59), so the
    reference's zfp_decompress stops at a different point than
    zfp_compress.  The oracle must reproduce both sides: the stream, the
    decoded values and where decompression stops."""
    t = TYPES[tname]
    n = [8, 8, 12]
    a = smooth_field(n, t, seed=31, special=0.05)
    a[: a.size // 3] = 0
    params = (300, 2500, 40, ZFP_MIN_EXP - 1)
    api = ref_api()
    base, off, s = make_layout(a, n, "as_is")
    ref_words, ref_size = api_compress(api, base, off, n, s, t, params)
    words, size = oracle_stream(base, off, n, s, t, params)
    assert size == ref_size and np.array_equal(words, ref_words)
    ref_out = np.zeros_like(base)
    ref_dsize = api_decompress(api, ref_words, ref_out, off, n, s, t, params)
    out = np.zeros_like(base)
    bits = oracle_decompress(words, out, off, n, s, t, params)
    assert (bits + 63) // 64 * 8 == ref_dsize
    assert np.array_equal(out.view(np.uint8), ref_out.view(np.uint8))


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("tname", list(TYPES))
@pytest.mark.parametrize("shape", list(SHAPES))
def test_modes(shape, tname, mode):
    n = SHAPES[shape]
    t = TYPES[tname]
    a = smooth_field(n, t, seed=zlib.crc32(f"{shape}{tname}".encode()) & 0xffff, special=0.03)
    compare(a, n, t, MODES[mode](t, len(n)))


@pytest.mark.parametrize("layout", ["reversed", "interleaved", "permuted", "permuted_neg"])
@pytest.mark.parametrize("tname", ["f32", "f64", "i32"])
def test_layouts(layout, tname):
    n = [9, 6, 5]
    t = TYPES[tname]
    a = smooth_field(n, t, seed=4)
    for mode in ("rate7", "prec12", "reversible"):
        compare(a, n, t, MODES[mode](t, 3), layout=layout)


@pytest.mark.parametrize("header", [1, 2, 4, 7])
def test_headers(header):
    n = [10, 7, 6]
    a = smooth_field(n, T_FLOAT, seed=8)
    for mode in ("rate7", "prec12", "acc1e-2", "expert", "reversible"):
        compare(a, n, T_FLOAT, MODES[mode](T_FLOAT, 3), header=header)


def _special(ztype):
    if ztype == T_DOUBLE:
        tiny, sub, big, mx = 1e-300, 5e-324, 1e300, np.finfo(np.float64).max
    else:
        tiny, sub, big, mx = 1e-30, 1e-45, 1e38, float(np.finfo(np.float32).max)
    r = np.random.default_rng(17)
    blocks = [np.zeros(64), np.full(64, -0.0), r.standard_normal(64) * tiny, r.standard_normal(64) * tiny * 1e10,
              np.where(r.random(64) < 0.5, sub, -sub), r.standard_normal(64) * big,
              np.concatenate([[mx, -mx], r.standard_normal(62)]),
              np.concatenate([[np.inf], r.standard_normal(63)]),
              np.concatenate([[-np.inf, 1.5, -1.75, 3.0], r.standard_normal(60)]),
              np.concatenate([[np.nan], r.standard_normal(63)]),
              np.concatenate([[np.nan], np.zeros(63)]),
              np.concatenate([[1.0, -1.0], np.full(62, tiny)])]
    return np.concatenate(blocks).astype(np.float64 if ztype == T_DOUBLE else np.float32)


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("tname", ["f32", "f64"])
def test_special_values(tname, mode):
    t = TYPES[tname]
    a = _special(t)
    compare(a, [4, 4, a.size // 16], t, MODES[mode](t, 3))


@pytest.mark.parametrize("threads", [1, 3, 8])
def test_openmp_streams_equal_serial(threads):
    """The reference's OpenMP policy produces the serial stream; so does the
    oracle (the CPU baseline times both)."""
    api = ref_api()
    n = [24, 20, 12]
    a = smooth_field(n, T_DOUBLE, seed=2)
    for params in (stream_params_rate(16, T_DOUBLE, 3), stream_params_precision(20), stream_params_accuracy(1e-4)):
        w_omp, s_omp = api_compress(api, a, 0, n, None, T_DOUBLE, params, policy=1, threads=threads)
        words, size = oracle_stream(a, 0, n, None, T_DOUBLE, params)
        assert s_omp == size and np.array_equal(w_omp, words)
