"""GPU parity: zfp_compress / zfp_decompress with zfp_exec_hip against the
oracle (the CPU restatement of serial zfp), bit for bit, on seeded inputs.

Mirrors the reference's end-to-end tests (tests/src/endtoend/
zfpEndtoendBase.c:386-464): the compressed stream must equal the serial
stream word for word, decompression must consume exactly the bytes
compression produced, and the decompressed array must equal the serial
decompression bit for bit.  Covers 1D-4D, the four scalar types, fixed-rate,
fixed-precision, fixed-accuracy and expert modes, partial blocks, the
reference's four memory layouts plus a negative-stride permuted view, streams
that start after a header, and host and device memory.
"""
import zlib

import numpy as np
import pytest

import zfp_amd as Z
from hiputil import Run
from zfptest import (T_DOUBLE, T_FLOAT, T_INT32, T_INT64, NP_OF_TYPE, make_layout, oracle_decompress,
                     oracle_stream, read_layout, smooth_field, stream_params_accuracy, stream_params_precision,
                     stream_params_rate)

pytestmark = pytest.mark.gpu

TYPES = {"i32": T_INT32, "i64": T_INT64, "f32": T_FLOAT, "f64": T_DOUBLE}
SHAPES = {"1d": [103], "2d": [13, 11], "3d": [9, 10, 7], "4d": [5, 6, 7, 5], "3d_full": [16, 16, 8]}


def mode_params(name, ztype, dims):
    if name == "rate8":
        return stream_params_rate(8, ztype, dims)
    if name == "rate21":
        return stream_params_rate(21.5, ztype, dims)
    if name == "rate5_aligned":
        return stream_params_rate(5, ztype, dims, align=True)
    if name == "prec16":
        return stream_params_precision(16)
    if name == "prec28":
        return stream_params_precision(28)
    if name == "acc1e-3":
        return stream_params_accuracy(1e-3)
    if name == "acc1e-7":
        return stream_params_accuracy(1e-7)
    if name == "expert":
        return (100, 600, 20, -30)
    if name == "reversible":   # zfp_stream_set_reversible (src/zfp.c:750-757)
        return (1, 16658, 64, -1075)
    if name == "rev_expert":   # reversible with a budget, minbits padding and a precision cap
        return (300, 2500, 40, -1075)
    raise ValueError(name)


MODES = ["rate8", "rate21", "rate5_aligned", "prec16", "prec28", "acc1e-3", "acc1e-7", "expert",
         "reversible", "rev_expert"]


def check(a, n, ztype, params, layout="as_is", header=0, device=True, device_stream=None):
    base, off, s = make_layout(a, n, layout)
    run = Run(base, off, n, s, ztype, params, header=header, device=device, device_stream=device_stream)
    try:
        size = run.compress()
        exp_words, exp_size = oracle_stream(base, off, n, s, ztype, params, header)
        assert size == exp_size, f"stream size {size} != serial {exp_size} ({run.reason})"
        got = run.stream_words()
        if not np.array_equal(got, exp_words):
            bad = np.nonzero(got != exp_words)[0]
            raise AssertionError(f"stream differs from serial at {bad.size} words, first word {bad[0]}")
        # serial decompression of the serial stream
        ref = np.zeros_like(base)
        hbits = run.header_bits
        obits = oracle_decompress(exp_words, ref, off, n, s, ztype, params, offset=hbits)
        consumed, out = run.decompress(np.zeros_like(base))
        # zfp_decompress returns the aligned stream size after the last block
        # read: the compressed size, except for reversible streams with
        # minbits > 1, whose all-zero blocks are read as minbits bits
        expect = (hbits + obits + 63) // 64 * 8
        assert consumed == expect, f"decompress consumed {consumed} bytes, serial {expect} (compress wrote {size})"
        got_vals = read_layout(out, off, n, s)
        exp_vals = read_layout(ref, off, n, s)
        if not np.array_equal(got_vals.view(np.uint8), exp_vals.view(np.uint8)):
            bad = np.nonzero(got_vals.view(f"u{got_vals.itemsize}") != exp_vals.view(f"u{exp_vals.itemsize}"))[0]
            raise AssertionError(f"decompressed values differ at {bad.size} elements, first {bad[0]}")
        if layout == "interleaved":
            # the gaps between interleaved elements are left alone
            assert np.array_equal(out[1::2], np.zeros_like(out[1::2]))
    finally:
        run.close()


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("tname", list(TYPES))
@pytest.mark.parametrize("shape", list(SHAPES))
def test_modes_types_shapes(shape, tname, mode):
    n = SHAPES[shape]
    ztype = TYPES[tname]
    a = smooth_field(n, ztype, seed=zlib.crc32(f"{shape}/{tname}".encode()) & 0xffff)
    check(a, n, ztype, mode_params(mode, ztype, len(n)))


@pytest.mark.parametrize("rate", [17, 24, 32, 48])
@pytest.mark.parametrize("tname", list(TYPES))
@pytest.mark.parametrize("layout", ["as_is", "permuted_neg"])
def test_long_fixed_rate_blocks(layout, tname, rate):
    """3D fixed-rate blocks of more than 16 words, word-aligned, every type:
    the slot coders with their longest slots (the wave's blocks copied as
    one run of words).  300 blocks, so five waves, the last one partial, and
    partial blocks on every axis."""
    n = [37, 22, 19]
    ztype = TYPES[tname]
    a = smooth_field(n, ztype, seed=zlib.crc32(f"long/{tname}/{rate}".encode()) & 0xffff,
                     special=0.02 if ztype in (T_FLOAT, T_DOUBLE) else 0.0)
    p = stream_params_rate(rate, ztype, 3)
    assert p[1] % 64 == 0 and p[1] > 16 * 64
    check(a, n, ztype, p, layout=layout)


@pytest.mark.parametrize("layout", ["reversed", "interleaved", "permuted", "permuted_neg"])
@pytest.mark.parametrize("mode", ["rate8", "prec16", "acc1e-3"])
@pytest.mark.parametrize("tname", ["f32", "f64", "i64"])
def test_layouts(layout, mode, tname):
    n = [10, 9, 6]
    ztype = TYPES[tname]
    a = smooth_field(n, ztype, seed=11)
    check(a, n, ztype, mode_params(mode, ztype, 3), layout=layout)


@pytest.mark.parametrize("mode", ["rate8", "rate21", "prec28", "acc1e-7"])
@pytest.mark.parametrize("header", [7, 1, 2])
def test_header_prefixed_stream(mode, header):
    """zfpy writes a full header first (python/zfpy.pyx:172-176), so the codec
    starts at a bit offset inside a word that holds buffered header bits."""
    n = [12, 7, 5]
    a = smooth_field(n, T_DOUBLE, seed=5)
    check(a, n, T_DOUBLE, mode_params(mode, T_DOUBLE, 3), header=header)


@pytest.mark.parametrize("mode", ["rate8", "prec16"])
@pytest.mark.parametrize("where", ["host_both", "host_field", "host_stream"])
def test_host_memory(mode, where):
    n = [17, 9, 6]
    a = smooth_field(n, T_FLOAT, seed=3)
    device = where == "host_stream"
    dstream = where == "host_field"
    check(a, n, T_FLOAT, mode_params(mode, T_FLOAT, 3), device=device, device_stream=dstream)
    check(a, n, T_FLOAT, mode_params(mode, T_FLOAT, 3), layout="interleaved", device=device,
          device_stream=dstream)


def special_blocks(ztype):
    """Blocks that exercise the corner semantics of the block-floating-point
    cast: zeros, subnormals, huge values, blocks whose maximum is so small
    that the scale factor overflows, infinities and NaNs."""
    if ztype == T_DOUBLE:
        tiny, sub, big = 1e-300, 5e-324, 1e300
        maxv, inf = np.finfo(np.float64).max, np.inf
    else:
        tiny, sub, big = np.float32(1e-30), np.float32(1e-45), np.float32(1e38)
        maxv, inf = np.finfo(np.float32).max, np.float32(np.inf)
    dt = NP_OF_TYPE[ztype]
    rng = np.random.default_rng(99)
    blocks = [
        np.zeros(64),
        np.full(64, -0.0),
        rng.standard_normal(64) * tiny,          # max below the scale-overflow threshold
        rng.standard_normal(64) * tiny * 1e10,
        np.where(rng.random(64) < 0.5, sub, -sub),
        rng.standard_normal(64) * big,
        np.concatenate([[maxv, -maxv], rng.standard_normal(62)]),
        np.concatenate([[inf], rng.standard_normal(63)]),
        np.concatenate([[-inf, 1.5, -1.75, 3.0], rng.standard_normal(60)]),
        np.concatenate([[np.nan], rng.standard_normal(63)]),
        np.concatenate([[np.nan], np.zeros(63)]),
        np.concatenate([[1.0, -1.0], np.full(62, tiny)]),
        rng.standard_normal(64),
    ]
    return np.concatenate(blocks).astype(dt)


@pytest.mark.parametrize("mode", ["rate8", "rate21", "prec28", "acc1e-7", "expert"])
@pytest.mark.parametrize("tname", ["f32", "f64"])
def test_special_values(tname, mode):
    ztype = TYPES[tname]
    a = special_blocks(ztype)
    n = [4, 4, 4 * (a.size // 64)]  # one block per z-slab
    check(a, n, ztype, mode_params(mode, ztype, 3))


@pytest.mark.parametrize("tname", ["i32", "i64"])
def test_integer_extremes(tname):
    ztype = TYPES[tname]
    bits = 32 if ztype == T_INT32 else 64
    lim = (1 << (bits - 2)) - 1  # inputs the transform supports (|x| < 2^(P-2))
    rng = np.random.default_rng(4)
    a = rng.integers(-lim, lim, size=4 * 4 * 4 * 6, dtype=NP_OF_TYPE[ztype])
    a[:64] = lim
    a[64:128] = -lim
    for mode in ["rate8", "prec16", "expert"]:
        check(a, [4, 4, 24], ztype, mode_params(mode, ztype, 3))


def test_variable_rate_without_index():
    """A variable-rate stream with no recorded block index is decoded after a
    sequential parse of the block lengths on the device."""
    api = Z.api_lib()
    n = [13, 9, 8]
    a = smooth_field(n, T_DOUBLE, seed=21)
    params = stream_params_precision(20)
    base, off, s = make_layout(a, n, "as_is")
    run = Run(base, off, n, s, T_DOUBLE, params)
    try:
        size = run.compress()
        Z.hip_lib().zfp_hip_index_clear()
        consumed, out = run.decompress(np.zeros_like(base))
        assert consumed == size
        exp_words, _ = oracle_stream(base, off, n, s, T_DOUBLE, params)
        ref = np.zeros_like(base)
        oracle_decompress(exp_words, ref, off, n, s, T_DOUBLE, params)
        assert np.array_equal(out.view(np.uint64), ref.view(np.uint64))
    finally:
        run.close()
    assert api is not None


@pytest.mark.parametrize("tname", list(TYPES))
@pytest.mark.parametrize("kind", ["smooth", "noise", "specials", "zeros"])
def test_reversible_is_lossless_and_serial(tname, kind):
    """Reversible mode (src/template/rev*.c): the stream equals serial zfp's
    and decompression returns every input bit (losslessness is what the mode
    promises; the blocks cover both the block-floating-point and the
    type-punned integer paths and all-zero blocks)."""
    n = [10, 9, 7]
    ztype = TYPES[tname]
    a = smooth_field(n, ztype, seed=23, special=0.05 if kind == "specials" else 0.0)
    if kind == "noise":
        rng = np.random.default_rng(3)
        a = rng.integers(0, 256, size=a.nbytes, dtype=np.uint8).view(a.dtype).copy()
        if a.dtype.kind == "f":
            a[~np.isfinite(a)] = 0   # NaN payloads are kept too, but keep the input plain
    if kind == "zeros":
        a[: a.size // 2] = 0
        if a.dtype.kind == "f":
            a[a.size // 2: a.size // 2 + 9] = -0.0
    params = mode_params("reversible", ztype, 3)
    check(a, n, ztype, params)
    base, off, s = make_layout(a, n, "as_is")
    run = Run(base, off, n, s, ztype, params)
    try:
        run.compress()
        _, out = run.decompress(np.zeros_like(base))
        assert np.array_equal(out.view(np.uint8), base.view(np.uint8)), "reversible round trip lost bits"
    finally:
        run.close()


def test_config0_fixed_rate_3d_float():
    """The 3D float32 128^3, rate-8 workload: a 2 MiB stream, identical to
    serial zfp's."""
    n = [128, 128, 128]
    a = smooth_field(n, T_FLOAT, seed=2024)
    check(a, n, T_FLOAT, stream_params_rate(8, T_FLOAT, 3))


def _split_field(n, seed):
    """Smooth doubles with a constant block column and zero planes: at fixed
    rates up to 16 bits per value the constant blocks code planes below 32
    and the smooth ones stop above, so the 3D double decoder's second pass
    (blocks that go below plane 32) has work."""
    a = smooth_field(n, T_DOUBLE, seed=seed)
    v = a.reshape(n[::-1])
    v[..., :4] = v.flat[0]
    v[..., -2:] = 0
    return a


@pytest.mark.parametrize("rate", [4, 8, 16])
@pytest.mark.parametrize("layout", ["as_is", "permuted_neg"])
@pytest.mark.parametrize("shape", ["3d", "3d_full"])
def test_split_decode(shape, layout, rate):
    n = SHAPES[shape]
    a = _split_field(n, seed=zlib.crc32(f"split/{shape}".encode()) & 0xffff)
    check(a, n, T_DOUBLE, stream_params_rate(rate, T_DOUBLE, 3, align=True), layout=layout)


def test_split_decode_every_block_redone():
    """A constant 256^3 double field at rate 16: every block codes planes
    below 32, so all 262144 blocks take the second pass, two per lane of its
    grid."""
    n = [256, 256, 256]
    a = np.full(n[0] * n[1] * n[2], 1.0 / 3.0)
    check(a, n, T_DOUBLE, stream_params_rate(16, T_DOUBLE, 3, align=True))
