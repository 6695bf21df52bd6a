"""Pins of the oracle (oracle/zfp_oracle.c), the CPU restatement of serial
zfp that every GPU parity test trusts.  Each check is anchored outside the
restatement itself:

* the reference's own end-to-end golden checksums
  (tests/constants/checksums/*.h, as tests/golden/ref_checksums.json) over its
  own smooth-data generator (tests/utils/genSmoothRandNums.c, built into
  oracle/_ref/), for 1D-4D x int32/int64/float/double x fixed-rate,
  fixed-precision and fixed-accuracy x three parameters;
* testzfp's known compressed sizes (tests/testzfp.cpp:849-955) over its
  integer-refinement generator, restated below and verified by testzfp's own
  data checksums; these also cover reversible mode;
* closed-form invariants: negabinary is base -2, the budgeted coder emits a
  prefix of the unbudgeted code, the coder pair is lossless with all planes,
  fixed-rate streams are blocks * maxbits long.
"""
import json
import os

import numpy as np
import pytest

from zfptest import (COMPRESSED_BITSTREAM, DECOMPRESSED_ARRAY, MODE_FIXED_ACCURACY, MODE_FIXED_PRECISION,
                     MODE_FIXED_RATE, ORIGINAL_INPUT, ROOT, T_DOUBLE, T_FLOAT, T_INT32, T_INT64, TYPE_NAMES,
                     ZFP_MIN_EXP, blocks_of, hash64, hash_array, have_ref, oracle_compress, oracle_decompress,
                     oracle_lib, ref_checksum, ref_mode_params, ref_smooth_input, stream_params_accuracy,
                     stream_params_precision, stream_params_rate, u64p)

needs_ref = pytest.mark.skipif(not have_ref(), reason="oracle/_ref (reference build) not present")


@needs_ref
@pytest.mark.parametrize("ztype", [T_INT32, T_INT64, T_FLOAT, T_DOUBLE], ids=lambda t: TYPE_NAMES[t])
@pytest.mark.parametrize("dims", [1, 2, 3, 4])
def test_reference_golden_checksums(dims, ztype):
    """Serial-zfp end-to-end checksums of the reference
    (tests/src/endtoend/zfpEndtoendBase.c:386-464)."""
    a, side = ref_smooth_input(ztype, dims)
    n = [side] * dims
    assert hash_array(a) == ref_checksum(dims, ztype, ORIGINAL_INPUT, n, 0, 0), "generator/hash mismatch"
    modes = [MODE_FIXED_RATE, MODE_FIXED_PRECISION]
    if ztype in (T_FLOAT, T_DOUBLE):
        modes.append(MODE_FIXED_ACCURACY)
    for mode in modes:
        for param in range(3):
            p = ref_mode_params(mode, param, ztype, dims)
            words, bits = oracle_compress(a, 0, n, None, ztype, p)
            nw = (bits + 63) // 64
            assert hash64(words[:nw]) == ref_checksum(dims, ztype, COMPRESSED_BITSTREAM, n, mode, param), \
                f"stream checksum, mode {mode} param {param}"
            out = np.zeros_like(a)
            assert oracle_decompress(words, out, 0, n, None, ztype, p) == bits
            assert hash_array(out) == ref_checksum(dims, ztype, DECOMPRESSED_ARRAY, n, mode, param), \
                f"decompressed checksum, mode {mode} param {param}"


# ---------------------------------------------------------------------------
# testzfp (tests/testzfp.cpp) known answers
# ---------------------------------------------------------------------------

def _trunc_div(s, d):
    return np.sign(s) * (np.abs(s) // d)


def _refine(f):
    """One refinement step of testzfp (testzfp.cpp:44-134): every axis doubles;
    odd positions get the (-1, 9, 9, -1)/16 cubic stencil (cyclic), even
    positions copy; the integer sum is divided once, truncating."""
    w = np.array([-1, 9, 9, -1], dtype=np.int64)
    s = f.astype(np.int64)
    for axis in range(f.ndim):
        m = s.shape[axis]
        even = 16 * s
        odd = sum(w[i] * np.roll(s, -(i - 1), axis=axis) for i in range(4))
        shape = list(s.shape)
        shape[axis] = 2 * m
        t = np.empty(shape, dtype=np.int64)
        sl_e = [slice(None)] * f.ndim
        sl_o = [slice(None)] * f.ndim
        sl_e[axis] = slice(0, None, 2)
        sl_o[axis] = slice(1, None, 2)
        t[tuple(sl_e)] = even
        t[tuple(sl_o)] = odd
        s = t
    return _trunc_div(s, 16 ** f.ndim).astype(np.int64)


def refinement_array(dims, dtype):
    """testzfp's small test array, 4096 values (testzfp.cpp:136-291)."""
    n = 1 << (12 // dims)
    f = np.zeros([4] * dims, dtype=np.int64)
    for idx in np.ndindex(*([2] * dims)):
        c = tuple(i + 1 for i in idx)  # coordinates 1..2 on every axis
        parity = 0
        for v in c:
            parity ^= v
        f[c] = 0x10000 * (1 - 2 * (parity & 1))
    m = 4
    while m < n:
        f = _refine(f)
        m *= 2
    # arrays are indexed [w][z][y][x] here and stored x fastest, as in C
    return np.ldexp(f.reshape(-1).astype(dtype), -12)


def _oaat_bytes(b):
    h = 0
    for v in b:
        h = (h + v) & 0xffffffff
        h = (h + (h << 10)) & 0xffffffff
        h ^= h >> 6
    h = (h + (h << 3)) & 0xffffffff
    h ^= h >> 11
    h = (h + (h << 15)) & 0xffffffff
    return h


def _known():
    with open(os.path.join(ROOT, "tests", "golden", "testzfp_known_sizes.json")) as fh:
        return json.load(fh)


@pytest.mark.parametrize("tname", ["float", "double"])
@pytest.mark.parametrize("dims", [1, 2, 3, 4])
def test_testzfp_known_sizes(dims, tname):
    known = _known()
    dtype = np.float32 if tname == "float" else np.float64
    ztype = T_FLOAT if tname == "float" else T_DOUBLE
    a = refinement_array(dims, dtype)
    assert _oaat_bytes(a.tobytes()) == known["data_checksum"][tname][dims - 1], "testzfp data differ"
    side = 1 << (12 // dims)
    n = [side] * dims

    def size_of(params):
        _, bits = oracle_compress(a, 0, n, None, ztype, params)
        return (bits + 63) // 64 * 8

    for p, exp in zip(known["precision"][tname]["params"], known["precision"][tname]["bytes"][dims - 1]):
        assert size_of(stream_params_precision(p)) == exp, f"precision {p}"
    eps = float(np.finfo(dtype).eps)
    for name, exp in zip(known["accuracy"][tname]["params"], known["accuracy"][tname]["bytes"][dims - 1]):
        tol = {"1e-3": float(dtype(1e-3)), "2eps": 2 * eps, "0": 0.0}[name]
        assert size_of(stream_params_accuracy(tol)) == exp, f"accuracy {name}"
    rev = (1, 16658, 64, ZFP_MIN_EXP - 1)
    assert size_of(rev) == known["reversible"][tname][dims - 1]
    # reversible mode reconstructs exactly
    words, bits = oracle_compress(a, 0, n, None, ztype, rev)
    out = np.zeros_like(a)
    oracle_decompress(words, out, 0, n, None, ztype, rev)
    assert np.array_equal(out.view(np.uint8), a.view(np.uint8))


# ---------------------------------------------------------------------------
# closed-form invariants
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("prec", [32, 64])
def test_negabinary_is_base_minus_two(prec):
    """to_negabinary(x) has base -2 digits of x (mod 2^prec), and inverts."""
    lib = oracle_lib()
    rng = np.random.default_rng(1)
    lim = 1 << (prec - 2)
    vals = [0, 1, -1, 2, -2, lim - 1, -lim + 1] + [int(v) for v in rng.integers(-lim, lim, 200)]
    for x in vals:
        u = lib.oracle_to_negabinary(x, prec)
        digits = sum((((u >> i) & 1) * (-2) ** i) for i in range(prec))
        assert (digits - x) % (1 << prec) == 0
        assert lib.oracle_from_negabinary(u, prec) == x


@pytest.mark.parametrize("size,prec", [(4, 32), (16, 64), (64, 32), (64, 64), (256, 64)])
def test_plane_coder_lossless_with_all_planes(size, prec):
    lib = oracle_lib()
    rng = np.random.default_rng(size + prec)
    for trial in range(20):
        bits_used = int(rng.integers(1, prec + 1))
        u = rng.integers(0, 1 << min(bits_used, 63), size=size, dtype=np.uint64)
        if trial % 5 == 0:
            u[:] = 0
        words = np.zeros(size * (prec + 2) // 64 + 4, dtype=np.uint64)
        nb = lib.oracle_encode_planes(u.ctypes.data_as(u64p), size, prec, 1 << 30, prec,
                                      words.ctypes.data_as(u64p), 0)
        back = np.zeros(size, dtype=np.uint64)
        nr = lib.oracle_decode_planes(back.ctypes.data_as(u64p), size, prec, 1 << 30, prec,
                                      words.ctypes.data_as(u64p), 0)
        assert nb == nr
        assert np.array_equal(back, u)


@pytest.mark.parametrize("ztype", [T_FLOAT, T_DOUBLE, T_INT64])
@pytest.mark.parametrize("dims", [1, 2, 3, 4])
def test_fixed_rate_stream_is_prefix_of_full_code(dims, ztype):
    """A block coded with budget B equals the first B bits of the unbudgeted
    code (encode.c:90-132: every emitted bit decrements the budget), padded
    with zeros to minbits = B."""
    lib = oracle_lib()
    rng = np.random.default_rng(dims * 10 + ztype)
    n = 1 << (2 * dims)
    dt = {T_FLOAT: np.float32, T_DOUBLE: np.float64, T_INT64: np.int64}[ztype]
    for trial in range(10):
        if ztype == T_INT64:
            blk = rng.integers(-(1 << 40), 1 << 40, size=n).astype(dt)
        else:
            blk = (rng.standard_normal(n) * 10.0 ** rng.integers(-5, 5)).astype(dt)
        full = np.zeros(4 * n + 8, dtype=np.uint64)
        fb = lib.oracle_encode_block(ztype, dims, blk.ctypes.data, 0, 1 << 20, 64, ZFP_MIN_EXP,
                                     full.ctypes.data_as(u64p), 0)
        for budget in sorted(set(int(b) for b in rng.integers(13, fb + 20, size=6))):
            part = np.zeros_like(full)
            pb = lib.oracle_encode_block(ztype, dims, blk.ctypes.data, budget, budget, 64, ZFP_MIN_EXP,
                                         part.ctypes.data_as(u64p), 0)
            assert pb == budget
            keep = min(budget, fb)
            bits_full = np.unpackbits(full.view(np.uint8), bitorder="little")
            bits_part = np.unpackbits(part.view(np.uint8), bitorder="little")
            assert np.array_equal(bits_part[:keep], bits_full[:keep])
            assert not bits_part[keep:].any()


@pytest.mark.parametrize("rate", [4, 8, 13.5])
def test_fixed_rate_size_law(rate):
    """Fixed-rate streams hold exactly blocks * maxbits bits
    (tests/src/endtoend/serialExecBase.c:92-138)."""
    n = [13, 10, 6]
    a = np.random.default_rng(3).standard_normal(13 * 10 * 6)
    p = stream_params_rate(rate, T_DOUBLE, 3)
    _, bits = oracle_compress(a, 0, n, None, T_DOUBLE, p)
    assert bits == blocks_of(n) * p[1]


PAR_CASES = [
    ([37, 29, 23], T_DOUBLE, stream_params_rate(16, T_DOUBLE, 3), None),
    ([203, 97], T_FLOAT, stream_params_precision(16), None),
    ([31, 45], T_FLOAT, stream_params_accuracy(1e-3), [45, 1]),          # transposed strides
    ([9, 10, 11, 13], T_DOUBLE, stream_params_accuracy(1e-7), None),
    ([301], T_INT64, (1, 16658, 64, ZFP_MIN_EXP - 1), None),             # reversible
    ([17, 18, 19], T_INT32, (40, 700, 30, ZFP_MIN_EXP), None),           # expert, minbits padding
]


@pytest.mark.parametrize("k", range(len(PAR_CASES)))
def test_parallel_oracle_equals_single_call(k):
    """The slab-parallel oracle used on the full-size workloads writes the
    single-call stream word for word and decodes it to the same values: the
    blocks of a run of whole block rows are a contiguous part of the stream,
    so the slab streams concatenated bit for bit are the whole stream."""
    from zfptest import oracle_compress_parallel, oracle_decompress_parallel, smooth_field
    n, ztype, p, s = PAR_CASES[k]
    a = smooth_field(n, ztype, seed=900 + k, special=0.01 if ztype == T_DOUBLE else 0.0)
    if s:   # the same logical field in a transposed layout
        a = a.reshape(n[::-1]).T.copy().reshape(-1)
    words, bits = oracle_compress(a, 0, n, s, ztype, p)
    for threads in (1, 3, 8):
        pw, pbits, starts, slabs = oracle_compress_parallel(a, 0, n, s, ztype, p, threads=threads)
        assert pbits == bits and len(slabs) > (1 if threads > 1 else 0)
        nw = (bits + 63) // 64
        assert np.array_equal(pw[:nw], words[:nw])
        ref = np.zeros_like(a)
        assert oracle_decompress(words, ref, 0, n, s, ztype, p) == bits
        out = np.zeros_like(a)
        assert oracle_decompress_parallel(pw, out, 0, n, s, ztype, p, starts, slabs, threads=threads) == bits
        assert np.array_equal(out.view(np.uint8), ref.view(np.uint8))
