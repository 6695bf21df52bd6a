"""Shared helpers for the tests: loaders for the oracle (our CPU restatement,
oracle/liboracle.so), the reference builds in oracle/_ref/ (when present) and
small utilities (layouts, checksum keys, seeded fields).

Only tests, __graft_entry__.smoke() and bench.py's CPU-baseline leg use the
oracle; the product under zfp_amd/ never does.
"""
import ctypes
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
REF_DIR = os.path.join(ORACLE_DIR, "_ref")

T_INT32, T_INT64, T_FLOAT, T_DOUBLE = 1, 2, 3, 4
NP_OF_TYPE = {T_INT32: np.int32, T_INT64: np.int64, T_FLOAT: np.float32, T_DOUBLE: np.float64}
TYPE_OF_NP = {np.dtype(v): k for k, v in NP_OF_TYPE.items()}
TYPE_NAMES = {T_INT32: "Int32", T_INT64: "Int64", T_FLOAT: "Float", T_DOUBLE: "Double"}

ZFP_MIN_BITS, ZFP_MAX_BITS, ZFP_MAX_PREC, ZFP_MIN_EXP = 1, 16658, 64, -1074

u64p = ctypes.POINTER(ctypes.c_uint64)
szp = ctypes.POINTER(ctypes.c_size_t)
pdp = ctypes.POINTER(ctypes.c_ssize_t)
u32p = ctypes.POINTER(ctypes.c_uint32)

_oracle = None


def oracle_lib():
    global _oracle
    if _oracle is None:
        path = os.path.join(ORACLE_DIR, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built; run __graft_entry__.build()")
        lib = ctypes.CDLL(path)
        lib.oracle_compress.restype = ctypes.c_uint64
        lib.oracle_compress.argtypes = [ctypes.c_int, szp, pdp, ctypes.c_void_p,
                                        ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_int,
                                        u64p, ctypes.c_uint64, u32p]
        lib.oracle_decompress.restype = ctypes.c_uint64
        lib.oracle_decompress.argtypes = [ctypes.c_int, szp, pdp, ctypes.c_void_p,
                                          ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_int,
                                          u64p, ctypes.c_uint64, u32p]
        lib.oracle_encode_block.restype = ctypes.c_uint
        lib.oracle_encode_block.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_int,
                                            u64p, ctypes.c_uint64]
        lib.oracle_decode_block.restype = ctypes.c_uint
        lib.oracle_decode_block.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
                                            ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_int,
                                            u64p, ctypes.c_uint64]
        lib.oracle_xform.restype = None
        lib.oracle_xform.argtypes = [ctypes.POINTER(ctypes.c_int64), ctypes.c_int, ctypes.c_uint, ctypes.c_int]
        lib.oracle_to_negabinary.restype = ctypes.c_uint64
        lib.oracle_to_negabinary.argtypes = [ctypes.c_int64, ctypes.c_uint]
        lib.oracle_from_negabinary.restype = ctypes.c_int64
        lib.oracle_from_negabinary.argtypes = [ctypes.c_uint64, ctypes.c_uint]
        lib.oracle_encode_planes.restype = ctypes.c_uint
        lib.oracle_encode_planes.argtypes = [u64p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                             ctypes.c_uint, u64p, ctypes.c_uint64]
        lib.oracle_decode_planes.restype = ctypes.c_uint
        lib.oracle_decode_planes.argtypes = [u64p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                             ctypes.c_uint, u64p, ctypes.c_uint64]
        lib.oracle_order_table.restype = ctypes.POINTER(ctypes.c_ubyte)
        lib.oracle_order_table.argtypes = [ctypes.c_int]
        lib.oracle_hash_words64.restype = ctypes.c_uint64
        lib.oracle_hash_words64.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_ssize_t]
        lib.oracle_hash_words32.restype = ctypes.c_uint32
        lib.oracle_hash_words32.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_ssize_t]
        _oracle = lib
    return _oracle


def have_ref():
    return os.path.exists(os.path.join(REF_DIR, "libzfp_ref.so")) and \
        os.path.exists(os.path.join(REF_DIR, "libzfp_reftests.so"))


_reftests = None


def reftests_lib():
    """The reference's own test-data generator and hashes (tests/utils/*.c)."""
    global _reftests
    if _reftests is None:
        lib = ctypes.CDLL(os.path.join(REF_DIR, "libzfp_reftests.so"))
        for name in ("generateSmoothRandFloats", "generateSmoothRandDoubles"):
            getattr(lib, name).restype = None
            getattr(lib, name).argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.POINTER(ctypes.c_void_p),
                                           szp, szp]
        for name in ("generateSmoothRandInts32", "generateSmoothRandInts64"):
            getattr(lib, name).restype = None
            getattr(lib, name).argtypes = [ctypes.c_size_t, ctypes.c_int, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_void_p), szp, szp]
        _reftests = lib
    return _reftests


_libc = ctypes.CDLL(None)
_libc.free.argtypes = [ctypes.c_void_p]


def ref_smooth_input(ztype, dims):
    """The reference end-to-end test input for (type, dims): its own
    generator with its own sizes (tests/src/endtoend/zfpEndtoendBase.c:19-23,
    60-101).  Returns (array in x-fastest order, side length)."""
    lib = reftests_lib()
    ptr = ctypes.c_void_p()
    side = ctypes.c_size_t()
    total = ctypes.c_size_t()
    fp = ztype in (T_FLOAT, T_DOUBLE)
    minel = 1000000 if fp else 4096
    if ztype == T_FLOAT:
        lib.generateSmoothRandFloats(minel, dims, ctypes.byref(ptr), ctypes.byref(side), ctypes.byref(total))
    elif ztype == T_DOUBLE:
        lib.generateSmoothRandDoubles(minel, dims, ctypes.byref(ptr), ctypes.byref(side), ctypes.byref(total))
    elif ztype == T_INT32:
        lib.generateSmoothRandInts32(minel, dims, 30, ctypes.byref(ptr), ctypes.byref(side), ctypes.byref(total))
    else:
        lib.generateSmoothRandInts64(minel, dims, 62, ctypes.byref(ptr), ctypes.byref(side), ctypes.byref(total))
    dt = np.dtype(NP_OF_TYPE[ztype])
    buf = (ctypes.c_char * (total.value * dt.itemsize)).from_address(ptr.value)
    arr = np.frombuffer(bytes(buf), dtype=dt).copy()
    _libc.free(ptr)
    return arr, side.value


# --------------------------------------------------------------------------
# reference checksum tables (fixture) and keys
# --------------------------------------------------------------------------

_tables = None


def ref_checksum_tables():
    global _tables
    if _tables is None:
        with open(os.path.join(ROOT, "tests", "golden", "ref_checksums.json")) as f:
            _tables = json.load(f)["tables"]
    return _tables


ARRAY_TEST = 2
ORIGINAL_INPUT, COMPRESSED_BITSTREAM, DECOMPRESSED_ARRAY = 0, 1, 2
MODE_FIXED_RATE, MODE_FIXED_PRECISION, MODE_FIXED_ACCURACY, MODE_REVERSIBLE = 2, 3, 4, 5


def checksum_key(subject, n, mode, param):
    """tests/utils/zfpChecksums.c:75-136 of the reference."""
    k1 = ((((ARRAY_TEST << 2) + subject) << 3) + mode << 4) + param
    dims = sum(1 for v in n if v)
    if dims == 1:
        k2 = n[0] - 1
    elif dims == 2:
        k2 = ((n[0] - 1) << 24) + (n[1] - 1)
    elif dims == 3:
        k2 = ((((n[0] - 1) << 16) + (n[1] - 1)) << 16) + (n[2] - 1)
    else:
        k2 = ((((((n[0] - 1) << 12) + (n[1] - 1)) << 12) + (n[2] - 1)) << 12) + (n[3] - 1)
    return k1, k2


def ref_checksum(dims, ztype, subject, n, mode, param):
    k1, k2 = checksum_key(subject, n, mode, param)
    for a, b, c in ref_checksum_tables()[f"{dims}d{TYPE_NAMES[ztype]}"]:
        if a == k1 and b == k2:
            return c
    return None


def ref_mode_params(mode, param, ztype, dims):
    """(minbits, maxbits, maxprec, minexp) for the reference test parameters
    (tests/utils/zfpCompressionParams.c and src/zfp.c:759-811)."""
    if mode == MODE_FIXED_PRECISION:
        return stream_params_precision(1 << (param + 3))
    if mode == MODE_FIXED_RATE:
        return stream_params_rate(float(1 << (param + 3)), ztype, dims)
    if mode == MODE_FIXED_ACCURACY:
        return stream_params_accuracy(2.0 ** -(1 << param))
    if mode == MODE_REVERSIBLE:
        return (ZFP_MIN_BITS, ZFP_MAX_BITS, ZFP_MAX_PREC, ZFP_MIN_EXP - 1)
    raise ValueError(mode)


def stream_params_rate(rate, ztype, dims, align=False):
    """zfp_stream_set_rate (src/zfp.c:759-784)."""
    import math
    n = 1 << (2 * dims)
    bits = int(math.floor(n * rate + 0.5))
    if ztype == T_FLOAT:
        bits = max(bits, 9)
    elif ztype == T_DOUBLE:
        bits = max(bits, 12)
    if align:
        bits = (bits + 63) & ~63
    return (bits, bits, ZFP_MAX_PREC, ZFP_MIN_EXP)


def stream_params_precision(prec):
    return (ZFP_MIN_BITS, ZFP_MAX_BITS, min(prec, ZFP_MAX_PREC) if prec else ZFP_MAX_PREC, ZFP_MIN_EXP)


def stream_params_accuracy(tol):
    import math
    emin = ZFP_MIN_EXP
    if tol > 0:
        _, e = math.frexp(tol)
        emin = e - 1
    return (ZFP_MIN_BITS, ZFP_MAX_BITS, ZFP_MAX_PREC, emin)


# --------------------------------------------------------------------------
# oracle calls
# --------------------------------------------------------------------------

def field_args(n, s):
    nl = list(n) + [0] * (4 - len(n))
    sl = list(s) if s else []
    sl = sl + [0] * (4 - len(sl))
    return (ctypes.c_size_t * 4)(*nl), (ctypes.c_ssize_t * 4)(*sl)


def blocks_of(n):
    b = 1
    for v in n:
        if v:
            b *= (v + 3) // 4
    return b


def max_stream_words(n, ztype, params, offset=0):
    """Upper bound on stream words (zfp_stream_maximum_size, src/zfp.c:711-742)."""
    minbits, maxbits, maxprec, minexp = params
    dims = sum(1 for v in n if v)
    vals = 1 << (2 * dims)
    p = 32 if ztype in (T_INT32, T_FLOAT) else 64
    rev = minexp < ZFP_MIN_EXP
    hdr = {T_INT32: 5 if rev else 0, T_INT64: 6 if rev else 0,
           T_FLOAT: 15 if rev else 9, T_DOUBLE: 19 if rev else 12}[ztype]
    mb = hdr + vals - 1 + vals * min(maxprec, p)
    mb = max(min(mb, maxbits), minbits)
    return (offset + 148 + blocks_of(n) * mb + 63) // 64 + 1


def oracle_compress(arr_base, elem_offset, n, s, ztype, params, offset=0, prefix=None, block_bits=False):
    """Compress through the oracle.  arr_base is the numpy buffer, elem_offset
    the element index of the field origin inside it (for negative strides).
    Returns (words, nbits[, per-block bits])."""
    lib = oracle_lib()
    nwords = max_stream_words(n, ztype, params, offset)
    words = np.zeros(nwords, dtype=np.uint64)
    if prefix is not None:
        words[:len(prefix)] = prefix
    nn, ss = field_args(n, s)
    ptr = arr_base.ctypes.data + elem_offset * arr_base.itemsize
    bb = np.zeros(blocks_of(n), dtype=np.uint32) if block_bits else None
    bits = lib.oracle_compress(ztype, nn, ss, ctypes.c_void_p(ptr), *params,
                               words.ctypes.data_as(u64p), offset,
                               bb.ctypes.data_as(u32p) if bb is not None else None)
    if block_bits:
        return words, bits, bb
    return words, bits


def oracle_decompress(words, arr_base, elem_offset, n, s, ztype, params, offset=0):
    lib = oracle_lib()
    nn, ss = field_args(n, s)
    ptr = arr_base.ctypes.data + elem_offset * arr_base.itemsize
    w = np.ascontiguousarray(words, dtype=np.uint64)
    return lib.oracle_decompress(ztype, nn, ss, ctypes.c_void_p(ptr), *params,
                                 w.ctypes.data_as(u64p), offset, None)


def bit_concat(parts):
    """Concatenate bit streams: parts is a list of (words, bits) whose words
    are zero past their bits.  Returns (words, start bit of each part)."""
    total = sum(int(b) for _, b in parts)
    out = np.zeros((total + 63) // 64 + 1, dtype=np.uint64)
    starts = []
    pos = 0
    for w, b in parts:
        b = int(b)
        starts.append(pos)
        nw = (b + 63) // 64
        if nw:
            w = np.ascontiguousarray(w[:nw], dtype=np.uint64)
            base, sh = pos // 64, pos % 64
            if sh == 0:
                out[base:base + nw] |= w
            else:
                out[base:base + nw] |= w << np.uint64(sh)
                out[base + 1:base + nw + 1] |= w >> np.uint64(64 - sh)
        pos += b
    return out, starts


def _slabs(n, count):
    """Ranges [a, b) of the slowest axis that hold whole rows of blocks: the
    blocks of one range are a contiguous run of the stream (x fastest)."""
    nb = (n[-1] + 3) // 4
    k = max(1, min(count, nb))
    edges = [4 * (nb * i // k) for i in range(k + 1)]
    edges[-1] = n[-1]
    return [(edges[i], edges[i + 1]) for i in range(k) if edges[i] < edges[i + 1]]


def _strides(n, s):
    default = [1]
    for d in range(1, len(n)):
        default.append(default[-1] * n[d - 1])
    s = list(s or [])
    return [s[d] if d < len(s) and s[d] else default[d] for d in range(len(n))]


def oracle_threads():
    return max(1, min(16, os.cpu_count() or 1))


def oracle_compress_parallel(arr_base, elem_offset, n, s, ztype, params, threads=None):
    """oracle_compress of a large field, in parallel threads over slabs of
    whole block rows along the slowest axis, each compressed on its own and
    the slab streams concatenated bit for bit: the stream order runs slab by
    slab, so the result is the single-call stream.  Returns (words, bits,
    start bit of each slab, slabs)."""
    from concurrent.futures import ThreadPoolExecutor
    threads = threads or oracle_threads()
    st = _strides(n, s)
    slabs = _slabs(list(n), 4 * threads)

    def one(r):
        a, b = r
        nn = list(n)
        nn[-1] = b - a
        return oracle_compress(arr_base, elem_offset + a * st[-1], nn, st, ztype, params)

    with ThreadPoolExecutor(threads) as ex:
        parts = list(ex.map(one, slabs))
    words, starts = bit_concat(parts)
    return words, sum(int(b) for _, b in parts), starts, slabs


def oracle_decompress_parallel(words, arr_base, elem_offset, n, s, ztype, params, starts, slabs, threads=None):
    """oracle_decompress of the stream oracle_compress_parallel wrote, slab by
    slab in parallel threads; returns the bits consumed."""
    from concurrent.futures import ThreadPoolExecutor
    threads = threads or oracle_threads()
    st = _strides(n, s)
    w = np.ascontiguousarray(words, dtype=np.uint64)

    def one(k):
        a, b = slabs[k]
        nn = list(n)
        nn[-1] = b - a
        return oracle_decompress(w, arr_base, elem_offset + a * st[-1], nn, st, ztype, params, offset=starts[k])

    with ThreadPoolExecutor(threads) as ex:
        return sum(ex.map(one, range(len(slabs))))


def hash64(arr):
    a = np.ascontiguousarray(arr).view(np.uint64)
    return oracle_lib().oracle_hash_words64(a.ctypes.data, a.size, 1)


def hash32(arr):
    a = np.ascontiguousarray(arr).view(np.uint32)
    return oracle_lib().oracle_hash_words32(a.ctypes.data, a.size, 1)


def hash_array(arr):
    return hash32(arr) if arr.dtype.itemsize == 4 else hash64(arr)


# --------------------------------------------------------------------------
# layouts: the same logical field stored four ways (the reference's
# AS_IS / REVERSED / INTERLEAVED / PERMUTED, tests/utils/stridedOperations.c)
# --------------------------------------------------------------------------

def make_layout(a, n, layout):
    """a: logical array in x-fastest order (flat).  Returns (base, elem_offset,
    strides) such that element (x,y,z,w) lives at base[elem_offset +
    x*s0 + y*s1 + z*s2 + w*s3] and equals a[x + nx*(y + ny*(z + nz*w))]."""
    dims = sum(1 for v in n if v)
    nn = list(n[:dims])
    logical = a.reshape(list(reversed(nn)))  # [w][z][y][x]
    default = [1]
    for d in range(1, dims):
        default.append(default[-1] * nn[d - 1])
    if layout == "as_is":
        return a.copy(), 0, [0] * dims
    if layout == "reversed":
        base = a[::-1].copy()
        return base, a.size - 1, [-v for v in default]
    if layout == "interleaved":
        base = np.zeros(2 * a.size, dtype=a.dtype)
        base[::2] = a
        return base, 0, [2 * v for v in default]
    if layout in ("permuted", "permuted_neg"):
        # memory order has x slowest (the reference's PERMUTED layout,
        # tests/utils/stridedOperations.c:117-133); "permuted_neg" also
        # runs x backwards, like the permuted view of the 512^3 workload
        rev = list(reversed(nn))
        strides = [1]
        for d in range(1, dims):
            strides.append(strides[-1] * rev[d - 1])
        strides = list(reversed(strides))  # x gets the largest stride
        base = np.empty(a.size, dtype=a.dtype)
        idx = np.indices(nn[::-1]).reshape(dims, -1)  # [w,z,y,x] coordinate rows
        coords = idx[::-1]  # x, y, z, w
        off = np.zeros(a.size, dtype=np.int64)
        flip = layout == "permuted_neg"
        for d in range(dims):
            c = (nn[0] - 1 - coords[d]) if (flip and d == 0) else coords[d]
            off += c * strides[d]
        base[off] = logical.reshape(-1)
        if not flip:
            return base, 0, strides
        # element x lives at (nx-1-x)*s0: origin at the far end, stride -s0
        s0 = strides[0]
        return base, (nn[0] - 1) * s0, [-s0] + strides[1:]
    raise ValueError(layout)


def read_layout(base, elem_offset, n, s):
    """Inverse of make_layout: gather the logical array back."""
    dims = sum(1 for v in n if v)
    nn = list(n[:dims])
    idx = np.indices(nn[::-1]).reshape(dims, -1)
    coords = idx[::-1]
    default = [1]
    for d in range(1, dims):
        default.append(default[-1] * nn[d - 1])
    strides = [sv if sv else dv for sv, dv in zip(s, default)]
    off = np.full(idx.shape[1], elem_offset, dtype=np.int64)
    for d in range(dims):
        off += coords[d] * strides[d]
    return base[off]


# --------------------------------------------------------------------------
# stream header bits (src/zfp.c:1182-1217) and seeded fields
# --------------------------------------------------------------------------

def field_metadata(ztype, n):
    """52-bit field metadata (src/zfp.c:287-330)."""
    dims = len(n)
    width = {1: 48, 2: 24, 3: 16, 4: 12}[dims]
    meta = 0
    for v in reversed(n):
        meta = (meta << width) + (v - 1)
    meta = (meta << 2) + (dims - 1)
    return (meta << 2) + (ztype - 1)


def compression_mode(params):
    """zfp_stream_compression_mode (src/zfp.c:567-608)."""
    minbits, maxbits, maxprec, minexp = params
    if minbits > maxbits or not (0 < maxprec <= 64):
        return 0
    if (minbits, maxbits, maxprec, minexp) == (ZFP_MIN_BITS, ZFP_MAX_BITS, ZFP_MAX_PREC, ZFP_MIN_EXP):
        return 1
    if minbits == maxbits and 1 <= maxbits <= ZFP_MAX_BITS and maxprec >= ZFP_MAX_PREC and minexp == ZFP_MIN_EXP:
        return MODE_FIXED_RATE
    if minbits <= ZFP_MIN_BITS and maxbits >= ZFP_MAX_BITS and maxprec >= 1 and minexp == ZFP_MIN_EXP:
        return MODE_FIXED_PRECISION
    if minbits <= ZFP_MIN_BITS and maxbits >= ZFP_MAX_BITS and maxprec >= ZFP_MAX_PREC and minexp >= ZFP_MIN_EXP:
        return MODE_FIXED_ACCURACY
    if minbits <= ZFP_MIN_BITS and maxbits >= ZFP_MAX_BITS and maxprec >= ZFP_MAX_PREC and minexp < ZFP_MIN_EXP:
        return MODE_REVERSIBLE
    return 1


def stream_mode_word(params):
    """12- or 64-bit mode word (src/zfp.c:635-690); returns (value, bits)."""
    minbits, maxbits, maxprec, minexp = params
    mode = compression_mode(params)
    if mode == MODE_FIXED_RATE and maxbits <= 2048:
        return maxbits - 1, 12
    if mode == MODE_FIXED_PRECISION and maxprec <= 128:
        return maxprec - 1 + 2048, 12
    if mode == MODE_FIXED_ACCURACY and minexp <= 843:
        return minexp - ZFP_MIN_EXP + 2048 + 128 + 1, 12
    if mode == MODE_REVERSIBLE:
        return 2048 + 128, 12
    cl = lambda v, hi: max(1, min(v, hi)) - 1
    mode = 0
    mode = (mode << 15) + max(0, min(minexp + 16495, 0x7fff))
    mode = (mode << 7) + cl(maxprec, 0x80)
    mode = (mode << 15) + cl(maxbits, 0x8000)
    mode = (mode << 15) + cl(minbits, 0x8000)
    mode = (mode << 12) + 0xfff
    return mode, 64


def header_prefix(ztype, n, params, mask=7):
    """Words and bit count of the header zfp_write_header(mask) writes."""
    fields = []
    if mask & 1:
        fields += [(ord("z"), 8), (ord("f"), 8), (ord("p"), 8), (5, 8)]
    if mask & 2:
        fields.append((field_metadata(ztype, n), 52))
    if mask & 4:
        fields.append(stream_mode_word(params))
    acc, nbits = 0, 0
    for v, b in fields:
        acc |= (v & ((1 << b) - 1)) << nbits
        nbits += b
    words = [(acc >> (64 * i)) & 0xffffffffffffffff for i in range((nbits + 63) // 64)]
    return np.array(words, dtype=np.uint64), nbits


def oracle_stream(arr_base, elem_offset, n, s, ztype, params, header=0):
    """The flushed stream zfp_compress produces after an optional header:
    (words, size in bytes)."""
    prefix, hbits = header_prefix(ztype, n, params, header) if header else (None, 0)
    words, bits = oracle_compress(arr_base, elem_offset, n, s, ztype, params, offset=hbits, prefix=prefix)
    nw = (hbits + bits + 63) // 64
    return words[:nw].copy(), nw * 8


def smooth_field(n, ztype, seed, noise=1e-3, special=0.0):
    """Seeded smooth field (sum of sinusoids + noise), optionally with a
    fraction `special` of values replaced by 0, +-1e-300 and +-1e300
    (doubles) or 0, +-1e-40, +-1e38 (floats).  Integer types are scaled to
    stay well inside the transform's range."""
    rng = np.random.default_rng(seed)
    dims = len(n)
    grids = np.meshgrid(*[np.linspace(0, 1, v, endpoint=False) for v in reversed(n)], indexing="ij")
    f = np.zeros(grids[0].shape)
    for _ in range(8):
        k = rng.integers(1, 7, size=dims)
        ph = rng.uniform(0, 2 * np.pi)
        amp = rng.uniform(0.1, 1.0)
        arg = sum(2 * np.pi * k[d] * grids[dims - 1 - d] for d in range(dims)) + ph
        f += amp * np.sin(arg)
    f += noise * rng.standard_normal(f.shape)
    f = f.reshape(-1)
    if ztype == T_DOUBLE:
        out = f.astype(np.float64)
        if special:
            idx = rng.random(out.size) < special
            choice = rng.integers(0, 5, size=out.size)
            vals = np.array([0.0, 1e-300, -1e-300, 1e300, -1e300])
            out[idx] = vals[choice[idx]]
        return out
    if ztype == T_FLOAT:
        out = f.astype(np.float32)
        if special:
            idx = rng.random(out.size) < special
            choice = rng.integers(0, 5, size=out.size)
            vals = np.array([0.0, 1e-40, -1e-40, 1e38, -1e38], dtype=np.float32)
            out[idx] = vals[choice[idx]]
        return out
    if ztype == T_INT32:
        return np.round(f * (1 << 26)).astype(np.int32)
    return np.round(f * (1 << 58)).astype(np.int64)


# --------------------------------------------------------------------------
# the reference libzfp (oracle/_ref/libzfp_ref.so) through its own C API
# --------------------------------------------------------------------------

_refapi = None


def ref_api():
    """The reference library, with the signatures of include/zfp.h (the same
    table the product binding uses: both export the same names)."""
    global _refapi
    if _refapi is None:
        import zfp_amd
        lib = ctypes.CDLL(os.path.join(REF_DIR, "libzfp_ref.so"))
        for name, (res, args) in zfp_amd._API.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _refapi = lib
    return _refapi


def api_compress(api, base, elem_offset, n, s, ztype, params, header=0, policy=0, threads=0):
    """zfp_compress through a zfp C API (reference or ours) on host memory;
    returns (stream words, size in bytes)."""
    dims = len(n)
    ptr = base.ctypes.data + elem_offset * base.itemsize
    if dims == 1:
        f = api.zfp_field_1d(ctypes.c_void_p(ptr), ztype, n[0])
    elif dims == 2:
        f = api.zfp_field_2d(ctypes.c_void_p(ptr), ztype, n[0], n[1])
    elif dims == 3:
        f = api.zfp_field_3d(ctypes.c_void_p(ptr), ztype, n[0], n[1], n[2])
    else:
        f = api.zfp_field_4d(ctypes.c_void_p(ptr), ztype, n[0], n[1], n[2], n[3])
    if s and any(s):
        ss = list(s) + [0] * (4 - len(s))
        api.zfp_field_set_stride_4d(f, *ss)
    z = api.zfp_stream_open(None)
    assert api.zfp_stream_set_params(z, *params)
    if policy == 1:
        assert api.zfp_stream_set_omp_threads(z, threads)
    cap = api.zfp_stream_maximum_size(z, f)
    buf = np.zeros(cap // 8 + 1, dtype=np.uint64)
    bs = api.stream_open(ctypes.c_void_p(buf.ctypes.data), cap)
    api.zfp_stream_set_bit_stream(z, bs)
    api.zfp_stream_rewind(z)
    if header:
        assert api.zfp_write_header(z, f, header)
    size = api.zfp_compress(z, f)
    api.stream_close(bs)
    api.zfp_stream_close(z)
    api.zfp_field_free(f)
    return buf[: size // 8].copy(), size


def api_decompress(api, words, base, elem_offset, n, s, ztype, params, header=0):
    """zfp_decompress through a zfp C API on host memory; returns bytes consumed."""
    dims = len(n)
    ptr = base.ctypes.data + elem_offset * base.itemsize
    nn = list(n) + [0] * (4 - dims)
    f = api.zfp_field_4d(ctypes.c_void_p(ptr), ztype, *nn)
    if s and any(s):
        ss = list(s) + [0] * (4 - len(s))
        api.zfp_field_set_stride_4d(f, *ss)
    z = api.zfp_stream_open(None)
    assert api.zfp_stream_set_params(z, *params)
    w = np.ascontiguousarray(words, dtype=np.uint64).copy()
    bs = api.stream_open(ctypes.c_void_p(w.ctypes.data), w.size * 8)
    api.zfp_stream_set_bit_stream(z, bs)
    api.zfp_stream_rewind(z)
    if header:
        g = api.zfp_field_alloc()
        assert api.zfp_read_header(z, g, header)
        api.zfp_field_free(g)
    size = api.zfp_decompress(z, f)
    api.stream_close(bs)
    api.zfp_stream_close(z)
    api.zfp_field_free(f)
    return size
