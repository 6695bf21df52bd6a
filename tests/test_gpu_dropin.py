"""The drop-in, proven on the reference's own code: oracle/_ref/libzfp_hiphook.so
is the reference library (its src/*.c, including its own src/zfp.c) with the
zfp_exec_hip policy added by the integration guide's three hooks, calling
this repository's libzfp_hip.so.  Through the reference's zfp_compress /
zfp_decompress, the hip policy must give what the reference's serial policy
gives in the same library: the same stream bytes, the same returned sizes,
the same decompressed values -- for every mode, with and without a header in
front of the stream (src/zfp.c:917-955, 1051-1180;
tests/src/endtoend/zfpEndtoendBase.c:386-464 checks the same for its
policies)."""
import ctypes
import os

import numpy as np
import pytest

import zfp_amd as Z
from zfptest import (ROOT, T_DOUBLE, T_FLOAT, T_INT32, T_INT64, make_layout, smooth_field, stream_params_accuracy,
                     stream_params_precision, stream_params_rate)

pytestmark = pytest.mark.gpu
HOOK = os.path.join(ROOT, "oracle", "_ref", "libzfp_hiphook.so")
needs_hook = pytest.mark.skipif(not os.path.exists(HOOK), reason="oracle/_ref/libzfp_hiphook.so not built")

_lib = None


def hooked():
    global _lib
    if _lib is None:
        Z.hip_lib()   # the codec library first, bound to torch's HIP runtime
        lib = ctypes.CDLL(HOOK)
        for name, (res, args) in Z._API.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def _field(lib, ptr, ztype, n, s):
    nn = list(n) + [0] * (4 - len(n))
    f = lib.zfp_field_4d(ctypes.c_void_p(ptr), ztype, *nn)
    if s and any(s):
        lib.zfp_field_set_stride_4d(f, *(list(s) + [0] * (4 - len(s))))
    return f


def _compress(lib, policy, base, off, n, s, ztype, params, header):
    f = _field(lib, base.ctypes.data + off * base.itemsize, ztype, n, s)
    z = lib.zfp_stream_open(None)
    assert lib.zfp_stream_set_params(z, *params)
    assert lib.zfp_stream_set_execution(z, policy)
    assert lib.zfp_stream_execution(z) == policy
    cap = lib.zfp_stream_maximum_size(z, f)
    buf = np.zeros(cap // 8 + 1, dtype=np.uint64)
    bs = lib.stream_open(ctypes.c_void_p(buf.ctypes.data), cap)
    lib.zfp_stream_set_bit_stream(z, bs)
    lib.zfp_stream_rewind(z)
    if header:
        assert lib.zfp_write_header(z, f, Z.ZFP_HEADER_FULL)
    size = lib.zfp_compress(z, f)
    lib.stream_close(bs)
    lib.zfp_stream_close(z)
    lib.zfp_field_free(f)
    return buf[: size // 8].copy(), size


def _decompress(lib, policy, words, like, off, n, s, ztype, params, header):
    out = np.zeros_like(like)
    f = _field(lib, out.ctypes.data + off * out.itemsize, ztype, n, s)
    z = lib.zfp_stream_open(None)
    assert lib.zfp_stream_set_params(z, *params)
    assert lib.zfp_stream_set_execution(z, policy)
    w = words.copy()
    bs = lib.stream_open(ctypes.c_void_p(w.ctypes.data), w.size * 8)
    lib.zfp_stream_set_bit_stream(z, bs)
    lib.zfp_stream_rewind(z)
    if header:
        g = lib.zfp_field_alloc()
        assert lib.zfp_read_header(z, g, Z.ZFP_HEADER_FULL)
        lib.zfp_field_free(g)
    size = lib.zfp_decompress(z, f)
    lib.stream_close(bs)
    lib.zfp_stream_close(z)
    lib.zfp_field_free(f)
    return size, out


CASES = [
    ("3d f64 rate 8", T_DOUBLE, [19, 17, 11], stream_params_rate(8, T_DOUBLE, 3), "as_is"),
    ("3d f64 rate 16 permuted", T_DOUBLE, [16, 12, 20], stream_params_rate(16, T_DOUBLE, 3), "permuted_neg"),
    ("2d f32 precision 16", T_FLOAT, [61, 47], stream_params_precision(16), "as_is"),
    ("3d f64 accuracy 1e-5", T_DOUBLE, [21, 22, 23], stream_params_accuracy(1e-5), "reversed"),
    ("1d i32 expert", T_INT32, [1001], (80, 400, 22, -1074), "as_is"),
    ("2d i64 rate 21.5", T_INT64, [30, 29], stream_params_rate(21.5, T_INT64, 2), "interleaved"),
    ("4d f64 precision 24", T_DOUBLE, [9, 8, 7, 6], stream_params_precision(24), "as_is"),
    ("3d f32 reversible", T_FLOAT, [13, 14, 15], (1, 16658, 64, -1075), "as_is"),
]


@needs_hook
@pytest.mark.parametrize("header", [0, 1], ids=["raw", "header"])
@pytest.mark.parametrize("k", range(len(CASES)), ids=[c[0] for c in CASES])
def test_reference_library_with_hip_policy(k, header):
    lib = hooked()
    name, ztype, n, params, layout = CASES[k]
    a = smooth_field(n, ztype, seed=800 + k, special=0.01 if ztype == T_DOUBLE else 0.0)
    base, off, s = make_layout(a, n, layout)
    ws, ss = _compress(lib, Z.zfp_exec_serial, base, off, n, s, ztype, params, header)
    wh, sh = _compress(lib, Z.zfp_exec_hip, base, off, n, s, ztype, params, header)
    assert ss > 0 and sh == ss, f"{name}: hip policy returned {sh}, serial {ss}"
    assert np.array_equal(wh, ws), f"{name}: streams differ"
    ds, outs = _decompress(lib, Z.zfp_exec_serial, ws, base, off, n, s, ztype, params, header)
    dh, outh = _decompress(lib, Z.zfp_exec_hip, ws, base, off, n, s, ztype, params, header)
    assert dh == ds, f"{name}: hip decompression returned {dh}, serial {ds}"
    assert np.array_equal(outh.view(np.uint8), outs.view(np.uint8)), f"{name}: decompressed values differ"


@needs_hook
def test_policy_switching_keeps_the_reference_semantics():
    """zfp_stream_set_execution admits zfp_exec_hip next to the reference's
    policies; leaving OpenMP for hip frees OpenMP's parameters (the
    reference's rule for serial) and the OpenMP setters work again after."""
    lib = hooked()
    z = lib.zfp_stream_open(None)
    assert lib.zfp_stream_execution(z) == Z.zfp_exec_serial
    assert lib.zfp_stream_set_execution(z, Z.zfp_exec_hip)
    assert lib.zfp_stream_execution(z) == Z.zfp_exec_hip
    assert lib.zfp_stream_set_omp_threads(z, 3)
    assert lib.zfp_stream_execution(z) == Z.zfp_exec_omp and lib.zfp_stream_omp_threads(z) == 3
    assert lib.zfp_stream_set_execution(z, Z.zfp_exec_hip)
    assert lib.zfp_stream_omp_threads(z) == 0
    assert not lib.zfp_stream_set_execution(z, 7)
    assert lib.zfp_stream_execution(z) == Z.zfp_exec_hip
    assert lib.zfp_stream_set_execution(z, Z.zfp_exec_serial)
    lib.zfp_stream_close(z)
