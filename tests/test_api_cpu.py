"""The C libraries without a GPU: they load, export every function their
headers declare, and the host-side API (fields, streams, modes, headers, bit
stream I/O) behaves like the reference library's (oracle/_ref/libzfp_ref.so)
call for call.  Compression with zfp_exec_hip on a machine without a gfx950
device declines with 0 and leaves the stream untouched, as the reference's
CUDA tests require of unsupported requests
(tests/src/endtoend/cudaExecBase.c:36-163)."""
import ctypes
import os
import re

import numpy as np
import pytest

import zfp_amd as Z
from zfptest import ROOT, have_ref, ref_api

needs_ref = pytest.mark.skipif(not have_ref(), reason="oracle/_ref (reference build) not present")


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = set(re.findall(r"\b((?:zfp|stream)_[a-z0-9_]+)\s*\(", text))
    return names


def test_headers_declare_what_the_libraries_export():
    hip = ctypes.CDLL(Z.LIB_HIP)
    api = ctypes.CDLL(Z.LIB_API)
    for name in sorted(_declared("zfp_hip.h")):
        assert hasattr(hip, name), f"libzfp_hip.so does not export {name}"
    for header in ("zfp.h", os.path.join("zfp", "bitstream.h")):
        for name in sorted(_declared(header)):
            if name.startswith("zfp_hip"):
                continue
            assert hasattr(api, name), f"libzfp_amd.so does not export {name}"
    # the bindings declare only exported functions
    for lib, names in ((hip, Z.exported_symbols()["libzfp_hip.so"]), (api, Z.exported_symbols()["libzfp_amd.so"])):
        for name in names:
            assert hasattr(lib, name)


def test_stream_word_bits_and_version():
    api = ctypes.CDLL(Z.LIB_API)
    assert ctypes.c_size_t.in_dll(api, "stream_word_bits").value == 64
    assert ctypes.c_uint.in_dll(api, "zfp_codec_version").value == 5
    assert Z.hip_lib().zfp_hip_version()


def _both():
    return [("ours", Z.api_lib()), ("ref", ref_api())]


@needs_ref
def test_mode_setters_match_reference():
    results = {}
    for tag, api in _both():
        out = []
        z = api.zfp_stream_open(None)
        for rate in (0.5, 1, 4, 8, 13.7, 16, 32, 64):
            for ztype in (1, 2, 3, 4):
                for dims in (1, 2, 3, 4):
                    for align in (0, 1):
                        r = api.zfp_stream_set_rate(z, rate, ztype, dims, align)
                        out.append((r, api.zfp_stream_mode(z), api.zfp_stream_compression_mode(z),
                                    api.zfp_stream_rate(z, dims)))
        for prec in (0, 1, 5, 16, 64, 100):
            out.append((api.zfp_stream_set_precision(z, prec), api.zfp_stream_mode(z),
                        api.zfp_stream_precision(z)))
        for tol in (0.0, 1e-300, 1e-7, 1e-3, 0.5, 1.0, 3.0, 1e10):
            out.append((api.zfp_stream_set_accuracy(z, tol), api.zfp_stream_mode(z),
                        api.zfp_stream_accuracy(z)))
        api.zfp_stream_set_reversible(z)
        out.append((api.zfp_stream_mode(z), api.zfp_stream_compression_mode(z)))
        for p in ((1, 16658, 64, -1074), (100, 600, 20, -30), (5, 4, 10, 0), (1, 2, 0, 0), (1, 2, 65, 0),
                  (7, 40000, 128, 2000)):
            ok = api.zfp_stream_set_params(z, *p)
            out.append((ok, api.zfp_stream_mode(z), api.zfp_stream_compression_mode(z)))
        for mode in (0, 1, 2047, 2048, 2100, 2175, 2176, 2177, 3000, 4094, 0x123456789abcdfff):
            out.append((api.zfp_stream_set_mode(z, mode), api.zfp_stream_mode(z)))
        api.zfp_stream_close(z)
        results[tag] = out
    assert results["ours"] == results["ref"]


@needs_ref
def test_field_functions_match_reference():
    results = {}
    buf = np.zeros(100000, dtype=np.float64)
    for tag, api in _both():
        out = []
        for n in ([7], [5, 9], [4, 6, 3], [3, 4, 5, 6], [1 << 20], [70000, 3]):
            for ztype in (1, 2, 3, 4):
                f = api.zfp_field_4d(ctypes.c_void_p(buf.ctypes.data + 8 * 5000), ztype, *(n + [0] * (4 - len(n))))
                for strides in ([0, 0, 0, 0], [-1, 0, 0, 0], [2, 30, 500, 0], [-3, -40, 700, 9]):
                    api.zfp_field_set_stride_4d(f, *strides)
                    meta = api.zfp_field_metadata(f)
                    out.append((api.zfp_field_dimensionality(f), api.zfp_field_blocks(f), meta,
                                api.zfp_field_size_bytes(f),
                                (api.zfp_field_begin(f) or 0) - buf.ctypes.data))
                    g = api.zfp_field_alloc()
                    if meta != 0xffffffffffffffff:
                        out.append(api.zfp_field_set_metadata(g, meta))
                        out.append(api.zfp_field_metadata(g))
                    api.zfp_field_free(g)
                z = api.zfp_stream_open(None)
                for p in ((64, 64, 64, -1074), (1, 16658, 12, -1074), (1, 16658, 64, -20), (1, 16658, 64, -1075)):
                    api.zfp_stream_set_params(z, *p)
                    out.append(api.zfp_stream_maximum_size(z, f))
                api.zfp_stream_close(z)
                api.zfp_field_free(f)
        results[tag] = out
    assert results["ours"] == results["ref"]


@needs_ref
def test_headers_match_reference():
    results = {}
    for tag, api in _both():
        out = []
        for n, ztype in (([100], 3), ([33, 17], 4), ([5, 6, 7], 1), ([9, 8, 7, 6], 2)):
            f = api.zfp_field_4d(None, ztype, *(n + [0] * (4 - len(n))))
            for p in ((64, 64, 64, -1074), (1, 16658, 12, -1074), (1, 16658, 64, -20), (1, 16658, 64, -1075),
                      (100, 600, 20, -30), (3000, 3000, 64, -1074)):
                z = api.zfp_stream_open(None)
                api.zfp_stream_set_params(z, *p)
                for mask in (1, 2, 4, 7):
                    buf = np.zeros(8, dtype=np.uint64)
                    bs = api.stream_open(ctypes.c_void_p(buf.ctypes.data), 64)
                    api.zfp_stream_set_bit_stream(z, bs)
                    bits = api.zfp_write_header(z, f, mask)
                    api.zfp_stream_flush(z)
                    out.append((bits, tuple(buf)))
                    api.zfp_stream_rewind(z)
                    g = api.zfp_field_alloc()
                    z2 = api.zfp_stream_open(bs)
                    rbits = api.zfp_read_header(z2, g, mask)
                    mn, mx, mp, me = ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint(), ctypes.c_int()
                    api.zfp_stream_params(z2, ctypes.byref(mn), ctypes.byref(mx), ctypes.byref(mp), ctypes.byref(me))
                    out.append((rbits, api.zfp_field_metadata(g) if mask & 2 else 0, mn.value, mx.value, mp.value,
                                me.value))
                    api.zfp_stream_close(z2)
                    api.zfp_field_free(g)
                    api.stream_close(bs)
                api.zfp_stream_close(z)
            api.zfp_field_free(f)
        results[tag] = out
    assert results["ours"] == results["ref"]


@needs_ref
def test_bitstream_ops_match_reference():
    rng = np.random.default_rng(5)
    ops = []
    for _ in range(3000):
        k = int(rng.integers(0, 6))
        n = int(rng.integers(0, 65))
        ops.append((k, n, int(rng.integers(0, 1 << 63)) * 2 + int(rng.integers(0, 2))))
    results = {}
    for tag, api in _both():
        buf = np.zeros(4096, dtype=np.uint64)
        bs = api.stream_open(ctypes.c_void_p(buf.ctypes.data), buf.nbytes)
        trace = []
        for k, n, v in ops[:1500]:
            if k in (0, 1, 2):
                trace.append(api.stream_write_bits(bs, v, n))
            elif k == 3:
                trace.append(api.stream_write_bit(bs, v & 1))
            elif k == 4:
                api.stream_pad(bs, n)
            else:
                trace.append(api.stream_flush(bs))
            trace.append(api.stream_wtell(bs))
        trace.append(api.stream_flush(bs))
        trace.append(api.stream_size(bs))
        api.stream_rewind(bs)
        for k, n, v in ops[1500:]:
            if k in (0, 1, 2):
                trace.append(api.stream_read_bits(bs, n))
            elif k == 3:
                trace.append(api.stream_read_bit(bs))
            elif k == 4:
                api.stream_skip(bs, n)
            else:
                trace.append(api.stream_align(bs))
            trace.append(api.stream_rtell(bs))
        api.stream_rseek(bs, 12345)
        trace.append(api.stream_read_bits(bs, 40))
        api.stream_wseek(bs, 777)
        api.stream_write_bits(bs, 0x5a5a, 16)
        api.stream_flush(bs)
        trace.append(tuple(buf[:64]))
        api.stream_close(bs)
        results[tag] = trace
    assert results["ours"] == results["ref"]


def test_execution_policy_contract():
    api = Z.api_lib()
    z = api.zfp_stream_open(None)
    assert api.zfp_stream_execution(z) == Z.zfp_exec_serial
    assert api.zfp_stream_set_execution(z, Z.zfp_exec_hip)
    assert api.zfp_stream_execution(z) == Z.zfp_exec_hip
    assert not api.zfp_stream_set_execution(z, Z.zfp_exec_cuda)
    assert api.zfp_stream_execution(z) == Z.zfp_exec_hip
    assert api.zfp_stream_set_omp_threads(z, 4) and api.zfp_stream_omp_threads(z) == 4
    assert api.zfp_stream_set_execution(z, Z.zfp_exec_serial)
    api.zfp_stream_close(z)


def test_unsupported_requests_leave_stream_untouched():
    """Serial execution is not provided by this library, and zfp_exec_hip
    without a usable device declines: both return 0 without touching the
    stream."""
    api = Z.api_lib()
    a = np.arange(64, dtype=np.float64)
    f = api.zfp_field_3d(ctypes.c_void_p(a.ctypes.data), Z.zfp_type_double, 4, 4, 4)
    z = api.zfp_stream_open(None)
    api.zfp_stream_set_rate(z, 8, Z.zfp_type_double, 3, 0)
    buf = np.full(64, 0x1234, dtype=np.uint64)
    bs = api.stream_open(ctypes.c_void_p(buf.ctypes.data), buf.nbytes)
    api.zfp_stream_set_bit_stream(z, bs)
    assert api.zfp_compress(z, f) == 0  # serial policy
    if Z.hip_lib().zfp_hip_device_check() != 0:
        assert api.zfp_stream_set_execution(z, Z.zfp_exec_hip)
        assert api.zfp_compress(z, f) == 0
        assert api.zfp_decompress(z, f) == 0
        assert Z.hip_lib().zfp_hip_last_error()
    assert api.stream_wtell(bs) == 0
    assert (buf == 0x1234).all()
    api.stream_close(bs)
    api.zfp_stream_close(z)
    api.zfp_field_free(f)
