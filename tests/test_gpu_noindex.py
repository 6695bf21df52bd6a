"""Variable-rate streams that arrive without a block index (a file written by
serial zfp, read by the `zfp` tool or zfpy): the decoder recovers every
block's start by following chains of block lengths in parallel segments
(k_chain, codec_kernels.h) and then decodes all blocks in parallel.

* Streams written by the reference library itself (serial zfp_compress,
  oracle/_ref/libzfp_ref.so) in every variable-length mode, including
  reversible and expert modes, 1D-4D, at a bit offset after a header, in
  buffers of exactly the stream's words, followed by arbitrary words, or
  followed by zeros: decompression equals serial zfp_decompress bit for bit
  and consumes exactly the stream's bits.
* The 2D float32 8191 x 8193 field at precision 16, written by serial zfp:
  decoded without an index in less time than serial zfp_decompress takes
  on this host, bit for bit equal.
"""
import ctypes
import time

import numpy as np
import pytest
import torch

import bench
import zfp_amd as Z
from zfptest import (T_DOUBLE, T_FLOAT, T_INT32, T_INT64, api_compress, api_decompress, have_ref, ref_api,
                     smooth_field, stream_params_accuracy, stream_params_precision)

pytestmark = [pytest.mark.gpu, pytest.mark.skipif(not have_ref(), reason="reference library not built")]
SZ4 = ctypes.c_size_t * 4
PD4 = ctypes.c_ssize_t * 4


def _decode(ztype, n, p, d_words, nwords, out, offset=0):
    nn = SZ4(*(list(n) + [0] * (4 - len(n))))
    bits = ctypes.c_uint64()
    rc = Z.hip_lib().zfp_hip_decode(ztype, nn, PD4(0, 0, 0, 0), ctypes.c_void_p(out.data_ptr()), *p,
                                    ctypes.c_void_p(d_words.data_ptr()), nwords, offset, ctypes.byref(bits), None)
    return rc, bits.value


CASES = [
    (T_FLOAT, [257, 131], stream_params_precision(16)),
    (T_FLOAT, [4001], stream_params_accuracy(1e-4)),
    (T_DOUBLE, [40, 36, 28], stream_params_accuracy(1e-6)),
    (T_DOUBLE, [20, 12, 16, 8], stream_params_precision(24)),
    (T_INT32, [61, 67], (1, 16658, 20, -1074)),
    (T_INT64, [21, 18, 23], (1, 16658, 40, -1074)),
    (T_FLOAT, [33, 35, 37], (1, 16658, 64, -1075)),       # reversible
    (T_DOUBLE, [97, 101], (300, 2500, 40, -1075)),         # reversible, minbits > 1
    (T_DOUBLE, [45, 47, 9], (100, 600, 20, -30)),          # expert
]


@pytest.mark.parametrize("tail", ["exact", "garbage", "zeros"])
@pytest.mark.parametrize("k", range(len(CASES)))
def test_serial_stream_without_index(k, tail):
    ztype, n, p = CASES[k]
    print(f"case {k} {ztype} {n} {p} tail {tail}", flush=True)
    a = smooth_field(n, ztype, seed=70 + k, special=0.02 if ztype in (T_FLOAT, T_DOUBLE) else 0.0)
    words, size = api_compress(ref_api(), a, 0, n, None, ztype, p)
    ref = np.zeros_like(a)
    assert api_decompress(ref_api(), words, ref, 0, n, None, ztype, p) == size
    rng = np.random.default_rng(k)
    extra = {"exact": np.zeros(0, dtype=np.uint64),
             "garbage": rng.integers(0, 2 ** 63, size=4096, dtype=np.int64).view(np.uint64),
             "zeros": np.zeros(4096, dtype=np.uint64)}[tail]
    buf = np.concatenate([words, extra])
    d_words = torch.from_numpy(buf.view(np.int64)).to("cuda:0")
    Z.hip_lib().zfp_hip_index_clear()
    out = torch.zeros(a.size, dtype=torch.from_numpy(a).dtype, device="cuda:0")
    rc, bits = _decode(ztype, n, p, d_words, buf.size, out)
    assert rc == 0
    assert (bits + 63) // 64 * 8 == size
    assert np.array_equal(out.cpu().numpy().view(np.uint8), ref.view(np.uint8))


def test_serial_stream_after_header_bits():
    """A stream that starts 77 bits into the buffer (as after a zfp header)."""
    ztype, n, p = T_FLOAT, [301, 203], stream_params_precision(20)
    a = smooth_field(n, ztype, seed=5)
    words, size = api_compress(ref_api(), a, 0, n, None, ztype, p)
    ref = np.zeros_like(a)
    api_decompress(ref_api(), words, ref, 0, n, None, ztype, p)
    # shift the stream by 77 bits
    nw = words.size + 2
    shifted = np.zeros(nw, dtype=np.uint64)
    for i, w in enumerate(words.tolist()):
        shifted[i + 1] |= np.uint64((w << 13) & 0xffffffffffffffff)
        shifted[i + 2] |= np.uint64(w >> 51)
    shifted[1] |= np.uint64(0x1555)  # the header's bits, not zero
    shifted[0] = np.uint64(0xfedcba9876543210)
    d_words = torch.from_numpy(shifted.view(np.int64)).to("cuda:0")
    Z.hip_lib().zfp_hip_index_clear()
    out = torch.zeros(a.size, dtype=torch.float32, device="cuda:0")
    rc, bits = _decode(ztype, n, p, d_words, nw, out, offset=77)
    assert rc == 0 and (bits + 63) // 64 * 8 == size
    assert np.array_equal(out.cpu().numpy().view(np.uint8), ref.view(np.uint8))


def test_full_size_2d_stream_faster_than_serial_decompress():
    """The 2D float32 8191 x 8193 field at precision 16 (the benchmark's
    field), compressed by serial zfp, decoded here without an index."""
    fields = bench.vr_fields(torch)
    ztype, n, f = fields["2d_f32_8191x8193"]
    del fields
    host = f.cpu().numpy().reshape(-1)
    p = Z.stream_params("precision", 16, ztype, 2)
    print("serial compress", flush=True)
    words, size = api_compress(ref_api(), host, 0, n, None, ztype, p)
    ref = np.zeros_like(host)
    print("serial decompress", flush=True)
    t0 = time.perf_counter()
    assert api_decompress(ref_api(), words, ref, 0, n, None, ztype, p) == size
    t_serial = time.perf_counter() - t0
    d_words = torch.from_numpy(words.view(np.int64)).to("cuda:0")
    out = torch.zeros_like(f)
    times = []
    for _ in range(3):
        Z.hip_lib().zfp_hip_index_clear()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        rc, bits = _decode(ztype, n, p, d_words, words.size, out)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        assert rc == 0 and (bits + 63) // 64 * 8 == size
    print(f"no-index decode {min(times) * 1e3:.1f} ms (runs {[round(t * 1e3, 1) for t in times]}), "
          f"serial zfp_decompress {t_serial * 1e3:.0f} ms", flush=True)
    assert np.array_equal(out.cpu().numpy().reshape(-1).view(np.uint32), ref.view(np.uint32))
    assert min(times) < t_serial
    Z.hip_lib().zfp_hip_index_clear()
