"""zfp_hip_encode_noalloc: the allocation-free variant of zfp_hip_encode (no
counterpart in the reference; for callers that cannot allocate device
memory).  Its streams must equal zfp_hip_encode's -- and so serial zfp's --
bit for bit, call after call on the shared device scratch, and the calls
must leave the device's free memory unchanged."""
import ctypes

import numpy as np
import pytest

import zfp_amd as Z
from zfptest import (T_DOUBLE, T_FLOAT, T_INT32, oracle_compress, smooth_field, stream_params_accuracy,
                     stream_params_precision, stream_params_rate)

pytestmark = pytest.mark.gpu

SZ4 = ctypes.c_size_t * 4
PD4 = ctypes.c_ssize_t * 4


def _call(fn, ztype, n, p, d_a, d_w, offset=0):
    bits = ctypes.c_uint64()
    rc = fn(ztype, SZ4(*(list(n) + [0] * (4 - len(n)))), PD4(0, 0, 0, 0), ctypes.c_void_p(d_a.data_ptr()), *p,
            ctypes.c_void_p(d_w.data_ptr()), offset, ctypes.byref(bits), None)
    return rc, bits.value


CASES = [
    # (type, first shape, second (smaller) shape, params)
    (T_FLOAT, [203, 171], [61, 47], stream_params_precision(16)),
    (T_DOUBLE, [40, 36, 28], [13, 9, 11], stream_params_accuracy(1e-7)),
    (T_DOUBLE, [24, 20, 16, 12], [9, 8, 7, 5], stream_params_precision(28)),
    (T_INT32, [1000], [77], (100, 900, 24, -40)),
    (T_DOUBLE, [64, 64, 64], [20, 24, 28], stream_params_rate(16, T_DOUBLE, 3)),
    (T_FLOAT, [33, 29, 7], [5, 6, 7], (1, 16658, 64, -1075)),   # reversible
]


@pytest.fixture(scope="module")
def warmed():
    """One call up front: the first call on a device looks up the scratch
    array's device address (and HIP may load the code object holding it)."""
    import torch
    a = torch.zeros(16, dtype=torch.float32, device="cuda:0")
    w = torch.zeros(64, dtype=torch.int64, device="cuda:0")
    rc, _ = _call(Z.hip_lib().zfp_hip_encode_noalloc, T_FLOAT, [16], stream_params_precision(8), a, w)
    assert rc == 0
    torch.cuda.synchronize()


@pytest.mark.parametrize("k", range(len(CASES)))
def test_noalloc_matches_encode(warmed, k):
    import torch
    ztype, n1, n2, p = CASES[k]
    lib = Z.hip_lib()
    for call, n in enumerate((n1, n2)):   # two calls in a row, the second smaller
        a = smooth_field(n, ztype, seed=700 + 10 * k + call)
        ew, eb = oracle_compress(a, 0, n, None, ztype, p)
        d_a = torch.from_numpy(a).to("cuda:0")
        w_ref = torch.full((ew.size + 4,), -1, dtype=torch.int64, device="cuda:0")
        w_new = torch.full((ew.size + 4,), -1, dtype=torch.int64, device="cuda:0")
        rc, bits_ref = _call(lib.zfp_hip_encode, ztype, n, p, d_a, w_ref)
        assert rc == 0 and bits_ref == eb
        torch.cuda.synchronize()
        free0, _ = torch.cuda.mem_get_info()
        rc, bits_new = _call(lib.zfp_hip_encode_noalloc, ztype, n, p, d_a, w_new)
        torch.cuda.synchronize()
        free1, _ = torch.cuda.mem_get_info()
        assert rc == 0, lib.zfp_hip_last_error().decode()
        assert bits_new == bits_ref
        nw = (bits_new + 63) // 64
        assert torch.equal(w_new[:nw], w_ref[:nw]), f"call {call}: stream differs from zfp_hip_encode's"
        assert np.array_equal(w_new[:nw].cpu().numpy().view(np.uint64), ew[:nw])
        assert free1 == free0, "the allocation-free encode changed the device's free memory"


def test_noalloc_refuses_fields_beyond_its_scratch():
    import torch
    lib = Z.hip_lib()
    n = [4 * 2300, 4 * 2200]   # 5,060,000 blocks
    d_a = torch.zeros(1, dtype=torch.float32, device="cuda:0")
    d_w = torch.zeros(1, dtype=torch.int64, device="cuda:0")
    rc, _ = _call(lib.zfp_hip_encode_noalloc, T_FLOAT, n, stream_params_precision(16), d_a, d_w)
    assert rc == Z.ZFP_HIP_EINVAL
    assert "scratch" in lib.zfp_hip_last_error().decode()
