"""Drives the product (zfp_amd's C API with zfp_exec_hip) from the tests, the
way the reference's end-to-end tests drive zfp_compress/zfp_decompress
(tests/src/endtoend/zfpEndtoendBase.c): open a zfp_stream on a buffer of
zfp_stream_maximum_size bytes, set the mode and the policy, compress, rewind,
decompress.  Device buffers are torch tensors (allocation only)."""
import ctypes

import numpy as np

import zfp_amd as Z
from zfptest import NP_OF_TYPE


def _torch():
    import torch
    return torch


def to_device(arr):
    torch = _torch()
    return torch.from_numpy(np.ascontiguousarray(arr)).to("cuda:0")


def make_field(api, ptr, ztype, n, s):
    dims = len(n)
    if dims == 1:
        f = api.zfp_field_1d(ctypes.c_void_p(ptr), ztype, n[0])
        if s and any(s):
            api.zfp_field_set_stride_1d(f, s[0])
    elif dims == 2:
        f = api.zfp_field_2d(ctypes.c_void_p(ptr), ztype, n[0], n[1])
        if s and any(s):
            api.zfp_field_set_stride_2d(f, s[0], s[1])
    elif dims == 3:
        f = api.zfp_field_3d(ctypes.c_void_p(ptr), ztype, n[0], n[1], n[2])
        if s and any(s):
            api.zfp_field_set_stride_3d(f, s[0], s[1], s[2])
    else:
        f = api.zfp_field_4d(ctypes.c_void_p(ptr), ztype, n[0], n[1], n[2], n[3])
        if s and any(s):
            api.zfp_field_set_stride_4d(f, s[0], s[1], s[2], s[3])
    return f


class Run:
    """One compress + decompress through zfp_compress/zfp_decompress with
    zfp_exec_hip.  base/elem_offset/s describe the layout (see
    zfptest.make_layout); header is a ZFP_HEADER_* mask written first."""

    def __init__(self, base, elem_offset, n, s, ztype, params, header=0, device=True, device_stream=None):
        self.api = api = Z.api_lib()
        self.n, self.s, self.ztype, self.params = list(n), s, ztype, params
        self.device = device
        self.device_stream = device if device_stream is None else device_stream
        itemsize = np.dtype(NP_OF_TYPE[ztype]).itemsize
        if device:
            self.dbase = to_device(base)
            ptr = self.dbase.data_ptr() + elem_offset * itemsize
        else:
            self.hbase = np.ascontiguousarray(base).copy()
            ptr = self.hbase.ctypes.data + elem_offset * itemsize
        self.elem_offset = elem_offset
        self.field = make_field(api, ptr, ztype, n, s)
        self.zfp = api.zfp_stream_open(None)
        assert api.zfp_stream_set_params(self.zfp, *params)
        assert api.zfp_stream_set_execution(self.zfp, Z.zfp_exec_hip)
        self.cap = api.zfp_stream_maximum_size(self.zfp, self.field)
        nwords = self.cap // 8 + 1
        if self.device_stream:
            torch = _torch()
            self.dbuf = torch.zeros(nwords, dtype=torch.int64, device="cuda:0")
            bptr = self.dbuf.data_ptr()
        else:
            self.hbuf = np.zeros(nwords, dtype=np.uint64)
            bptr = self.hbuf.ctypes.data
        self.bs = api.stream_open(ctypes.c_void_p(bptr), self.cap)
        api.zfp_stream_set_bit_stream(self.zfp, self.bs)
        self.header = header
        self.header_bits = 0

    def compress(self):
        api = self.api
        api.zfp_stream_rewind(self.zfp)
        if self.header:
            self.header_bits = api.zfp_write_header(self.zfp, self.field, self.header)
            assert self.header_bits > 0
        self.size = api.zfp_compress(self.zfp, self.field)
        self.reason = Z.hip_lib().zfp_hip_last_error().decode()
        return self.size

    def stream_words(self):
        if self.device_stream:
            return self.dbuf[: self.size // 8].cpu().numpy().view(np.uint64).copy()
        return self.hbuf[: self.size // 8].copy()

    def decompress(self, out_base):
        """Decompress into a fresh buffer shaped like out_base; returns
        (consumed bytes, buffer as numpy)."""
        api = self.api
        itemsize = np.dtype(NP_OF_TYPE[self.ztype]).itemsize
        if self.device:
            dout = to_device(out_base)
            ptr = dout.data_ptr() + self.elem_offset * itemsize
        else:
            hout = np.ascontiguousarray(out_base).copy()
            ptr = hout.ctypes.data + self.elem_offset * itemsize
        api.zfp_field_set_pointer(self.field, ctypes.c_void_p(ptr))
        api.zfp_stream_rewind(self.zfp)
        if self.header:
            f2 = api.zfp_field_alloc()
            got = api.zfp_read_header(self.zfp, f2, self.header)
            api.zfp_field_free(f2)
            assert got == self.header_bits
        consumed = api.zfp_decompress(self.zfp, self.field)
        res = dout.cpu().numpy() if self.device else hout
        return consumed, res

    def close(self):
        api = self.api
        api.stream_close(self.bs)
        api.zfp_stream_close(self.zfp)
        api.zfp_field_free(self.field)
