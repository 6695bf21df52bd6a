"""zfpy-compatible NumPy interface over libzfp_amd, with an `execution`
keyword that selects the MI355X policy.

The functions, arguments, defaults, exceptions and stream layout follow the
reference's Cython binding (python/zfpy.pyx):

* compress_numpy(arr, tolerance=-1, rate=-1, precision=-1, write_header=True)
  (zfpy.pyx:140-187): at most one of tolerance/rate/precision; none of them
  selects reversible mode; the full header (magic, field metadata, mode) goes
  first unless write_header is False; returns the stream as bytes.
* decompress_numpy(buffer) (zfpy.pyx:335-356): reads the full header and
  returns a new C-ordered array.
* _decompress(buffer, ztype, shape, out=None, tolerance=-1, rate=-1,
  precision=-1) (zfpy.pyx:259-333): a headerless stream with the metadata
  given by the caller, optionally into `out`.
* header(buffer) (zfpy.pyx:358-400): the header as a dict.

The only addition is `execution="hip"` on the compressing and decompressing
functions: the stream is compressed or decompressed under zfp_exec_hip.
This library computes with no other policy, so any other value raises
ValueError.  Arrays are NumPy host arrays (any strides); the policy stages
them through device memory.
"""
import ctypes

import numpy as np

from . import (ZFP_HEADER_FULL, ZFP_HEADER_MAGIC, ZFP_HEADER_META, ZFP_HEADER_MODE, api_lib,
               zfp_exec_hip, zfp_mode_expert, zfp_mode_fixed_accuracy, zfp_mode_fixed_precision,
               zfp_mode_fixed_rate, zfp_mode_null, zfp_mode_reversible, zfp_type_double, zfp_type_float,
               zfp_type_int32, zfp_type_int64, zfp_type_none)

HEADER_MAGIC = ZFP_HEADER_MAGIC
HEADER_META = ZFP_HEADER_META
HEADER_MODE = ZFP_HEADER_MODE
HEADER_FULL = ZFP_HEADER_FULL

type_none = zfp_type_none
type_int32 = zfp_type_int32
type_int64 = zfp_type_int64
type_float = zfp_type_float
type_double = zfp_type_double

mode_null = zfp_mode_null
mode_expert = zfp_mode_expert
mode_fixed_rate = zfp_mode_fixed_rate
mode_fixed_precision = zfp_mode_fixed_precision
mode_fixed_accuracy = zfp_mode_fixed_accuracy
mode_reversible = zfp_mode_reversible

_DTYPE_OF = {zfp_type_int32: np.int32, zfp_type_int64: np.int64, zfp_type_float: np.float32,
             zfp_type_double: np.float64}
_MODE_NAME = {zfp_mode_null: "null", zfp_mode_expert: "expert", zfp_mode_reversible: "reversible",
              zfp_mode_fixed_accuracy: "tolerance", zfp_mode_fixed_precision: "precision",
              zfp_mode_fixed_rate: "rate"}
_POLICIES = {"hip": zfp_exec_hip}


def dtype_to_ztype(dtype):
    for z, d in _DTYPE_OF.items():
        if dtype == d:
            return z
    raise TypeError("Unknown dtype: {}".format(dtype))


def ztype_to_dtype(ztype):
    try:
        return _DTYPE_OF[ztype]
    except KeyError:
        raise ValueError("Unsupported zfp_type {}".format(ztype))


def zmode_to_str(zmode):
    try:
        return _MODE_NAME[zmode]
    except KeyError:
        raise ValueError("Unsupported zfp_mode {}".format(zmode))


def _policy(execution):
    try:
        return _POLICIES[execution]
    except KeyError:
        raise ValueError("Unsupported execution {!r}: this library computes with execution='hip' only"
                         .format(execution))


def _set_mode(api, stream, ztype, ndim, tolerance, rate, precision):
    # the order of zfpy.pyx:205-220: tolerance, rate, precision, else reversible
    if tolerance >= 0:
        api.zfp_stream_set_accuracy(stream, tolerance)
    elif rate >= 0:
        api.zfp_stream_set_rate(stream, rate, ztype, ndim, 0)
    elif precision >= 0:
        api.zfp_stream_set_precision(stream, precision)
    else:
        api.zfp_stream_set_reversible(stream)


def _select(api, stream, execution):
    if not api.zfp_stream_set_execution(stream, _policy(execution)):
        raise RuntimeError("{} execution not available".format(execution))


def _field_of(api, arr):
    """zfp_field over a NumPy array, extents and strides in zfp's x-first
    order (zfpy.pyx:93-120)."""
    ndim = arr.ndim
    if ndim < 1 or ndim > 4:
        raise RuntimeError("Greater than 4 dimensions not supported")
    ztype = dtype_to_ztype(arr.dtype)
    ptr = ctypes.c_void_p(arr.ctypes.data)
    shape = arr.shape[::-1]
    strides = [s // arr.itemsize for s in arr.strides[::-1]]
    make = (api.zfp_field_1d, api.zfp_field_2d, api.zfp_field_3d, api.zfp_field_4d)[ndim - 1]
    field = make(ptr, ztype, *shape)
    (api.zfp_field_set_stride_1d, api.zfp_field_set_stride_2d, api.zfp_field_set_stride_3d,
     api.zfp_field_set_stride_4d)[ndim - 1](field, *strides)
    return field


def _buffer(data):
    """A read-only byte view of `data` and its address (kept alive by the view)."""
    if data is None:
        raise TypeError("compressed_data cannot be None")
    view = np.frombuffer(data, dtype=np.uint8)
    if view.size == 0:
        raise ValueError("compressed_data is empty")
    # the stream reader loads whole 64-bit words: give it an aligned copy
    # padded to a word boundary
    words = np.zeros((view.size + 7) // 8, dtype=np.uint64)
    words.view(np.uint8)[:view.size] = view
    return words


def compress_numpy(arr, tolerance=-1, rate=-1, precision=-1, write_header=True, execution="hip"):
    if arr is None:
        raise TypeError("Input array cannot be None")
    if sum(1 for x in (tolerance, rate, precision) if x >= 0) > 1:
        raise ValueError("Only one of tolerance, rate, or precision can be set")
    api = api_lib()
    arr = np.asarray(arr)
    field = _field_of(api, arr)
    stream = api.zfp_stream_open(None)
    bs = None
    try:
        # zfpy passes zfp_type_none to the rate setter here (zfpy.pyx:160-163)
        _set_mode(api, stream, zfp_type_none, arr.ndim, tolerance, rate, precision)
        _select(api, stream, execution)
        cap = api.zfp_stream_maximum_size(stream, field)
        out = np.zeros((cap + 7) // 8, dtype=np.uint64)
        bs = api.stream_open(ctypes.c_void_p(out.ctypes.data), cap)
        api.zfp_stream_set_bit_stream(stream, bs)
        api.zfp_stream_rewind(stream)
        if write_header and api.zfp_write_header(stream, field, HEADER_FULL) == 0:
            raise RuntimeError("Failed to write header to stream")
        size = api.zfp_compress(stream, field)
        if size == 0:
            raise RuntimeError("Failed to write to stream")
        return out.view(np.uint8)[:size].tobytes()
    finally:
        api.zfp_field_free(field)
        api.zfp_stream_close(stream)
        if bs:
            api.stream_close(bs)


def _open(api, data):
    words = _buffer(data)
    bs = api.stream_open(ctypes.c_void_p(words.ctypes.data), words.nbytes)
    stream = api.zfp_stream_open(bs)
    return words, bs, stream


def decompress_numpy(compressed_data, execution="hip"):
    api = api_lib()
    words, bs, stream = _open(api, compressed_data)
    field = api.zfp_field_alloc()
    try:
        if api.zfp_read_header(stream, field, HEADER_FULL) == 0:
            raise ValueError("Failed to read required zfp header")
        _select(api, stream, execution)
        f = field.contents
        dtype = ztype_to_dtype(f.type)
        shape = tuple(int(x) for x in (f.nw, f.nz, f.ny, f.nx) if x > 0)
        out = np.empty(shape, dtype=dtype)
        api.zfp_field_set_pointer(field, ctypes.c_void_p(out.ctypes.data))
        if api.zfp_decompress(stream, field) == 0:
            raise RuntimeError("error during zfp decompression")
        return out
    finally:
        api.zfp_field_free(field)
        api.zfp_stream_close(stream)
        api.stream_close(bs)
        del words


def _decompress(compressed_data, ztype, shape, out=None, tolerance=-1, rate=-1, precision=-1,
                execution="hip"):
    if compressed_data is None:
        raise TypeError("compressed_data cannot be None")
    if compressed_data is out:
        raise ValueError("Cannot decompress in-place")
    try:
        if len(shape) > 4:
            raise ValueError("User-provided shape has too many dimensions (up to 4 supported)")
        if len(shape) <= 0:
            raise ValueError("User-provided shape needs at least one dimension")
    except TypeError:
        raise TypeError("User-provided shape is not an iterable")
    api = api_lib()
    words, bs, stream = _open(api, compressed_data)
    field = api.zfp_field_alloc()
    try:
        api.zfp_stream_rewind(stream)
        zshape = [int(x) for x in reversed(list(shape))] + [0] * (4 - len(shape))
        f = field.contents
        f.nx, f.ny, f.nz, f.nw = zshape
        api.zfp_field_set_type(field, ztype)
        ndim = sum(1 for x in zshape if x > 0)
        _set_mode(api, stream, ztype, ndim, tolerance, rate, precision)
        _select(api, stream, execution)
        dtype = ztype_to_dtype(ztype)
        if out is None:
            output = np.empty(tuple(x for x in shape if x > 0), dtype=dtype)
        elif isinstance(out, np.ndarray):
            output = out
            if out.dtype != dtype:
                raise ValueError("Out ndarray has dtype {} but decompression is using {}. Use "
                                 "out=ndarray.data to avoid this check.".format(out.dtype, dtype))
            user_shape = [x for x in shape if x > 0]
            if list(out.shape) != list(user_shape):
                raise ValueError("Out ndarray has shape {} but decompression is using {}.  Use "
                                 "out=ndarray.data to avoid this check.".format(out.shape, user_shape))
        else:
            output = np.frombuffer(out, dtype=dtype).reshape(shape)
        if not output.flags.c_contiguous:
            raise ValueError("out must be C-contiguous")
        api.zfp_field_set_pointer(field, ctypes.c_void_p(output.ctypes.data))
        if api.zfp_decompress(stream, field) == 0:
            raise RuntimeError("error during zfp decompression")
        return output
    finally:
        api.zfp_field_free(field)
        api.zfp_stream_close(stream)
        api.stream_close(bs)
        del words


def header(compressed_data):
    """Stream header information as a dict (zfpy.pyx:358-400)."""
    api = api_lib()
    words, bs, stream = _open(api, compressed_data)
    field = api.zfp_field_alloc()
    try:
        if api.zfp_read_header(stream, field, HEADER_FULL) == 0:
            raise ValueError("Failed to read required zfp header")
        mode = api.zfp_stream_compression_mode(stream)
        f = field.contents
        ndim = sum(1 for x in (f.nx, f.ny, f.nz, f.nw) if x > 0)
        minbits, maxbits, maxprec, minexp = (ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint(), ctypes.c_int())
        api.zfp_stream_params(stream, ctypes.byref(minbits), ctypes.byref(maxbits), ctypes.byref(maxprec),
                              ctypes.byref(minexp))
        return {
            "nx": int(f.nx), "ny": int(f.ny), "nz": int(f.nz), "nw": int(f.nw),
            "type": ztype_to_dtype(f.type),
            "mode": zmode_to_str(mode),
            "config": {
                "mode": int(mode),
                "tolerance": float(api.zfp_stream_accuracy(stream)),
                "rate": float(api.zfp_stream_rate(stream, ndim)),
                "precision": int(api.zfp_stream_precision(stream)),
                # zfpy reports minbits under "maxbits" too (zfpy.pyx:392-397)
                "expert": {"minbits": int(minbits.value), "maxbits": int(minbits.value),
                           "maxprec": int(maxprec.value), "minexp": int(minexp.value)},
            },
        }
    finally:
        api.zfp_field_free(field)
        api.zfp_stream_close(stream)
        api.stream_close(bs)
        del words
