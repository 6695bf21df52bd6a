"""zfp_amd -- zfp's block codec on AMD Instinct MI355X (gfx950).

The product is two in-tree shared libraries built from zfp_amd/csrc:

* libzfp_hip.so  the gfx950 codec behind the C ABI of include/zfp_hip.h
  (zfp_hip_compress / zfp_hip_decompress take zfp_stream/zfp_field like the
  reference's cuda_compress / cuda_decompress; zfp_hip_encode / zfp_hip_decode
  take device pointers and plain sizes);
* libzfp_amd.so  zfp's high-level C API (include/zfp.h, same names, structs
  and return values as LLNL zfp 1.0.1 src/zfp.c) with the zfp_exec_hip policy.

This module loads them with ctypes and declares their signatures; it is a
binding, not a second implementation.  Loading fails loudly when the
libraries are missing: there is no CPU fallback.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_HIP = os.path.join(HERE, "libzfp_hip.so")
LIB_API = os.path.join(HERE, "libzfp_amd.so")

# enums of include/zfp.h
zfp_exec_serial, zfp_exec_omp, zfp_exec_cuda, zfp_exec_hip = 0, 1, 2, 3
zfp_type_none, zfp_type_int32, zfp_type_int64, zfp_type_float, zfp_type_double = 0, 1, 2, 3, 4
(zfp_mode_null, zfp_mode_expert, zfp_mode_fixed_rate, zfp_mode_fixed_precision,
 zfp_mode_fixed_accuracy, zfp_mode_reversible) = range(6)
ZFP_HEADER_NONE, ZFP_HEADER_MAGIC, ZFP_HEADER_META, ZFP_HEADER_MODE, ZFP_HEADER_FULL = 0, 1, 2, 4, 7
ZFP_MIN_BITS, ZFP_MAX_BITS, ZFP_MAX_PREC, ZFP_MIN_EXP = 1, 16658, 64, -1074

ZFP_HIP_OK, ZFP_HIP_EINVAL, ZFP_HIP_ENODEV, ZFP_HIP_ENOMEM, ZFP_HIP_ERUNTIME, ZFP_HIP_ENOINDEX = range(6)


class zfp_execution(ctypes.Structure):
    _fields_ = [("policy", ctypes.c_int), ("params", ctypes.c_void_p)]


class zfp_stream(ctypes.Structure):
    """include/zfp.h:86-93 of the reference."""
    _fields_ = [("minbits", ctypes.c_uint), ("maxbits", ctypes.c_uint), ("maxprec", ctypes.c_uint),
                ("minexp", ctypes.c_int), ("stream", ctypes.c_void_p), ("exec", zfp_execution)]


class zfp_field(ctypes.Structure):
    """include/zfp.h:131-136 of the reference."""
    _fields_ = [("type", ctypes.c_int), ("nx", ctypes.c_size_t), ("ny", ctypes.c_size_t),
                ("nz", ctypes.c_size_t), ("nw", ctypes.c_size_t), ("sx", ctypes.c_ssize_t),
                ("sy", ctypes.c_ssize_t), ("sz", ctypes.c_ssize_t), ("sw", ctypes.c_ssize_t),
                ("data", ctypes.c_void_p)]


class bitstream(ctypes.Structure):
    """State layout of include/zfp/bitstream.inl:133-143 of the reference."""
    _fields_ = [("bits", ctypes.c_size_t), ("buffer", ctypes.c_uint64), ("ptr", ctypes.c_void_p),
                ("begin", ctypes.c_void_p), ("end", ctypes.c_void_p)]


_P = ctypes.POINTER
zs_p = _P(zfp_stream)
zf_p = _P(zfp_field)
bs_p = _P(bitstream)

# name: (restype, argtypes)
_API = {
    "zfp_type_size": (ctypes.c_size_t, [ctypes.c_int]),
    "zfp_stream_open": (zs_p, [bs_p]),
    "zfp_stream_close": (None, [zs_p]),
    "zfp_stream_bit_stream": (bs_p, [zs_p]),
    "zfp_stream_compression_mode": (ctypes.c_int, [zs_p]),
    "zfp_stream_rate": (ctypes.c_double, [zs_p, ctypes.c_uint]),
    "zfp_stream_precision": (ctypes.c_uint, [zs_p]),
    "zfp_stream_accuracy": (ctypes.c_double, [zs_p]),
    "zfp_stream_mode": (ctypes.c_uint64, [zs_p]),
    "zfp_stream_params": (None, [zs_p, _P(ctypes.c_uint), _P(ctypes.c_uint), _P(ctypes.c_uint), _P(ctypes.c_int)]),
    "zfp_stream_compressed_size": (ctypes.c_size_t, [zs_p]),
    "zfp_stream_maximum_size": (ctypes.c_size_t, [zs_p, zf_p]),
    "zfp_stream_rewind": (None, [zs_p]),
    "zfp_stream_set_bit_stream": (None, [zs_p, bs_p]),
    "zfp_stream_set_reversible": (None, [zs_p]),
    "zfp_stream_set_rate": (ctypes.c_double, [zs_p, ctypes.c_double, ctypes.c_int, ctypes.c_uint, ctypes.c_int]),
    "zfp_stream_set_precision": (ctypes.c_uint, [zs_p, ctypes.c_uint]),
    "zfp_stream_set_accuracy": (ctypes.c_double, [zs_p, ctypes.c_double]),
    "zfp_stream_set_mode": (ctypes.c_int, [zs_p, ctypes.c_uint64]),
    "zfp_stream_set_params": (ctypes.c_int, [zs_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_int]),
    "zfp_stream_execution": (ctypes.c_int, [zs_p]),
    "zfp_stream_set_execution": (ctypes.c_int, [zs_p, ctypes.c_int]),
    "zfp_stream_set_omp_threads": (ctypes.c_int, [zs_p, ctypes.c_uint]),
    "zfp_stream_omp_threads": (ctypes.c_uint, [zs_p]),
    "zfp_field_alloc": (zf_p, []),
    "zfp_field_1d": (zf_p, [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]),
    "zfp_field_2d": (zf_p, [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t]),
    "zfp_field_3d": (zf_p, [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t]),
    "zfp_field_4d": (zf_p, [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t, ctypes.c_size_t, ctypes.c_size_t,
                            ctypes.c_size_t]),
    "zfp_field_free": (None, [zf_p]),
    "zfp_field_begin": (ctypes.c_void_p, [zf_p]),
    "zfp_field_size_bytes": (ctypes.c_size_t, [zf_p]),
    "zfp_field_blocks": (ctypes.c_size_t, [zf_p]),
    "zfp_field_dimensionality": (ctypes.c_uint, [zf_p]),
    "zfp_field_metadata": (ctypes.c_uint64, [zf_p]),
    "zfp_field_set_metadata": (ctypes.c_int, [zf_p, ctypes.c_uint64]),
    "zfp_field_set_pointer": (None, [zf_p, ctypes.c_void_p]),
    "zfp_field_set_type": (ctypes.c_int, [zf_p, ctypes.c_int]),
    "zfp_field_type": (ctypes.c_int, [zf_p]),
    "zfp_field_set_stride_1d": (None, [zf_p, ctypes.c_ssize_t]),
    "zfp_field_set_stride_2d": (None, [zf_p, ctypes.c_ssize_t, ctypes.c_ssize_t]),
    "zfp_field_set_stride_3d": (None, [zf_p, ctypes.c_ssize_t, ctypes.c_ssize_t, ctypes.c_ssize_t]),
    "zfp_field_set_stride_4d": (None, [zf_p, ctypes.c_ssize_t, ctypes.c_ssize_t, ctypes.c_ssize_t,
                                       ctypes.c_ssize_t]),
    "zfp_compress": (ctypes.c_size_t, [zs_p, zf_p]),
    "zfp_decompress": (ctypes.c_size_t, [zs_p, zf_p]),
    "zfp_write_header": (ctypes.c_size_t, [zs_p, zf_p, ctypes.c_uint]),
    "zfp_read_header": (ctypes.c_size_t, [zs_p, zf_p, ctypes.c_uint]),
    "zfp_stream_flush": (ctypes.c_size_t, [zs_p]),
    "zfp_stream_align": (ctypes.c_size_t, [zs_p]),
    "stream_open": (bs_p, [ctypes.c_void_p, ctypes.c_size_t]),
    "stream_close": (None, [bs_p]),
    "stream_size": (ctypes.c_size_t, [bs_p]),
    "stream_wtell": (ctypes.c_uint64, [bs_p]),
    "stream_rtell": (ctypes.c_uint64, [bs_p]),
    "stream_rewind": (None, [bs_p]),
    "stream_write_bits": (ctypes.c_uint64, [bs_p, ctypes.c_uint64, ctypes.c_size_t]),
    "stream_read_bits": (ctypes.c_uint64, [bs_p, ctypes.c_size_t]),
    "stream_write_bit": (ctypes.c_uint, [bs_p, ctypes.c_uint]),
    "stream_read_bit": (ctypes.c_uint, [bs_p]),
    "stream_pad": (None, [bs_p, ctypes.c_uint64]),
    "stream_skip": (None, [bs_p, ctypes.c_uint64]),
    "stream_flush": (ctypes.c_size_t, [bs_p]),
    "stream_align": (ctypes.c_size_t, [bs_p]),
    "stream_rseek": (None, [bs_p, ctypes.c_uint64]),
    "stream_wseek": (None, [bs_p, ctypes.c_uint64]),
    "stream_copy": (None, [bs_p, bs_p, ctypes.c_uint64]),
}

_SZ4 = ctypes.c_size_t * 4
_PD4 = ctypes.c_ssize_t * 4


class zfp_hip_job(ctypes.Structure):
    """One field of a batch (include/zfp_hip.h)."""
    _fields_ = [("type", ctypes.c_int), ("n", _SZ4), ("s", _PD4), ("data", ctypes.c_void_p),
                ("minbits", ctypes.c_uint), ("maxbits", ctypes.c_uint), ("maxprec", ctypes.c_uint),
                ("minexp", ctypes.c_int), ("stream", ctypes.c_void_p), ("offset", ctypes.c_uint64),
                ("nwords", ctypes.c_uint64), ("index", ctypes.c_void_p)]


ZFP_HIP_BATCH_COMPRESS, ZFP_HIP_BATCH_DECOMPRESS = 0, 1

_HIP = {
    "zfp_hip_compress": (ctypes.c_size_t, [zs_p, zf_p]),
    "zfp_hip_decompress": (ctypes.c_size_t, [zs_p, zf_p]),
    "zfp_hip_encode": (ctypes.c_int, [ctypes.c_int, _SZ4, _PD4, ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint,
                                      ctypes.c_uint, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64,
                                      _P(ctypes.c_uint64), ctypes.c_void_p]),
    "zfp_hip_encode_noalloc": (ctypes.c_int, [ctypes.c_int, _SZ4, _PD4, ctypes.c_void_p, ctypes.c_uint,
                                              ctypes.c_uint, ctypes.c_uint, ctypes.c_int, ctypes.c_void_p,
                                              ctypes.c_uint64, _P(ctypes.c_uint64), ctypes.c_void_p]),
    "zfp_hip_decode": (ctypes.c_int, [ctypes.c_int, _SZ4, _PD4, ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint,
                                      ctypes.c_uint, ctypes.c_int, ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                      _P(ctypes.c_uint64), ctypes.c_void_p]),
    "zfp_hip_encode_blocks": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint, ctypes.c_size_t, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_int,
                                             ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                             _P(ctypes.c_uint64), ctypes.c_void_p]),
    "zfp_hip_decode_blocks": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint, ctypes.c_size_t, ctypes.c_void_p,
                                             ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_int,
                                             ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint64,
                                             _P(ctypes.c_uint64), ctypes.c_void_p]),
    "zfp_hip_encode_blocks_host": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint, ctypes.c_size_t, ctypes.c_void_p,
                                                  _P(ctypes.c_uint64), ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                                  ctypes.c_int, _P(ctypes.c_uint64), ctypes.c_uint64,
                                                  ctypes.c_uint64]),
    "zfp_hip_decode_blocks_host": (ctypes.c_int, [ctypes.c_int, ctypes.c_uint, ctypes.c_size_t, ctypes.c_void_p,
                                                  _P(ctypes.c_uint64), ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                                  ctypes.c_int, _P(ctypes.c_uint64), ctypes.c_uint64,
                                                  ctypes.c_uint64]),
    "zfp_hip_index_size": (ctypes.c_size_t, [ctypes.c_void_p]),
    "zfp_hip_index_export_hybrid4": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                                    _P(ctypes.c_uint64), _P(ctypes.c_uint64)]),
    "zfp_hip_index_import_hybrid4": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_void_p,
                                                    ctypes.c_size_t, ctypes.c_uint64]),
    "zfp_hip_index_export_hybrid8": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, _P(ctypes.c_uint64), ctypes.c_size_t,
                                                    _P(ctypes.c_uint64), _P(ctypes.c_uint64)]),
    "zfp_hip_index_import_hybrid8": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint64,
                                                    _P(ctypes.c_uint64), ctypes.c_size_t, ctypes.c_uint64]),
    "zfp_hip_hybrid8_from_offsets": (ctypes.c_int, [ctypes.c_uint, _P(ctypes.c_uint64), ctypes.c_size_t,
                                                    _P(ctypes.c_uint64)]),
    "zfp_hip_offsets_from_hybrid8": (ctypes.c_int, [ctypes.c_uint, _P(ctypes.c_uint64), ctypes.c_size_t,
                                                    ctypes.c_uint64, _P(ctypes.c_uint64)]),
    "zfp_hip_hybrid4_from_offsets": (ctypes.c_int, [_P(ctypes.c_uint64), ctypes.c_size_t, ctypes.c_void_p]),
    "zfp_hip_offsets_from_hybrid4": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_uint64,
                                                    _P(ctypes.c_uint64)]),
    "zfp_hip_index_export": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_uint64), ctypes.c_size_t]),
    "zfp_hip_index_import": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_uint64), ctypes.c_size_t]),
    "zfp_hip_index_clear": (None, []),
    "zfp_hip_device_check": (ctypes.c_int, []),
    "zfp_hip_version": (ctypes.c_char_p, []),
    "zfp_hip_last_error": (ctypes.c_char_p, []),
    "zfp_hip_batch_create": (ctypes.c_void_p, [_P(zfp_hip_job), ctypes.c_size_t, ctypes.c_int, ctypes.c_void_p]),
    "zfp_hip_batch_run": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_uint64)]),
    "zfp_hip_batch_shape": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_uint), _P(ctypes.c_uint),
                                           _P(ctypes.c_uint)]),
    "zfp_hip_batch_set_shape": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint]),
    "zfp_hip_batch_destroy": (None, [ctypes.c_void_p]),
    "zfp_hip_server_create": (ctypes.c_void_p, [ctypes.c_uint64, ctypes.c_uint, ctypes.c_uint]),
    "zfp_hip_server_submit": (ctypes.c_int, [ctypes.c_void_p, _P(zfp_hip_job), ctypes.c_int, ctypes.c_uint64,
                                             ctypes.c_void_p, _P(ctypes.c_uint64)]),
    "zfp_hip_server_wait": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_uint64, _P(ctypes.c_uint64),
                                           _P(ctypes.c_double)]),
    "zfp_hip_server_stats": (ctypes.c_int, [ctypes.c_void_p, _P(ctypes.c_uint64), _P(ctypes.c_uint64),
                                            _P(ctypes.c_uint64), _P(ctypes.c_uint64)]),
    "zfp_hip_server_destroy": (None, [ctypes.c_void_p]),
}

_libs = {}


def _one_hip_runtime():
    """A process must hold a single HIP runtime.  PyTorch-ROCm ships its own
    libamdhip64 (SONAME libamdhip64.so.7, but linked by file name), so if our
    library loaded /opt/rocm's copy first, a later `import torch` would map a
    second runtime and one of the two would lose the device.  Importing
    torch first (when it is installed) makes libzfp_hip.so bind to torch's
    copy through the shared SONAME."""
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def _load(path, table):
    _one_hip_runtime()
    if not os.path.exists(path):
        raise RuntimeError(f"{os.path.basename(path)} is not built (run __graft_entry__.build() "
                           f"or `make -C zfp_amd`); the HIP codec has no fallback")
    lib = ctypes.CDLL(path)
    for name, (res, args) in table.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


def hip_lib():
    """libzfp_hip.so: the codec's C ABI (include/zfp_hip.h)."""
    if "hip" not in _libs:
        _libs["hip"] = _load(LIB_HIP, _HIP)
    return _libs["hip"]


def api_lib():
    """libzfp_amd.so: zfp's high-level C API with zfp_exec_hip (include/zfp.h)."""
    if "api" not in _libs:
        hip_lib()
        _libs["api"] = _load(LIB_API, _API)
    return _libs["api"]


def exported_symbols():
    """The C symbols each library must export (for load checks)."""
    return {"libzfp_hip.so": sorted(_HIP), "libzfp_amd.so": sorted(_API)}


def _args4(n, s):
    nl = list(n) + [0] * (4 - len(n))
    sl = list(s or []) + [0] * (4 - len(s or []))
    return _SZ4(*nl), _PD4(*sl)


def encode(ztype, n, s, d_data, params, d_words, offset=0, stream=None):
    """Device-level encode (zfp_hip_encode); returns bits written."""
    nn, ss = _args4(n, s)
    bits = ctypes.c_uint64()
    rc = hip_lib().zfp_hip_encode(ztype, nn, ss, ctypes.c_void_p(d_data), *params, ctypes.c_void_p(d_words),
                                  offset, ctypes.byref(bits), ctypes.c_void_p(stream or 0))
    if rc:
        raise RuntimeError(f"zfp_hip_encode failed with status {rc}")
    return bits.value


def decode(ztype, n, s, d_data, params, d_words, offset=0, stream=None, nwords=None):
    """Device-level decode (zfp_hip_decode); returns bits consumed.  nwords:
    readable words at d_words (None: the worst-case stream of the field
    after `offset`, i.e. a buffer sized like max_stream_words)."""
    nn, ss = _args4(n, s)
    if nwords is None:
        nwords = (offset + 63) // 64 + max_stream_words(ztype, n, params)
    bits = ctypes.c_uint64()
    rc = hip_lib().zfp_hip_decode(ztype, nn, ss, ctypes.c_void_p(d_data), *params, ctypes.c_void_p(d_words),
                                  nwords, offset, ctypes.byref(bits), ctypes.c_void_p(stream or 0))
    if rc:
        raise RuntimeError(f"zfp_hip_decode failed with status {rc}")
    return bits.value


def stream_params(mode, value, ztype, dims):
    """(minbits, maxbits, maxprec, minexp) that zfp_stream_set_rate /
    _set_precision / _set_accuracy give for `mode` in ("rate", "precision",
    "accuracy") -- computed by the library's own zfp.h API."""
    api = api_lib()
    zs = api.zfp_stream_open(None)
    try:
        if mode == "rate":
            api.zfp_stream_set_rate(zs, float(value), ztype, dims, 0)
        elif mode == "precision":
            api.zfp_stream_set_precision(zs, int(value))
        elif mode == "accuracy":
            api.zfp_stream_set_accuracy(zs, float(value))
        else:
            raise ValueError(f"unknown mode {mode!r}")
        z = zs.contents
        return (z.minbits, z.maxbits, z.maxprec, z.minexp)
    finally:
        api.zfp_stream_close(zs)


def max_stream_words(ztype, n, params):
    """Stream capacity in 64-bit words for a field of extents n
    (zfp_stream_maximum_size of the library's zfp.h API)."""
    api = api_lib()
    zs = api.zfp_stream_open(None)
    nn = list(n) + [0] * (4 - len(n))
    f = api.zfp_field_alloc()
    try:
        fl = f.contents
        fl.type = ztype
        fl.nx, fl.ny, fl.nz, fl.nw = nn
        api.zfp_stream_set_params(zs, *params)
        return (api.zfp_stream_maximum_size(zs, f) + 7) // 8
    finally:
        api.zfp_field_free(f)
        api.zfp_stream_close(zs)


def make_job(ztype, n, s, d_data, params, d_words, offset=0, nwords=0, d_index=None):
    """A zfp_hip_job for a device field and stream (params: minbits, maxbits,
    maxprec, minexp)."""
    nn, ss = _args4(n, s)
    j = zfp_hip_job()
    j.type = ztype
    j.n = nn
    j.s = ss
    j.data = d_data
    j.minbits, j.maxbits, j.maxprec, j.minexp = params
    j.stream = d_words
    j.offset = offset
    j.nwords = nwords
    j.index = d_index or None
    return j


class Batch:
    """A zfp_hip_batch: create once, run many times (graph replays)."""

    def __init__(self, jobs, direction=ZFP_HIP_BATCH_COMPRESS, stream=None):
        lib = hip_lib()
        self._n = len(jobs)
        self._jobs = (zfp_hip_job * self._n)(*jobs)
        self._h = lib.zfp_hip_batch_create(self._jobs, self._n, direction, ctypes.c_void_p(stream or 0))
        if not self._h:
            raise RuntimeError("zfp_hip_batch_create failed: " + lib.zfp_hip_last_error().decode())

    def run(self):
        """Runs the batch; returns the bits of every job."""
        bits = (ctypes.c_uint64 * self._n)()
        rc = hip_lib().zfp_hip_batch_run(self._h, bits)
        if rc:
            raise RuntimeError(f"zfp_hip_batch_run failed with status {rc}: "
                               + hip_lib().zfp_hip_last_error().decode())
        return list(bits)

    def shape(self):
        a, b, c = ctypes.c_uint(), ctypes.c_uint(), ctypes.c_uint()
        hip_lib().zfp_hip_batch_shape(self._h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(c))
        return a.value, b.value, c.value

    def set_shape(self, wg, waves_per_cu, blocks_per_task):
        rc = hip_lib().zfp_hip_batch_set_shape(self._h, wg, waves_per_cu, blocks_per_task)
        if rc:
            raise RuntimeError("zfp_hip_batch_set_shape failed: " + hip_lib().zfp_hip_last_error().decode())

    def close(self):
        if self._h:
            hip_lib().zfp_hip_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """A zfp_hip_server: submit jobs as they arrive, wait for their results."""

    def __init__(self, hbm_budget=64 << 30, max_batch_jobs=0, slots=0):
        self._h = hip_lib().zfp_hip_server_create(hbm_budget, max_batch_jobs, slots)
        if not self._h:
            raise RuntimeError("zfp_hip_server_create failed: " + hip_lib().zfp_hip_last_error().decode())

    def submit(self, job, direction, nbytes, stream=None):
        rid = ctypes.c_uint64()
        rc = hip_lib().zfp_hip_server_submit(self._h, ctypes.byref(job), direction, nbytes,
                                             ctypes.c_void_p(stream or 0), ctypes.byref(rid))
        if rc:
            raise RuntimeError(f"zfp_hip_server_submit failed with status {rc}: "
                               + hip_lib().zfp_hip_last_error().decode())
        return rid.value

    def wait(self, rid):
        """(bits, latency in ms) of request rid."""
        bits, lat = ctypes.c_uint64(), ctypes.c_double()
        rc = hip_lib().zfp_hip_server_wait(self._h, rid, ctypes.byref(bits), ctypes.byref(lat))
        if rc:
            raise RuntimeError(f"request {rid} failed with status {rc}")
        return bits.value, lat.value

    def stats(self):
        v = [ctypes.c_uint64() for _ in range(4)]
        hip_lib().zfp_hip_server_stats(self._h, *[ctypes.byref(x) for x in v])
        return dict(zip(("done", "batches", "max_batch", "max_inflight"), (x.value for x in v)))

    def close(self):
        if self._h:
            hip_lib().zfp_hip_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
