"""The callers of the codec without a GPU: the `zfp` command-line tool
(zfp_amd/zfp, the reference's utils/zfp.c with `-x hip`) and the zfpy
interface (zfp_amd.zfpy, the reference's python/zfpy.pyx with an
`execution` keyword).

Argument checking, messages and exit codes are compared with the
reference's own tool built from utils/zfp.c (oracle/_ref/zfp_ref), and
header parsing with the reference library.  Compression itself needs the
GPU (tests/test_gpu_tools.py); here it must decline: the tool reports
"compression failed" and exits 1, zfpy raises RuntimeError."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

import zfp_amd as Z
from zfp_amd import zfpy
from zfptest import ROOT, T_DOUBLE, T_FLOAT, T_INT32, T_INT64, have_ref, header_prefix, ref_api

CLI = os.path.join(ROOT, "zfp_amd", "zfp")
REF_CLI = os.path.join(ROOT, "oracle", "_ref", "zfp_ref")
needs_cli = pytest.mark.skipif(not os.path.exists(CLI), reason="zfp_amd/zfp not built")
needs_ref_cli = pytest.mark.skipif(not os.path.exists(REF_CLI), reason="oracle/_ref/zfp_ref not built")


def _gpu_usable():
    return Z.hip_lib().zfp_hip_device_check() == 0


def run(exe, args, cwd, stdin=None):
    p = subprocess.run([exe] + args, cwd=cwd, capture_output=True, input=stdin, timeout=60)
    return p.returncode, p.stderr.decode(), p.stdout


# argument errors (in.raw and in.zfp exist; missing.raw does not)
CASES = [
    [],
    ["-y"],
    ["-r"],
    ["-3", "4", "4"],
    ["-x", "bogus", "-i", "in.raw"],
    ["-t", "f16"],
    ["-d", "-3", "4", "4", "4", "-r", "8"],                          # no input
    ["-i", "in.raw", "-3", "4", "4", "4", "-r", "8"],                # no type
    ["-i", "in.raw", "-d", "-r", "8"],                               # no dims
    ["-i", "in.raw", "-d", "-3", "4", "4", "4"],                     # no mode
    ["-z", "in.zfp", "-o", "out.raw"],                               # no type, no header
    ["-z", "in.zfp", "-d", "-o", "out.raw"],                         # no dims
    ["-z", "in.zfp", "-d", "-3", "4", "4", "4", "-o", "out.raw"],    # no mode
    ["-z", "in.zfp", "-h", "-d", "-o", "out.raw"],                   # type and header
    ["-z", "in.zfp", "-h", "-s"],                                    # stats without input
    ["-i", "in.raw", "-d", "-3", "0", "4", "4", "-r", "8"],          # zero extent
    ["-i", "missing.raw", "-d", "-3", "4", "4", "4", "-r", "8"],     # unreadable
    ["-i", "in.raw", "-d", "-3", "8", "8", "8", "-r", "8"],          # short file
    ["-i", "in.raw", "-d", "-3", "4", "4", "4", "-c", "200", "100", "64", "-1074"],  # minbits > maxbits
]


@needs_cli
@needs_ref_cli
@pytest.mark.parametrize("k", range(len(CASES)))
def test_cli_argument_errors_match_reference(k, tmp_path):
    np.zeros(64, dtype=np.float64).tofile(tmp_path / "in.raw")
    (tmp_path / "in.zfp").write_bytes(b"\0" * 64)
    args = CASES[k]
    rc, err, _ = run(CLI, args, tmp_path)
    rrc, rerr, _ = run(REF_CLI, args, tmp_path)
    assert rc == rrc == 1
    if "Usage: zfp <options>" in rerr:
        assert "Usage: zfp <options>" in err
    else:
        assert err == rerr


@needs_cli
def test_cli_usage_names_hip():
    rc, err, _ = run(CLI, [], ROOT)
    assert rc == 1 and "-x hip" in err


@needs_cli
@needs_ref_cli
def test_cli_unavailable_policies(tmp_path):
    np.zeros(64, dtype=np.float64).tofile(tmp_path / "in.raw")
    base = ["-i", "in.raw", "-d", "-3", "4", "4", "4", "-r", "8", "-z", "o.zfp"]
    rc, err, _ = run(CLI, base + ["-x", "cuda"], tmp_path)
    # the reference built without CUDA says the same
    rrc, rerr, _ = run(REF_CLI, base + ["-x", "cuda"], tmp_path)
    assert (rc, err) == (rrc, rerr) == (1, "cuda execution not available\n")
    # no CPU codec in this library: serial and OpenMP compression fail
    for x in ("serial", "omp", "omp=4", "omp=4,16"):
        rc, err, _ = run(CLI, base + ["-x", x], tmp_path)
        assert (rc, err) == (1, "compression failed\n")


@needs_cli
def test_cli_without_gpu_declines(tmp_path):
    if _gpu_usable():
        pytest.skip("a gfx950 device is present")
    np.zeros(64, dtype=np.float64).tofile(tmp_path / "in.raw")
    rc, err, _ = run(CLI, ["-i", "in.raw", "-d", "-3", "4", "4", "4", "-r", "8", "-z", "o.zfp"], tmp_path)
    assert (rc, err) == (1, "compression failed\n")
    assert not (tmp_path / "o.zfp").exists()


# ---------------------------------------------------------------------------
# zfpy
# ---------------------------------------------------------------------------

def test_zfpy_constants():
    assert (zfpy.HEADER_MAGIC, zfpy.HEADER_META, zfpy.HEADER_MODE, zfpy.HEADER_FULL) == (1, 2, 4, 7)
    assert (zfpy.type_int32, zfpy.type_int64, zfpy.type_float, zfpy.type_double) == (1, 2, 3, 4)
    assert (zfpy.mode_null, zfpy.mode_expert, zfpy.mode_fixed_rate, zfpy.mode_fixed_precision,
            zfpy.mode_fixed_accuracy) == (0, 1, 2, 3, 4)
    for dt, zt in ((np.int32, 1), (np.int64, 2), (np.float32, 3), (np.float64, 4)):
        assert zfpy.dtype_to_ztype(np.dtype(dt)) == zt
        assert zfpy.ztype_to_dtype(zt) == dt
    with pytest.raises(TypeError):
        zfpy.dtype_to_ztype(np.dtype(np.float16))
    with pytest.raises(ValueError):
        zfpy.ztype_to_dtype(0)


def test_zfpy_argument_errors():
    a = np.zeros((4, 4))
    with pytest.raises(TypeError):
        zfpy.compress_numpy(None)
    with pytest.raises(ValueError):
        zfpy.compress_numpy(a, tolerance=1e-3, rate=8)
    with pytest.raises(ValueError):
        zfpy.compress_numpy(a, rate=8, execution="serial")
    with pytest.raises(TypeError):
        zfpy.decompress_numpy(None)
    with pytest.raises(TypeError):
        zfpy._decompress(None, zfpy.type_double, [4, 4])
    with pytest.raises(ValueError):
        zfpy._decompress(b"\0" * 64, zfpy.type_double, [2, 2, 2, 2, 2])
    with pytest.raises(ValueError):
        zfpy._decompress(b"\0" * 64, zfpy.type_double, [])
    with pytest.raises(TypeError):
        zfpy._decompress(b"\0" * 64, zfpy.type_double, 5)
    # a stream without the magic word
    with pytest.raises(ValueError, match="header"):
        zfpy.decompress_numpy(b"\0" * 64)
    with pytest.raises(ValueError, match="header"):
        zfpy.header(b"\0" * 64)


HEADERS = [
    (T_DOUBLE, [10, 20, 30], (1024, 1024, 64, -1074), "rate"),
    (T_FLOAT, [100, 7], (1, 16658, 20, -1074), "precision"),
    (T_DOUBLE, [5, 6, 7, 8], (1, 16658, 64, -20), "tolerance"),
    (T_INT32, [1000], (1, 16658, 64, -1075), "reversible"),
    (T_INT64, [9, 9], (100, 600, 20, -30), "expert"),
]


@pytest.mark.parametrize("k", range(len(HEADERS)))
def test_zfpy_header_reads_what_the_reference_writes(k):
    ztype, n, params, mode = HEADERS[k]
    words, bits = header_prefix(ztype, n, params)
    buf = words.tobytes()
    h = zfpy.header(buf)
    ext = list(n) + [0] * (4 - len(n))
    assert (h["nx"], h["ny"], h["nz"], h["nw"]) == tuple(ext)
    assert h["type"] == zfpy.ztype_to_dtype(ztype)
    assert h["mode"] == mode
    assert h["config"]["expert"] == {"minbits": params[0], "maxbits": params[0], "maxprec": params[2],
                                     "minexp": params[3]}
    if have_ref():
        # the same bytes through the reference library's reader
        api = ref_api()
        store = np.zeros(8, dtype=np.uint64)
        store[:words.size] = words
        bs = api.stream_open(ctypes.c_void_p(store.ctypes.data), store.nbytes)
        z = api.zfp_stream_open(bs)
        f = api.zfp_field_alloc()
        assert api.zfp_read_header(z, f, 7) == bits
        assert api.zfp_stream_compression_mode(z) == h["config"]["mode"]
        assert api.zfp_stream_rate(z, len(n)) == h["config"]["rate"]
        assert api.zfp_stream_precision(z) == h["config"]["precision"]
        assert api.zfp_stream_accuracy(z) == h["config"]["tolerance"]
        api.zfp_field_free(f)
        api.zfp_stream_close(z)
        api.stream_close(bs)


def test_zfpy_without_gpu_declines():
    if _gpu_usable():
        pytest.skip("a gfx950 device is present")
    with pytest.raises(RuntimeError, match="Failed to write to stream"):
        zfpy.compress_numpy(np.zeros((8, 8)), rate=8)
