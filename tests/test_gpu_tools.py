"""The callers of the codec on the GPU.

* The `zfp` command-line tool (zfp_amd/zfp) under `-x hip` against the
  reference's own tool (utils/zfp.c built as oracle/_ref/zfp_ref) under
  `-x serial`: identical compressed files (with and without the full header),
  identical reconstructions, identical statistics lines, for every type,
  1D-4D and every mode; stdin/stdout streaming.
* The zfpy interface (zfp_amd.zfpy) against the oracle: compress_numpy
  bytes equal the serial header + stream, decompress_numpy and _decompress
  equal serial decompression, reversible round trips are lossless (the
  reference's tests/python/test_numpy.py), strided and reversed views.
"""
import os
import subprocess

import numpy as np
import pytest

from zfp_amd import zfpy
from zfptest import (ROOT, T_DOUBLE, T_FLOAT, T_INT32, T_INT64, header_prefix, oracle_decompress, oracle_stream,
                     smooth_field, stream_params_accuracy, stream_params_precision, stream_params_rate)

pytestmark = pytest.mark.gpu

CLI = os.path.join(ROOT, "zfp_amd", "zfp")
REF_CLI = os.path.join(ROOT, "oracle", "_ref", "zfp_ref")
needs_ref_cli = pytest.mark.skipif(not os.path.exists(REF_CLI), reason="oracle/_ref/zfp_ref not built")

TFLAG = {T_INT32: ["-t", "i32"], T_INT64: ["-t", "i64"], T_FLOAT: ["-f"], T_DOUBLE: ["-d"]}
SHAPES = {1: [301], 2: [37, 19], 3: [21, 13, 9], 4: [7, 6, 5, 9]}
MODES = {"rate": ["-r", "12.5"], "precision": ["-p", "20"], "accuracy": ["-a", "1e-4"], "reversible": ["-R"],
         "expert": ["-c", "100", "800", "30", "-40"]}


def run(exe, args, cwd, stdin=None):
    print("run", os.path.basename(exe), " ".join(args), flush=True)
    p = subprocess.run([exe] + args, cwd=cwd, capture_output=True, input=stdin, timeout=120)
    return p.returncode, p.stderr.decode(), p.stdout


# every mode on 3D doubles and 2D floats; every type and dimensionality
CLI_CASES = ([(3, T_DOUBLE, m) for m in MODES] + [(2, T_FLOAT, m) for m in MODES]
             + [(1, T_INT32, "rate"), (4, T_INT32, "reversible"), (2, T_INT64, "precision"),
                (3, T_INT64, "expert"), (1, T_DOUBLE, "accuracy"), (4, T_FLOAT, "rate")])


@needs_ref_cli
@pytest.mark.parametrize("dims,ztype,mode", CLI_CASES)
def test_cli_matches_reference_tool(dims, ztype, mode, tmp_path):
    n = SHAPES[dims]
    a = smooth_field(n, ztype, seed=dims * 31 + ztype)
    a.tofile(tmp_path / "in.raw")
    meta = TFLAG[ztype] + [f"-{dims}"] + [str(v) for v in n] + MODES[mode]
    for header in ([], ["-h"]):
        rc, err, _ = run(CLI, ["-i", "in.raw", "-z", "ours.zfp", "-o", "ours.raw", "-s"] + meta + header, tmp_path)
        rrc, rerr, _ = run(REF_CLI, ["-i", "in.raw", "-z", "ref.zfp", "-o", "ref.raw", "-s", "-x", "serial"]
                           + meta + header, tmp_path)
        assert rc == rrc == 0, err
        assert (tmp_path / "ours.zfp").read_bytes() == (tmp_path / "ref.zfp").read_bytes()
        assert (tmp_path / "ours.raw").read_bytes() == (tmp_path / "ref.raw").read_bytes()
        # type, extents, sizes, ratio, rate and the error statistics
        assert err == rerr
        # decompression of the compressed file alone
        dmeta = [] if header else meta
        rc, err, _ = run(CLI, ["-z", "ours.zfp", "-o", "back.raw"] + dmeta + header, tmp_path)
        assert rc == 0, err
        assert (tmp_path / "back.raw").read_bytes() == (tmp_path / "ref.raw").read_bytes()


@needs_ref_cli
def test_cli_streams_through_stdin_and_stdout(tmp_path):
    n = [40, 30, 20]
    a = smooth_field(n, T_FLOAT, seed=3)
    meta = ["-f", "-3"] + [str(v) for v in n] + ["-r", "8", "-h"]
    rc, err, zbytes = run(CLI, ["-i", "-", "-z", "-", "-q"] + meta, tmp_path, stdin=a.tobytes())
    assert rc == 0 and err == ""
    rrc, _, rbytes = run(REF_CLI, ["-i", "-", "-z", "-", "-q", "-x", "serial"] + meta, tmp_path, stdin=a.tobytes())
    assert rrc == 0 and zbytes == rbytes
    rc, err, raw = run(CLI, ["-z", "-", "-o", "-", "-h", "-q"], tmp_path, stdin=zbytes)
    rrc, _, rraw = run(REF_CLI, ["-z", "-", "-o", "-", "-h", "-q"], tmp_path, stdin=rbytes)
    assert rc == rrc == 0 and raw == rraw and len(raw) == a.nbytes


# ---------------------------------------------------------------------------
# zfpy
# ---------------------------------------------------------------------------

def _zfpy_params(ztype, ndim, tolerance=-1, rate=-1, precision=-1):
    """The parameters compress_numpy sets (zfpy.pyx:205-220; the rate setter
    is passed zfp_type_none)."""
    if tolerance >= 0:
        return stream_params_accuracy(tolerance)
    if rate >= 0:
        return stream_params_rate(rate, 0, ndim)
    if precision >= 0:
        return stream_params_precision(precision)
    return (1, 16658, 64, -1075)


def _serial(arr, params, header=7):
    """Bytes of serial zfp on the array (as zfpy lays it out) and the
    serial reconstruction."""
    ztype = zfpy.dtype_to_ztype(arr.dtype)
    n = list(arr.shape[::-1])
    flat = np.ascontiguousarray(arr).reshape(-1)
    words, size = oracle_stream(flat, 0, n, None, ztype, params, header)
    ref = np.zeros_like(flat)
    hbits = header_prefix(ztype, n, params)[1] if header else 0
    oracle_decompress(words, ref, 0, n, None, ztype, params, offset=hbits)
    return words.view(np.uint8)[:size].tobytes(), ref.reshape(arr.shape)


KW = [dict(rate=8), dict(rate=21.5), dict(precision=16), dict(precision=30), dict(tolerance=1e-3),
      dict(tolerance=1e-8), dict()]


@pytest.mark.parametrize("kw", range(len(KW)))
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int32, np.int64])
@pytest.mark.parametrize("ndim", [1, 2, 3, 4])
def test_zfpy_matches_serial(ndim, dtype, kw):
    shape = [[157], [19, 23], [9, 11, 13], [5, 6, 7, 9]][ndim - 1]
    ztype = zfpy.dtype_to_ztype(np.dtype(dtype))
    arr = smooth_field(shape[::-1], ztype, seed=ndim * 7 + ztype).reshape(shape)
    kwargs = KW[kw]
    params = _zfpy_params(ztype, ndim, **kwargs)
    buf = zfpy.compress_numpy(arr, **kwargs)
    exp, ref = _serial(arr, params)
    assert buf == exp
    out = zfpy.decompress_numpy(buf)
    assert out.shape == arr.shape and out.dtype == arr.dtype
    assert np.array_equal(out.view(np.uint8), ref.view(np.uint8))
    if not kwargs:
        assert np.array_equal(out, arr)
    h = zfpy.header(buf)
    assert h["type"] == dtype and h["mode"] == {"rate": "rate", "precision": "precision",
                                                 "tolerance": "tolerance"}.get(next(iter(kwargs), ""), "reversible")


def test_zfpy_lossless_round_trips():
    """tests/python/test_numpy.py: test_different_dimensions and
    test_different_dtypes, seeded."""
    rng = np.random.default_rng(11)
    for dims in range(1, 5):
        for shape in ([5] * dims, list(range(2, 2 + dims))):
            a = rng.random(shape)
            assert np.array_equal(zfpy.decompress_numpy(zfpy.compress_numpy(a, write_header=True)), a)
    for dtype in (np.float32, np.float64):
        a = rng.random((5, 5)).astype(dtype)
        assert np.array_equal(zfpy.decompress_numpy(zfpy.compress_numpy(a)), a)
    for dtype in (np.int32, np.int64):
        a = rng.integers(0, 2 ** 30, size=(5, 5), dtype=dtype)
        assert np.array_equal(zfpy.decompress_numpy(zfpy.compress_numpy(a)), a)


def test_zfpy_headerless_decompress():
    """The advanced interface (zfpy._decompress) with the metadata given by
    the caller, into a new array and into `out`."""
    arr = smooth_field([33, 17], T_FLOAT, seed=5).reshape(17, 33)
    for kwargs in (dict(tolerance=1e-3), dict(rate=10), dict(precision=18)):
        buf = zfpy.compress_numpy(arr, write_header=False, **kwargs)
        exp, ref = _serial(arr, _zfpy_params(T_FLOAT, 2, **kwargs), header=0)
        assert buf == exp
        # decompression passes the real type to the rate setter
        # (zfpy.pyx:291); rates here are above the float minimum, so the
        # parameters agree
        got = zfpy._decompress(buf, zfpy.type_float, arr.shape, **kwargs)
        assert np.array_equal(got.view(np.uint8), ref.view(np.uint8))
        out = np.empty_like(arr)
        zfpy._decompress(buf, zfpy.type_float, arr.shape, out=out, **kwargs)
        assert np.array_equal(out.view(np.uint8), ref.view(np.uint8))
        with pytest.raises(ValueError):
            zfpy._decompress(buf, zfpy.type_float, arr.shape, out=np.empty(arr.shape, np.float64), **kwargs)


def test_zfpy_strided_views():
    """Non-contiguous arrays are compressed with their strides, as zfpy
    passes them (zfpy.pyx:93-120): the stream equals serial zfp on the same
    view."""
    base = smooth_field([24, 20, 16], T_DOUBLE, seed=9).reshape(16, 20, 24)
    for view in (base[:, ::-1, :], base[::2, :, 1::3], base.transpose(2, 0, 1)):
        buf = zfpy.compress_numpy(view, rate=16)
        exp, ref = _serial(np.ascontiguousarray(view), _zfpy_params(T_DOUBLE, 3, rate=16))
        assert buf == exp
        assert np.array_equal(zfpy.decompress_numpy(buf), ref)
