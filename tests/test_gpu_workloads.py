"""Parity at the benchmark's own workloads, at their full sizes: every stream
the GPU writes equals serial zfp's (the oracle's) word for word, and every
array it decompresses equals serial zfp's decompression bit for bit -- the
contract of the reference's end-to-end tests
(tests/src/endtoend/zfpEndtoendBase.c:386-464), applied to

* 3D float64 512^3 (1 GiB) at fixed rates 4, 16 and 32 bits/value, with 1%
  of the values 0, +-1e-300 or +-1e300, read from the contiguous layout and
  from the permuted view (z fastest in memory, x slowest with a negative
  stride);
* 2D float32 8191 x 8193 and 4D float64 96 x 96 x 96 x 24 at precision 16
  and 28 and accuracy 1e-3 and 1e-7;
* the batch of 1024 mixed arrays (the batched codec, one launch per pass);
* the serving trace of 200 requests on 32 resident 3D fields of 32^3 to
  384^3, with Poisson arrivals.

The fields are the benchmark's own generators (bench.py), made on the GPU and
copied to the host for the oracle.  The oracle runs in parallel threads over
slabs of whole block rows, concatenated bit for bit (the single-call stream:
pinned on the CPU by tests/test_oracle_pins.py).
"""
import ctypes
import math

import numpy as np
import pytest

import bench
import zfp_amd as Z
from zfptest import (oracle_compress, oracle_compress_parallel, oracle_decompress, oracle_decompress_parallel,
                     oracle_threads)

pytestmark = pytest.mark.gpu

NP_OF = {1: np.int32, 2: np.int64, 3: np.float32, 4: np.float64}


def _torch():
    import torch
    return torch


def _host_words(t, bits):
    return t[: (bits + 63) // 64].cpu().numpy().view(np.uint64)


# --------------------------------------------------------------------------
# 3D float64 512^3, fixed rate, contiguous and permuted view
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def field512():
    torch = _torch()
    n = 512
    f = bench.make_field(torch, n)
    fp = torch.flip(f.permute(2, 1, 0), dims=[0]).contiguous()
    host = f.cpu().numpy().reshape(-1)
    torch.cuda.synchronize()
    yield n, f, fp, host
    del f, fp
    torch.cuda.empty_cache()


@pytest.mark.parametrize("rate", [4, 16, 32])
def test_512cube_fixed_rate(field512, rate):
    torch = _torch()
    n, f, fp, host = field512
    N = n ** 3
    shape = [n, n, n]
    p = Z.stream_params("rate", rate, Z.zfp_type_double, 3)
    assert p == (64 * rate, 64 * rate, 64, -1074)
    print(f"512^3 f64 rate {rate}: oracle", flush=True)
    ow, obits, starts, slabs = oracle_compress_parallel(host, 0, shape, None, 4, p)
    assert obits == N * rate
    nw = obits // 64
    w = torch.zeros(nw + 8, dtype=torch.int64, device="cuda:0")
    print(f"512^3 f64 rate {rate}: GPU contiguous", flush=True)
    assert Z.encode(4, shape, None, f.data_ptr(), p, w.data_ptr()) == obits
    assert np.array_equal(_host_words(w, obits), ow[:nw]), "contiguous stream differs from serial zfp"
    # the permuted negative-stride view of the same field writes the same stream
    w.zero_()
    pv = (fp.data_ptr() + (n - 1) * n * n * 8, [-n * n, n, 1])
    print(f"512^3 f64 rate {rate}: GPU permuted", flush=True)
    assert Z.encode(4, shape, pv[1], pv[0], p, w.data_ptr()) == obits
    assert np.array_equal(_host_words(w, obits), ow[:nw]), "permuted-view stream differs from serial zfp"
    # decompression of serial zfp's stream, both layouts
    ref = np.zeros_like(host)
    assert oracle_decompress_parallel(ow, ref, 0, shape, None, 4, p, starts, slabs) == obits
    out = torch.empty_like(f)
    assert Z.decode(4, shape, None, out.data_ptr(), p, w.data_ptr(), nwords=nw) == obits
    assert np.array_equal(out.cpu().numpy().reshape(-1).view(np.uint64), ref.view(np.uint64)), \
        "contiguous decompression differs from serial zfp"
    outp = torch.empty_like(fp)
    assert Z.decode(4, shape, pv[1], outp.data_ptr() + (n - 1) * n * n * 8, p, w.data_ptr(), nwords=nw) == obits
    back = torch.flip(outp, dims=[0]).permute(2, 1, 0).contiguous()
    assert torch.equal(back.view(torch.int64), out.view(torch.int64)), "permuted-view decompression differs"


# --------------------------------------------------------------------------
# variable-rate modes: 2D float32 8191 x 8193, 4D float64 96^3 x 24
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def vr_fields():
    torch = _torch()
    fields = bench.vr_fields(torch)
    yield {k: (zt, n, f, f.cpu().numpy().reshape(-1)) for k, (zt, n, f) in fields.items()}
    del fields
    torch.cuda.empty_cache()


VR_MODES = [("precision", 16), ("precision", 28), ("accuracy", 1e-3), ("accuracy", 1e-7)]


@pytest.mark.parametrize("name", ["2d_f32_8191x8193", "4d_f64_96x96x96x24"])
@pytest.mark.parametrize("mode", VR_MODES, ids=[f"{m}{v}" for m, v in VR_MODES])
def test_variable_rate_full_size(vr_fields, name, mode):
    torch = _torch()
    ztype, n, f, host = vr_fields[name]
    p = Z.stream_params(mode[0], mode[1], ztype, len(n))
    print(f"{name} {mode}: oracle", flush=True)
    ow, obits, starts, slabs = oracle_compress_parallel(host, 0, n, None, ztype, p)
    nw = (obits + 63) // 64
    w = torch.zeros(Z.max_stream_words(ztype, n, p), dtype=torch.int64, device="cuda:0")
    print(f"{name} {mode}: GPU", flush=True)
    assert Z.encode(ztype, n, None, f.data_ptr(), p, w.data_ptr()) == obits
    assert np.array_equal(_host_words(w, obits), ow[:nw]), "stream differs from serial zfp"
    ref = np.zeros_like(host)
    assert oracle_decompress_parallel(ow, ref, 0, n, None, ztype, p, starts, slabs) == obits
    out = torch.zeros_like(f)
    assert Z.decode(ztype, n, None, out.data_ptr(), p, w.data_ptr(), nwords=nw) == obits
    width = np.uint32 if ztype == 3 else np.uint64
    assert np.array_equal(out.cpu().numpy().reshape(-1).view(width), ref.view(width)), \
        "decompression differs from serial zfp"


# --------------------------------------------------------------------------
# the batch of 1024 mixed arrays
# --------------------------------------------------------------------------

MODES = {"rate8": ("rate", 8), "rate16": ("rate", 16), "prec20": ("precision", 20), "acc1e-4": ("accuracy", 1e-4)}


def _oracle_job(host, n, ztype, p):
    words, bits = oracle_compress(host, 0, n, None, ztype, p)
    rec = np.zeros_like(host)
    oracle_decompress(words, rec, 0, n, None, ztype, p)
    return words, bits, rec


def test_batch_1024_mixed_arrays():
    from concurrent.futures import ThreadPoolExecutor
    torch = _torch()
    fields = bench.batch_fields(torch, 1024, 20251203)
    assert len(fields) == 1024
    jobs_c, jobs_d, keep, expect_in = [], [], [], []
    for ztype, n, mode, t in fields:
        p = Z.stream_params(*MODES[mode], ztype, len(n))
        nw = Z.max_stream_words(ztype, n, p)
        nb = math.prod((v + 3) // 4 for v in n)
        w = torch.zeros(nw, dtype=torch.int64, device="cuda:0")
        var = p[0] != p[1]
        idx = torch.zeros(nb + 1, dtype=torch.int64, device="cuda:0") if var else None
        o = torch.zeros_like(t)
        jobs_c.append(Z.make_job(ztype, n, None, t.data_ptr(), p, w.data_ptr(), 0, nw,
                                 idx.data_ptr() if var else None))
        jobs_d.append(Z.make_job(ztype, n, None, o.data_ptr(), p, w.data_ptr(), 0, nw,
                                 idx.data_ptr() if var else None))
        keep.append((ztype, n, p, w, o, t, idx))   # idx too: the batches hold its address
        expect_in.append((t.cpu().numpy().reshape(-1), n, ztype, p))
    print("batch: oracle", flush=True)
    with ThreadPoolExecutor(oracle_threads()) as ex:
        expect = list(ex.map(lambda a: _oracle_job(*a), expect_in))
    print("batch: GPU", flush=True)
    bc = Z.Batch(jobs_c, Z.ZFP_HIP_BATCH_COMPRESS)
    bits = bc.run()
    bd = Z.Batch(jobs_d, Z.ZFP_HIP_BATCH_DECOMPRESS)
    dbits = bd.run()
    bc.close()
    bd.close()
    for k, ((ztype, n, p, w, o, t, _), (ew, eb, rec)) in enumerate(zip(keep, expect)):
        assert bits[k] == eb and dbits[k] == eb, f"job {k}: {n} type {ztype} params {p}"
        assert np.array_equal(_host_words(w, eb), ew[: (eb + 63) // 64]), f"job {k}: stream differs"
        width = np.uint32 if ztype in (1, 3) else np.uint64
        assert np.array_equal(o.cpu().numpy().reshape(-1).view(width), rec.view(width)), \
            f"job {k}: decompression differs"


# --------------------------------------------------------------------------
# the serving trace: 200 requests, Poisson arrivals
# --------------------------------------------------------------------------

def test_serving_trace_200_requests():
    import time
    from concurrent.futures import ThreadPoolExecutor
    torch = _torch()
    torch.manual_seed(20251204)
    rng = np.random.default_rng(20251204)
    pool = bench.serve_pool(torch, 32, rng)
    picks = bench.serve_picks(rng, 200, len(pool))
    print("serving: oracle", flush=True)
    hosts = [P["f"].cpu().numpy().reshape(-1) for P in pool]

    def one(k):
        P = pool[k]
        ow, obits, starts, slabs = oracle_compress_parallel(hosts[k], 0, P["n"], None, P["ztype"], P["p"], threads=4)
        rec = np.zeros_like(hosts[k])
        oracle_decompress_parallel(ow, rec, 0, P["n"], None, P["ztype"], P["p"], starts, slabs, threads=4)
        return ow, obits, rec

    with ThreadPoolExecutor(max(1, oracle_threads() // 4)) as ex:
        expect = list(ex.map(one, range(len(pool))))
    # the decompression requests read serial zfp's stream and its index
    for k, (P, (ow, obits, _)) in enumerate(zip(pool, expect)):
        w_in = np.zeros(P["nw"], dtype=np.uint64)   # the stream in a buffer of the job's nwords words
        k_w = min(P["nw"], ow.size)
        w_in[:k_w] = ow[:k_w]
        P["w_in"] = torch.from_numpy(w_in.view(np.int64)).to("cuda:0")
        if P["var"]:
            _, _, bb = oracle_compress(hosts[k], 0, P["n"], None, P["ztype"], P["p"], block_bits=True)
            P["offs"] = np.concatenate([[0], np.cumsum(bb.astype(np.uint64))]).astype(np.uint64)
            P["idx_in"] = torch.from_numpy(P["offs"].view(np.int64)).to("cuda:0")
    outs, jobs = [], []
    for fi, kind in picks:
        P = pool[fi]
        if kind == "c":
            w = torch.zeros(P["nw"], dtype=torch.int64, device="cuda:0")
            idx = torch.zeros(P["nb"] + 1, dtype=torch.int64, device="cuda:0") if P["var"] else None
            outs.append((w, idx))
            jobs.append((Z.make_job(P["ztype"], P["n"], None, P["f"].data_ptr(), P["p"], w.data_ptr(), 0, P["nw"],
                                    idx.data_ptr() if idx is not None else None), Z.ZFP_HIP_BATCH_COMPRESS))
        else:
            o = torch.zeros_like(P["f"])
            outs.append((o, None))
            jobs.append((Z.make_job(P["ztype"], P["n"], None, o.data_ptr(), P["p"], P["w_in"].data_ptr(), 0,
                                    P["nw"], P["idx_in"].data_ptr() if P["var"] else None),
                         Z.ZFP_HIP_BATCH_DECOMPRESS))
    torch.cuda.synchronize()
    nbytes = [pool[fi]["f"].numel() * pool[fi]["f"].element_size() + pool[fi]["nw"] * 8 for fi, _ in picks]
    # arrivals: a Poisson process at about the rate the server drains them
    arrivals = np.cumsum(rng.exponential(2e-4, size=len(picks)))
    print("serving: GPU", flush=True)
    sv = Z.Server(64 << 30, 0, 2)
    t0 = time.perf_counter()
    ids = []
    for i, (job, direction) in enumerate(jobs):
        while time.perf_counter() - t0 < arrivals[i]:
            pass
        ids.append(sv.submit(job, direction, nbytes[i]))
    res = [sv.wait(r) for r in ids]
    stats = sv.stats()
    sv.close()
    assert stats["done"] == len(picks)
    for i, ((fi, kind), (bits, _)) in enumerate(zip(picks, res)):
        P = pool[fi]
        ow, obits, rec = expect[fi]
        assert bits == obits, f"request {i} ({kind}, field {fi}): {bits} bits, serial zfp {obits}"
        if kind == "c":
            assert np.array_equal(_host_words(outs[i][0], obits), ow[: (obits + 63) // 64]), f"request {i}: stream"
            if P["var"]:
                assert torch.equal(outs[i][1], P["idx_in"]), f"request {i}: block index"
        else:
            width = np.uint32 if P["ztype"] == 3 else np.uint64
            assert np.array_equal(outs[i][0].cpu().numpy().reshape(-1).view(width), rec.view(width)), \
                f"request {i}: decompression"
    # and each equals the same job run alone (one call through zfp_hip_encode/decode)
    for i, (fi, kind) in enumerate(picks[:40]):
        P = pool[fi]
        if kind == "c":
            w = torch.zeros(P["nw"], dtype=torch.int64, device="cuda:0")
            assert Z.encode(P["ztype"], P["n"], None, P["f"].data_ptr(), P["p"], w.data_ptr()) == res[i][0]
            assert torch.equal(w, outs[i][0]), f"request {i}: differs from the job alone"
        else:
            if P["var"]:   # serial zfp's stream with its index, as the request had it
                offs = P["offs"]
                assert Z.hip_lib().zfp_hip_index_import(ctypes.c_void_p(P["w_in"].data_ptr()),
                                                        offs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
                                                        offs.size) == 0
            o = torch.zeros_like(P["f"])
            Z.decode(P["ztype"], P["n"], None, o.data_ptr(), P["p"], P["w_in"].data_ptr(), nwords=P["nw"])
            assert torch.equal(o.view(torch.uint8), outs[i][0].view(torch.uint8)), \
                f"request {i}: differs from the job alone"
