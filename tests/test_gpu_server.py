"""The serving loop (zfp_hip_server_*): requests admitted by HBM budget,
batched as they arrive, several batches in flight.

Each request's result must be the same alone and in the mix: a compression
request writes the stream that one zfp_hip_encode call writes for the same
field (and that the oracle writes), and a decompression request writes the
array that one zfp_hip_decode call writes (and that the oracle writes).  The
codec is integer and IEEE-exact, so "the same" is bit for bit."""
import time

import numpy as np
import pytest

import zfp_amd as Z
from zfptest import (T_DOUBLE, T_FLOAT, NP_OF_TYPE, oracle_compress, oracle_decompress, smooth_field,
                     stream_params_accuracy, stream_params_rate)

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


class Req:
    """A float field (3D unless dims says otherwise) with a mode; its
    expected stream and reconstruction come from the oracle, and device
    buffers for compress and decompress requests."""

    def __init__(self, rng, k, dims=3, reversible=False):
        torch = _torch()
        self.ztype = [T_FLOAT, T_DOUBLE][k % 2]
        self.n = [int(v) for v in rng.integers(6, 30 if dims < 4 else 11, size=dims)]
        if reversible:   # zfp_stream_set_reversible (src/zfp.c:750-757)
            self.params = (1, 16658, 64, -1075)
        elif k % 3 == 2:
            self.params = stream_params_accuracy(float(10.0 ** rng.uniform(-6, -2)))
        else:
            self.params = stream_params_rate(int(rng.integers(4, 33)), self.ztype, dims)
        a = smooth_field(self.n, self.ztype, seed=500 + k, special=0.01 if k % 4 == 0 else 0.0)
        self.host = a
        self.words, self.bits = oracle_compress(a, 0, self.n, None, self.ztype, self.params)
        self.rec = np.zeros_like(a)
        oracle_decompress(self.words, self.rec, 0, self.n, None, self.ztype, self.params)
        self.nwords = len(self.words)
        self.field = torch.from_numpy(a).to("cuda:0")
        self.out = torch.zeros_like(self.field)
        self.w = torch.zeros(self.nwords, dtype=torch.int64, device="cuda:0")
        nb = int(np.prod([(v + 3) // 4 for v in self.n]))
        self.var = self.params[0] != self.params[1] or self.params[3] < -1074
        self.idx = torch.zeros(nb + 1, dtype=torch.int64, device="cuda:0") if self.var else None
        # a stream and index for decompression requests, written up front
        self.w_in = torch.from_numpy(self.words.view(np.int64).copy()).to("cuda:0")
        if self.var:
            _, _, bb = oracle_compress(a, 0, self.n, None, self.ztype, self.params, block_bits=True)
            offs = np.concatenate([[0], np.cumsum(bb.astype(np.int64))])
            self.idx_in = torch.from_numpy(offs).to("cuda:0")
        self.nbytes = self.field.numel() * self.field.element_size() + self.nwords * 8

    def comp_job(self):
        return Z.make_job(self.ztype, self.n, None, self.field.data_ptr(), self.params, self.w.data_ptr(), 0,
                          self.nwords, self.idx.data_ptr() if self.var else None)

    def dec_job(self):
        return Z.make_job(self.ztype, self.n, None, self.out.data_ptr(), self.params, self.w_in.data_ptr(), 0,
                          self.nwords, self.idx_in.data_ptr() if self.var else None)

    def check_compress(self, bits):
        assert bits == self.bits
        got = self.w.cpu().numpy().view(np.uint64)[: (bits + 63) // 64]
        assert np.array_equal(got, self.words[: len(got)])
        if self.var:   # the index the server wrote
            _, _, bb = oracle_compress(self.host, 0, self.n, None, self.ztype, self.params, block_bits=True)
            offs = np.concatenate([[0], np.cumsum(bb.astype(np.int64))])
            assert np.array_equal(self.idx.cpu().numpy(), offs)

    def check_decompress(self, bits):
        assert bits == self.bits
        assert np.array_equal(self.out.cpu().numpy().view(np.uint8), self.rec.view(np.uint8))

    def alone(self):
        """The same two jobs, one call each (zfp_hip_encode / zfp_hip_decode)."""
        torch = _torch()
        w = torch.zeros_like(self.w)
        b1 = Z.encode(self.ztype, self.n, None, self.field.data_ptr(), self.params, w.data_ptr())
        o = torch.zeros_like(self.out)
        b2 = Z.decode(self.ztype, self.n, None, o.data_ptr(), self.params, w.data_ptr())
        return w, b1, o, b2


@pytest.fixture(scope="module")
def reqs():
    rng = np.random.default_rng(11)
    return [Req(rng, k) for k in range(24)]


def _serve(reqs, server, spacing_s=0.0):
    ids = []
    for k, r in enumerate(reqs):
        if k % 2 == 0:
            ids.append((r, "c", server.submit(r.comp_job(), Z.ZFP_HIP_BATCH_COMPRESS, r.nbytes)))
        else:
            ids.append((r, "d", server.submit(r.dec_job(), Z.ZFP_HIP_BATCH_DECOMPRESS, r.nbytes)))
        if spacing_s:
            time.sleep(spacing_s)
    return [(r, kind, *server.wait(i)) for r, kind, i in ids]


def _reset(reqs):
    for r in reqs:
        r.w.zero_()
        r.out.zero_()
        if r.var:
            r.idx.zero_()
    _torch().cuda.synchronize()


def test_mixed_requests_match_oracle_and_alone(reqs):
    _reset(reqs)
    sv = Z.Server(64 << 30)
    done = _serve(reqs, sv)
    for r, kind, bits, lat in done:
        assert lat >= 0
        if kind == "c":
            r.check_compress(bits)
        else:
            r.check_decompress(bits)
    st = sv.stats()
    sv.close()
    assert st["done"] == len(reqs)
    # alone: one call per job gives the same words and values
    torch = _torch()
    for r, kind, bits, _ in done:
        w, b1, o, b2 = r.alone()
        if kind == "c":
            assert b1 == bits and torch.equal(w[: (bits + 63) // 64], r.w[: (bits + 63) // 64])
        else:
            assert b2 == bits and torch.equal(o.view(torch.uint8), r.out.view(torch.uint8))


def test_4d_and_reversible_requests_match_oracle_and_alone(reqs):
    """4D and reversible requests (the batched codec's second kernel family)
    mixed with 3D lossy ones."""
    rng = np.random.default_rng(12)
    wide = [Req(rng, k, dims=4) for k in range(4)]
    wide += [Req(rng, k, dims=int(d), reversible=True) for k, d in enumerate([1, 2, 3, 4, 3, 4])]
    # alternate positions, so that wide requests are both compressions
    # (even places) and decompressions (odd places)
    mix = []
    for i, (a, b) in enumerate(zip(wide, reqs)):
        mix += [a, b] if i % 2 == 0 else [b, a]
    _reset(mix)
    sv = Z.Server(64 << 30)
    done = _serve(mix, sv)
    sv.close()
    torch = _torch()
    for r, kind, bits, _ in done:
        (r.check_compress if kind == "c" else r.check_decompress)(bits)
        w, b1, o, b2 = r.alone()
        if kind == "c":
            assert b1 == bits and torch.equal(w[: (bits + 63) // 64], r.w[: (bits + 63) // 64])
        else:
            assert b2 == bits and torch.equal(o.view(torch.uint8), r.out.view(torch.uint8))


def test_arrivals_over_time_overlap(reqs):
    """Requests arriving while others run join later batches."""
    _reset(reqs)
    sv = Z.Server(64 << 30, 0, 2)
    done = _serve(reqs, sv, spacing_s=0.0005)
    for r, kind, bits, _ in done:
        (r.check_compress if kind == "c" else r.check_decompress)(bits)
    st = sv.stats()
    sv.close()
    assert st["done"] == len(reqs) and st["batches"] >= 2


def test_tight_budget_admits_in_order(reqs):
    """A budget that fits one request at a time still completes every
    request, each with its own result."""
    _reset(reqs)
    sv = Z.Server(max(r.nbytes for r in reqs), 0, 2)
    done = _serve(reqs, sv)
    for r, kind, bits, _ in done:
        (r.check_compress if kind == "c" else r.check_decompress)(bits)
    st = sv.stats()
    sv.close()
    assert st["done"] == len(reqs)


def test_request_larger_than_budget_is_refused(reqs):
    sv = Z.Server(1024)
    with pytest.raises(RuntimeError, match="budget"):
        sv.submit(reqs[0].comp_job(), Z.ZFP_HIP_BATCH_COMPRESS, reqs[0].nbytes)
    sv.close()


def test_inputs_produced_on_a_side_stream(reqs):
    """The server starts a request after the work queued on the stream its
    inputs came from."""
    torch = _torch()
    r = reqs[1]
    side = torch.cuda.Stream()
    field = torch.empty_like(r.field)
    w = torch.zeros_like(r.w)
    idx = torch.zeros_like(r.idx) if r.var else None
    sv = Z.Server(64 << 30)
    with torch.cuda.stream(side):
        torch.cuda._sleep(2_000_000)   # keep the side stream busy before the copy
        field.copy_(r.field)
    job = Z.make_job(r.ztype, r.n, None, field.data_ptr(), r.params, w.data_ptr(), 0, r.nwords,
                     idx.data_ptr() if r.var else None)
    rid = sv.submit(job, Z.ZFP_HIP_BATCH_COMPRESS, r.nbytes, stream=side.cuda_stream)
    bits, _ = sv.wait(rid)
    sv.close()
    assert bits == r.bits
    got = w.cpu().numpy().view(np.uint64)[: (bits + 63) // 64]
    assert np.array_equal(got, r.words[: len(got)])


def test_requests_on_one_buffer_keep_their_order(reqs):
    """compress(A -> W) then decompress(W -> B), back to back: B is the
    oracle's reconstruction of A; decompress(W -> B2) then compress(C -> W):
    B2 is still A's reconstruction and W ends as C's stream; and two
    compressions into one buffer leave the second's stream."""
    torch = _torch()
    a = next(r for r in reqs if not r.var and r.ztype == T_DOUBLE)
    w = torch.zeros(max(a.nwords, 8), dtype=torch.int64, device="cuda:0")
    b = torch.zeros_like(a.field)
    b2 = torch.zeros_like(a.field)
    comp = Z.make_job(a.ztype, a.n, None, a.field.data_ptr(), a.params, w.data_ptr(), 0, a.nwords, None)
    dec = Z.make_job(a.ztype, a.n, None, b.data_ptr(), a.params, w.data_ptr(), 0, a.nwords, None)
    dec2 = Z.make_job(a.ztype, a.n, None, b2.data_ptr(), a.params, w.data_ptr(), 0, a.nwords, None)
    # a second field of the same shape and mode: A's field perturbed
    other = (a.field * 0.5 + 0.25).contiguous()
    host_other = other.cpu().numpy()
    ow, obits = oracle_compress(host_other, 0, a.n, None, a.ztype, a.params)
    comp2 = Z.make_job(a.ztype, a.n, None, other.data_ptr(), a.params, w.data_ptr(), 0, a.nwords, None)
    sv = Z.Server(64 << 30)
    for rep in range(20):
        w.zero_()
        b.zero_()
        b2.zero_()
        torch.cuda.synchronize()
        i1 = sv.submit(comp, Z.ZFP_HIP_BATCH_COMPRESS, a.nbytes)
        i2 = sv.submit(dec, Z.ZFP_HIP_BATCH_DECOMPRESS, a.nbytes)      # reads what i1 writes
        i3 = sv.submit(dec2, Z.ZFP_HIP_BATCH_DECOMPRESS, a.nbytes)     # reads what i1 writes
        i4 = sv.submit(comp2, Z.ZFP_HIP_BATCH_COMPRESS, a.nbytes)      # overwrites what i2, i3 read
        assert sv.wait(i1)[0] == a.bits
        assert sv.wait(i2)[0] == a.bits
        assert sv.wait(i3)[0] == a.bits
        assert sv.wait(i4)[0] == obits
        assert np.array_equal(b.cpu().numpy().view(np.uint8), a.rec.view(np.uint8)), f"repetition {rep}"
        assert np.array_equal(b2.cpu().numpy().view(np.uint8), a.rec.view(np.uint8)), f"repetition {rep}"
        got = w.cpu().numpy().view(np.uint64)[: (obits + 63) // 64]
        assert np.array_equal(got, ow[: len(got)]), f"repetition {rep}"
        # two compressions into one buffer: the later one's stream remains
        w.zero_()
        torch.cuda.synchronize()
        j1 = sv.submit(comp2, Z.ZFP_HIP_BATCH_COMPRESS, a.nbytes)
        j2 = sv.submit(comp, Z.ZFP_HIP_BATCH_COMPRESS, a.nbytes)
        sv.wait(j1)
        sv.wait(j2)
        got = w.cpu().numpy().view(np.uint64)[: (a.bits + 63) // 64]
        assert np.array_equal(got, a.words[: len(got)]), f"repetition {rep}"
    sv.close()
