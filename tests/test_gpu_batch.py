"""The batched codec (zfp_hip_batch_*): many fields of mixed type,
dimensionality (1-4), size, layout and mode (reversible included) in one
persistent launch per pass and family, replayed from a HIP graph.

Every job's stream must be the oracle's stream for that field bit for bit,
and every job's decompressed array the oracle's reconstruction, exactly as
for one field per call (the batch has no counterpart in the reference; its
results are defined by the serial codec's)."""
import numpy as np
import pytest

import zfp_amd as Z
from zfptest import (T_DOUBLE, T_FLOAT, T_INT32, T_INT64, NP_OF_TYPE, make_layout, max_stream_words,
                     oracle_compress, oracle_decompress, read_layout, smooth_field, stream_params_accuracy,
                     stream_params_precision, stream_params_rate)

pytestmark = pytest.mark.gpu

TYPES = [T_INT32, T_INT64, T_FLOAT, T_DOUBLE]


def _torch():
    import torch
    return torch


def _params(kind, ztype, dims):
    if kind == "reversible":   # zfp_stream_set_reversible (src/zfp.c:750-757)
        return (1, 16658, 64, -1075)
    if kind == "rev_int":      # reversible with a budget and minbits padding (integer types)
        return (300, 2500, 40, -1075)
    if kind == "rate8":
        return stream_params_rate(8, ztype, dims)
    if kind == "rate16":
        return stream_params_rate(16, ztype, dims)
    if kind == "rate5":   # blocks that are not whole words
        return stream_params_rate(5, ztype, dims)
    if kind == "prec20":
        return stream_params_precision(20)
    if kind == "acc":
        return stream_params_accuracy(1e-4)
    return (100, 900, 24, -40)   # expert


class Case:
    """One job: host field (numpy base + layout), device copies, expected
    stream and reconstruction from the oracle."""

    def __init__(self, rng, k, dims=None, mode=None):
        torch = _torch()
        dims = dims or int(rng.integers(1, 4))
        n = [int(v) for v in np.exp(rng.uniform(np.log(3), np.log({1: 400, 2: 40, 3: 14, 4: 7}[dims]), size=dims))]
        self.ztype = TYPES[k % 4]
        mode = mode or ["rate8", "rate16", "rate5", "prec20", "acc", "expert"][(k // 4) % 6]
        self.params = _params(mode, self.ztype, dims)
        self.n = n
        a = smooth_field(n, self.ztype, seed=100 + k, noise=1e-3, special=0.01 if k % 5 == 0 else 0.0)
        layout = ["as_is", "reversed", "permuted", "interleaved"][k % 4] if dims > 1 else ["as_is", "reversed"][k % 2]
        self.base, self.off, self.s = make_layout(a, n, layout)
        self.offset = [0, 0, 7, 64][k % 4]   # bit offset of the field in its stream
        self.words, self.bits = oracle_compress(self.base, self.off, n, self.s, self.ztype, self.params,
                                                offset=self.offset)
        self.nwords = max_stream_words(n, self.ztype, self.params, self.offset)
        self.d_base = torch.from_numpy(np.ascontiguousarray(self.base)).to("cuda:0")
        self.d_words = torch.zeros(self.nwords, dtype=torch.int64, device="cuda:0")
        nb = int(np.prod([(v + 3) // 4 for v in n]))
        self.var = self.params[0] != self.params[1] or self.params[3] < -1074
        self.d_index = torch.zeros(nb + 1, dtype=torch.int64, device="cuda:0") if self.var else None
        self.itemsize = np.dtype(NP_OF_TYPE[self.ztype]).itemsize
        # reconstruction of the oracle's stream, laid out like the field
        self.expect_rec = np.zeros_like(self.base)
        oracle_decompress(self.words, self.expect_rec, self.off, n, self.s, self.ztype, self.params,
                          offset=self.offset)

    def job(self, d_base=None):
        base = self.d_base if d_base is None else d_base
        return Z.make_job(self.ztype, self.n, self.s, base.data_ptr() + self.off * self.itemsize, self.params,
                          self.d_words.data_ptr(), self.offset, self.nwords,
                          self.d_index.data_ptr() if self.var else None)

    def stream(self):
        w = self.d_words.cpu().numpy().view(np.uint64)
        return w[: (self.offset + self.bits + 63) // 64]


def _cases(count, seed):
    rng = np.random.default_rng(seed)
    return [Case(rng, k) for k in range(count)]


def _wide_cases(seed):
    """4D jobs in the lossy modes, reversible jobs of every dimensionality,
    and some 1-3D lossy jobs, so that both kernel families run in one
    batch."""
    rng = np.random.default_rng(seed)
    out = []
    k = 0
    for mode in ["rate8", "rate5", "prec20", "acc", "expert"]:
        for _ in range(2):
            out.append(Case(rng, k, dims=4, mode=mode))
            k += 1
    for dims in (1, 2, 3, 4):
        for _ in range(4):
            mode = "rev_int" if TYPES[k % 4] in (T_INT32, T_INT64) and k % 8 < 4 else "reversible"
            out.append(Case(rng, k, dims=dims, mode=mode))
            k += 1
    for _ in range(6):
        out.append(Case(rng, k))
        k += 1
    return out


def _check_streams(cases, bits):
    for k, (c, b) in enumerate(zip(cases, bits)):
        assert b == c.bits, f"job {k}: {b} bits, oracle {c.bits}"
        assert np.array_equal(c.stream(), c.words[: len(c.stream())]), f"job {k}: stream differs from the oracle"


@pytest.fixture(scope="module")
def cases():
    return _cases(48, 7)


@pytest.fixture(scope="module")
def wide_cases():
    return _wide_cases(9)


def test_batch_compress_matches_oracle(cases):
    b = Z.Batch([c.job() for c in cases], Z.ZFP_HIP_BATCH_COMPRESS)
    bits = b.run()
    _check_streams(cases, bits)
    wg, wpc, bpt = b.shape()
    assert wg in (64, 128, 256) and wpc >= 1 and bpt % 64 == 0
    b.close()


def test_batch_graph_replays_rewrite_everything(cases):
    torch = _torch()
    b = Z.Batch([c.job() for c in cases], Z.ZFP_HIP_BATCH_COMPRESS)
    for rep in range(3):
        # eager work between replays: scribble over every stream buffer
        for c in cases:
            c.d_words.fill_(-1 if rep % 2 else 0x5A5A5A5A)
            if c.offset:
                # the bits below the field's offset belong to the caller
                head = torch.from_numpy(c.words[:1].view(np.int64)).to("cuda:0")
                c.d_words[:1] = head
        torch.cuda.synchronize()
        bits = b.run()
        _check_streams(cases, bits)
    b.close()


@pytest.mark.parametrize("shape", [(64, 2, 64), (256, 8, 256), (128, 4, 1024)])
def test_batch_shapes_give_identical_streams(cases, shape):
    b = Z.Batch([c.job() for c in cases], Z.ZFP_HIP_BATCH_COMPRESS)
    b.set_shape(*shape)
    bits = b.run()
    _check_streams(cases, bits)
    assert b.shape() == shape
    b.close()


def _decompress_and_check(cases):
    torch = _torch()
    comp = Z.Batch([c.job() for c in cases], Z.ZFP_HIP_BATCH_COMPRESS)
    comp.run()
    comp.close()
    outs = [torch.zeros_like(c.d_base) for c in cases]
    dec = Z.Batch([c.job(o) for c, o in zip(cases, outs)], Z.ZFP_HIP_BATCH_DECOMPRESS)
    for _ in range(2):
        for o in outs:
            o.fill_(0)
        bits = dec.run()
        for k, (c, o) in enumerate(zip(cases, outs)):
            assert bits[k] == c.bits, f"job {k}: consumed {bits[k]} bits, oracle {c.bits}"
            got = o.cpu().numpy()
            mine = read_layout(got, c.off, c.n, c.s)
            ref = read_layout(c.expect_rec, c.off, c.n, c.s)
            assert np.array_equal(mine.view(np.uint8), ref.view(np.uint8)), f"job {k}: reconstruction differs"
    dec.close()


def test_batch_decompress_matches_oracle(cases):
    _decompress_and_check(cases)


def test_batch_wide_compress_matches_oracle(wide_cases):
    """4D and reversible jobs (their own kernel family) beside 1-3D lossy
    ones, in one batch."""
    for c in wide_cases:
        c.d_words.zero_()
    b = Z.Batch([c.job() for c in wide_cases], Z.ZFP_HIP_BATCH_COMPRESS)
    for _ in range(2):
        bits = b.run()
        _check_streams(wide_cases, bits)
    b.close()


def test_batch_wide_decompress_matches_oracle(wide_cases):
    _decompress_and_check(wide_cases)


def test_batch_wide_reversible_is_lossless(wide_cases):
    for k, c in enumerate(wide_cases):
        if c.params != (1, 16658, 64, -1075):   # budgeted reversible jobs may drop planes
            continue
        got = read_layout(c.expect_rec, c.off, c.n, c.s)
        want = read_layout(c.base, c.off, c.n, c.s)
        assert np.array_equal(got.view(np.uint8), want.view(np.uint8)), f"job {k}: reversible job lost bits"


def test_batch_matches_per_field_calls(cases):
    """The batch and one zfp_hip_encode per field write the same words."""
    torch = _torch()
    b = Z.Batch([c.job() for c in cases], Z.ZFP_HIP_BATCH_COMPRESS)
    b.run()
    b.close()
    for k, c in enumerate(cases):
        w = torch.zeros_like(c.d_words)
        if c.offset:
            w[:1] = torch.from_numpy(c.words[:1].view(np.int64)).to("cuda:0")
        bits = Z.encode(c.ztype, c.n, c.s, c.d_base.data_ptr() + c.off * c.itemsize, c.params, w.data_ptr(),
                        c.offset)
        assert bits == c.bits
        nw = (c.offset + bits + 63) // 64
        assert torch.equal(w[:nw], c.d_words[:nw]), f"job {k}: per-field stream differs from the batch's"


def test_batch_rejects_what_it_cannot_run():
    torch = _torch()
    d = torch.zeros(4 * 4 * 4 * 4, dtype=torch.float64, device="cuda:0")
    w = torch.zeros(1024, dtype=torch.int64, device="cuda:0")
    # reversible float blocks with minbits > 1: serial zfp decodes such
    # streams only block after block
    with pytest.raises(RuntimeError, match="minbits"):
        Z.Batch([Z.make_job(T_DOUBLE, [4, 4, 4, 4], None, d.data_ptr(), (300, 2500, 64, -1075), w.data_ptr())])
    h = np.zeros(64, dtype=np.float64)
    with pytest.raises(RuntimeError, match="device memory"):
        Z.Batch([Z.make_job(T_DOUBLE, [64], None, h.ctypes.data, stream_params_rate(8, T_DOUBLE, 1),
                            w.data_ptr())])
    pv = stream_params_precision(16)
    with pytest.raises(RuntimeError, match="index"):
        Z.Batch([Z.make_job(T_DOUBLE, [64], None, d.data_ptr(), pv, w.data_ptr(), nwords=1024)],
                Z.ZFP_HIP_BATCH_DECOMPRESS)
