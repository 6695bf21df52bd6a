"""The block codec for compressed arrays (zfp_hip_encode_blocks /
zfp_hip_decode_blocks) against the oracle.

The reference's compressed arrays code one whole block at a time through
zfp::codec::zfp_base::encode_block/decode_block
(include/zfp/codec/zfpcodec.hpp:146-190): a block's bits at a bit offset of
the array's stream, fixed rate with word-aligned blocks (the arrays set
their rate aligned).  Here many blocks go in one launch:
* consecutive blocks: the stream equals serial zfp's stream of the blocks
  as a field 4 x ... x 4*count, in every mode;
* blocks at chosen slots: slot s holds exactly the bits serial zfp writes
  for that block alone, and every other word of the stream is untouched;
  decoding reads the blocks back from their slots;
* a slot past the end of the stream, or slots with a rate that is not
  word-aligned, fail before anything is written.
"""
import ctypes

import numpy as np
import pytest

import zfp_amd as Z
from hiputil import to_device
from zfptest import (T_DOUBLE, T_FLOAT, T_INT32, T_INT64, oracle_compress, oracle_decompress,
                     smooth_field, stream_params_accuracy, stream_params_precision, stream_params_rate)

pytestmark = pytest.mark.gpu

def _blocks(ztype, dims, count, seed):
    """count blocks of 4^dims values, one after another (x fastest)."""
    n = [4] * (dims - 1) + [4 * count]
    return smooth_field(n, ztype, seed=seed), n


def _torch():
    import torch
    return torch


def _encode(ztype, dims, count, blocks, slots, p, words, offset=0):
    bits = ctypes.c_uint64()
    rc = Z.hip_lib().zfp_hip_encode_blocks(ztype, dims, count, ctypes.c_void_p(blocks.data_ptr()),
                                           ctypes.c_void_p(slots.data_ptr() if slots is not None else 0),
                                           *p, ctypes.c_void_p(words.data_ptr()), words.numel(), offset,
                                           ctypes.byref(bits), None)
    return rc, bits.value


def _decode(ztype, dims, count, out, slots, p, words, offset=0):
    bits = ctypes.c_uint64()
    rc = Z.hip_lib().zfp_hip_decode_blocks(ztype, dims, count, ctypes.c_void_p(out.data_ptr()),
                                           ctypes.c_void_p(slots.data_ptr() if slots is not None else 0),
                                           *p, ctypes.c_void_p(words.data_ptr()), words.numel(), offset,
                                           ctypes.byref(bits), None)
    return rc, bits.value


MODES = {"rate8": lambda t, d: stream_params_rate(8, t, d), "prec20": lambda t, d: stream_params_precision(20),
         "acc1e-4": lambda t, d: stream_params_accuracy(1e-4), "reversible": lambda t, d: (1, 16658, 64, -1075)}


@pytest.mark.parametrize("mode", list(MODES))
@pytest.mark.parametrize("ztype", [T_FLOAT, T_DOUBLE, T_INT32, T_INT64])
@pytest.mark.parametrize("dims", [1, 2, 3, 4])
def test_consecutive_blocks_equal_the_field_stream(dims, ztype, mode):
    torch = _torch()
    count = 37
    vals, n = _blocks(ztype, dims, count, seed=dims * 10 + ztype)
    p = MODES[mode](ztype, dims)
    exp_words, exp_bits = oracle_compress(vals, 0, n, None, ztype, p)
    words = torch.zeros(exp_words.size + 8, dtype=torch.int64, device="cuda:0")
    rc, bits = _encode(ztype, dims, count, to_device(vals), None, p, words)
    assert rc == 0 and bits == exp_bits
    got = words.cpu().numpy().view(np.uint64)[:exp_words.size]
    assert np.array_equal(got, exp_words)
    ref = np.zeros_like(vals)
    oracle_decompress(exp_words, ref, 0, n, None, ztype, p)
    out = torch.zeros(vals.size, dtype=to_device(vals).dtype, device="cuda:0")
    rc, dbits = _decode(ztype, dims, count, out, None, p, words)
    assert rc == 0 and dbits == bits
    assert np.array_equal(out.cpu().numpy().view(np.uint8), ref.view(np.uint8))


@pytest.mark.parametrize("ztype", [T_FLOAT, T_DOUBLE, T_INT64])
@pytest.mark.parametrize("dims", [1, 2, 3, 4])
def test_blocks_at_slots(dims, ztype):
    torch = _torch()
    count = 29
    vals, n = _blocks(ztype, dims, count, seed=dims + 3 * ztype)
    p = stream_params_rate(16, ztype, dims, align=True)
    maxbits = p[1]
    assert maxbits % 64 == 0
    wpb = maxbits // 64
    nslots = 64
    rng = np.random.default_rng(dims)
    slots = rng.permutation(nslots)[:count].astype(np.uint64)
    sentinel = np.uint64(0x5A5A5A5A5A5A5A5A)
    words = torch.from_numpy(np.full(nslots * wpb, sentinel, dtype=np.uint64).view(np.int64)).to("cuda:0")
    d_slots = torch.from_numpy(slots.view(np.int64)).to("cuda:0")
    rc, bits = _encode(ztype, dims, count, to_device(vals), d_slots, p, words)
    assert rc == 0 and bits == count * maxbits
    got = words.cpu().numpy().view(np.uint64).reshape(nslots, wpb)
    per = 4 ** dims
    bn = [4] * dims
    for k in range(count):
        blk = np.ascontiguousarray(vals[per * k: per * (k + 1)])
        w, b = oracle_compress(blk, 0, bn, None, ztype, p)
        assert b == maxbits
        assert np.array_equal(got[int(slots[k])], w[:wpb]), f"block {k} at slot {slots[k]}"
    untouched = np.setdiff1d(np.arange(nslots), slots.astype(np.int64))
    assert np.all(got[untouched] == sentinel)
    # read them back from their slots, in another order
    order = rng.permutation(count)
    d_rslots = torch.from_numpy(slots[order].view(np.int64)).to("cuda:0")
    out = torch.zeros(vals.size, dtype=to_device(vals).dtype, device="cuda:0")
    rc, dbits = _decode(ztype, dims, count, out, d_rslots, p, words)
    assert rc == 0 and dbits == count * maxbits
    back = out.cpu().numpy()
    for j, k in enumerate(order):
        blk = np.ascontiguousarray(vals[per * k: per * (k + 1)])
        w, _ = oracle_compress(blk, 0, bn, None, ztype, p)
        ref = np.zeros_like(blk)
        oracle_decompress(w, ref, 0, bn, None, ztype, p)
        assert np.array_equal(back[per * j: per * (j + 1)].view(np.uint8), ref.view(np.uint8))


def test_slot_errors_leave_the_stream_alone():
    torch = _torch()
    count = 4
    vals, n = _blocks(T_DOUBLE, 3, count, seed=1)
    words = torch.full((64,), 7, dtype=torch.int64, device="cuda:0")
    p = stream_params_rate(16, T_DOUBLE, 3, align=True)      # 1024 bits = 16 words a block
    blocks = to_device(vals)
    # slot 4 of a 64-word stream would end at word 80
    d_slots = torch.tensor([0, 1, 2, 4], dtype=torch.int64, device="cuda:0")
    rc, _ = _encode(T_DOUBLE, 3, count, blocks, d_slots, p, words)
    assert rc == Z.ZFP_HIP_EINVAL
    assert torch.all(words == 7)
    # a rate that is not word-aligned
    d_slots = torch.tensor([0, 1, 2, 3], dtype=torch.int64, device="cuda:0")
    rc, _ = _encode(T_DOUBLE, 3, count, blocks, d_slots, stream_params_rate(15.5, T_DOUBLE, 3), words)
    assert rc == Z.ZFP_HIP_EINVAL
    assert torch.all(words == 7)
    # variable-length modes take consecutive blocks only
    rc, _ = _encode(T_DOUBLE, 3, count, blocks, d_slots, stream_params_precision(20), words)
    assert rc == Z.ZFP_HIP_EINVAL
    assert torch.all(words == 7)
    # bad dimensionality or count
    assert _encode(T_DOUBLE, 5, count, blocks, None, p, words)[0] == Z.ZFP_HIP_EINVAL
    assert _encode(T_DOUBLE, 3, 0, blocks, None, p, words)[0] == Z.ZFP_HIP_EINVAL
