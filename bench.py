"""Benchmark of the MI355X zfp block codec.

Workload (one GPU): fixed-rate compression and decompression of a 3D float64
512^3 field (1 GiB) resident in HBM, at rates 4, 16 and 32 bits/value, for
the contiguous layout and for the permuted view (z fastest in memory, x
slowest with a negative stride).  The field is a seeded sum of 8 random 3D
sinusoids (wavenumbers 1-6, amplitudes U(0.1, 1)) plus N(0, 1e-3) noise, with
1% of the values replaced by 0, +-1e-300 and +-1e300.

One step = compress + decompress at each of the three rates for both layouts
(12 codec passes over 1 GiB).  `value` is the aggregate uncompressed
throughput of those passes: 12 * 1 GiB / step time.  The roofline entry is
for the pass with the largest mean time per step: algorithmic bytes N*8 + C
per launch over its mean duration, timed with HIP events on the stream it is
launched on; its `traffic` is the HBM bytes per launch that the committed
rocprofv3 counter summary (profiles/) measured for that pass.  `passes`
gives every pass's time and HBM fraction.  The CPU
baseline is the reference's serial zfp_compress/zfp_decompress (and its
OpenMP compression) on a 256^3 slab (128 MiB) of the same field, timed on
this machine's host cores: one warm-up repetition, then five or more within
a budget of about 20 s of CPU time.

Beside the headline line's own workload, the same JSON line reports:
  batched        1024 mixed arrays in one persistent launch per pass (graph
                 replay) against one zfp_hip_encode/decode call per array;
  serving        the request loop at three Poisson offered loads (p50/p99,
                 overall and per request kind);
  variable_rate  precision 16/28 and accuracy 1e-3/1e-7 on a 2D float32
                 8191x8193 and a 4D float64 96x96x96x24 array.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 512]
       [--rates 4,16,32] [--layouts contiguous,permuted]
       [--no-cpu] [--no-batch] [--no-serve] [--no-vr] [--vr-only CASES]
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=10)
    p.add_argument("--warmup", type=int, default=2)
    p.add_argument("--size", type=int, default=512)
    p.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    p.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU baseline")
    p.add_argument("--no-batch", action="store_true", help="skip the batched-codec measurement")
    p.add_argument("--batch-arrays", type=int, default=1024)
    p.add_argument("--batch-reps", type=int, default=9)
    p.add_argument("--rates", default="4,16,32", help="fixed rates of the 512^3 passes")
    p.add_argument("--layouts", default="contiguous,permuted", help="layouts of the 512^3 passes")
    p.add_argument("--no-vr", action="store_true", help="skip the variable-rate measurement")
    p.add_argument("--no-serve", action="store_true", help="skip the serving-loop measurement")
    p.add_argument("--vr-only", default="",
                   help="variable-rate cases to measure, e.g. 2d_f32_8191x8193_precision16 (comma list; default all)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only copy 1 GiB to 1 GiB three times (the known-size kernel that scales the "
                        "rocprofv3 FETCH_SIZE/WRITE_SIZE counters)")
    p.add_argument("--serve-fields", type=int, default=32)
    p.add_argument("--serve-requests", type=int, default=200)
    return p.parse_args()


def make_field(torch, n, seed=20251202, special=0.01):
    """Seeded smooth field on the device, stored x fastest ([z][y][x])."""
    g = torch.Generator().manual_seed(seed)
    k = torch.randint(1, 7, (8, 3), generator=g).double()
    amp = 0.1 + 0.9 * torch.rand(8, generator=g, dtype=torch.float64)
    ph = 2 * math.pi * torch.rand(8, generator=g, dtype=torch.float64)
    dev = "cuda:0"
    c = torch.arange(n, device=dev, dtype=torch.float64) * (2 * math.pi / n)
    f = torch.zeros((n, n, n), device=dev, dtype=torch.float64)
    for i in range(8):
        ax = (k[i, 0] * c).view(1, 1, n)
        ay = (k[i, 1] * c).view(1, n, 1)
        az = (k[i, 2] * c).view(n, 1, 1)
        f += amp[i].item() * torch.sin(ax + ay + az + ph[i].item())
    gd = torch.Generator(device=dev).manual_seed(seed + 1)
    f += 1e-3 * torch.randn((n, n, n), device=dev, dtype=torch.float64, generator=gd)
    if special:
        flat = f.view(-1)
        pick = torch.rand(flat.numel(), device=dev, generator=gd) < special
        choice = torch.randint(0, 5, (flat.numel(),), device=dev, generator=gd)
        vals = torch.tensor([0.0, 1e-300, -1e-300, 1e300, -1e300], device=dev, dtype=torch.float64)
        flat[pick] = vals[choice[pick]]
    return f


COUNTERS = os.path.join(ROOT, "profiles", "codec_f64_3d_512_counters.json")


def measured_traffic(pass_name):
    """HBM bytes per pass from the committed rocprofv3 counter summary
    (FETCH_SIZE and WRITE_SIZE of every kernel of the pass, scaled by what
    the same counters read for a copy kernel of known size), or None when
    the summary has no entry for the pass."""
    if not os.path.exists(COUNTERS):
        return None
    with open(COUNTERS) as fh:
        c = json.load(fh)
    e = c.get("passes", {}).get(pass_name)
    if not e:
        return None
    return int(e["fetch_bytes"] + e["write_bytes"])


def pass_kernels(op, rate, layout):
    """The kernels one pass launches (codec_kernels.h launch_encode /
    launch_decode for 3D doubles at a word-aligned fixed rate)."""
    if op == "compress":
        if rate <= 16 and layout == "contiguous":
            return "k_encode_slot<double,3,3>"
        return "k_encode_slot<double,3,2>"
    if rate <= 16:
        return "k_decode<double,3,STAGE,false,HI_ONLY> + k_decode_list<double,3>"
    return "k_decode<double,3,STAGE>"


def gpu_bench(args):
    import torch
    import zfp_amd as Z

    if Z.hip_lib().zfp_hip_device_check() != 0:
        raise SystemExit("bench.py: no usable gfx950 device: " + Z.hip_lib().zfp_hip_last_error().decode())
    n = args.size
    N = n ** 3
    f = make_field(torch, n)
    # permuted view: element (x,y,z) at (n-1-x)*n*n + y*n + z
    fp = torch.flip(f.permute(2, 1, 0), dims=[0]).contiguous()
    torch.cuda.synchronize()
    layouts = {
        "contiguous": (f.data_ptr(), [0, 0, 0, 0]),
        "permuted": (fp.data_ptr() + (n - 1) * n * n * 8, [-n * n, n, 1, 0]),
    }
    layouts = {k: v for k, v in layouts.items() if k in args.layouts.split(",")}
    out = torch.empty_like(f)
    rates = [int(r) for r in args.rates.split(",")]
    streams = {r: torch.zeros(N * r // 64 + 8, dtype=torch.int64, device="cuda:0") for r in rates}
    st = torch.cuda.Stream()  # default priority
    sh = st.cuda_stream
    shape = [n, n, n]
    params = {r: (64 * r, 64 * r, 64, -1074) for r in rates}  # zfp_stream_set_rate(r, double, 3D)

    # decompression writes the field back in the layout it was read from
    out_ptr = {"contiguous": out.data_ptr(), "permuted": out.data_ptr() + (n - 1) * n * n * 8}

    def run_pass(op, r, layout):
        ptr, s = layouts[layout]
        if op == "compress":
            return Z.encode(4, shape, s, ptr, params[r], streams[r].data_ptr(), 0, stream=sh)
        return Z.decode(4, shape, s, out_ptr[layout], params[r], streams[r].data_ptr(), 0, stream=sh)

    passes = [(op, r, lay) for lay in layouts for r in rates for op in ("compress", "decompress")]
    # parity properties at full size: both layouts give the same stream, the
    # fixed-rate size law holds, decompression consumes what compression
    # wrote, and both layouts decompress to the same values
    checks = {}
    lay0 = next(iter(layouts))
    with torch.cuda.stream(st):  # torch's copies and the codec run in order on one stream
        for r in rates:
            bits_c = run_pass("compress", r, lay0)
            bits_d = run_pass("decompress", r, lay0)
            checks[f"rate{r}"] = {"bits": bits_c, "size_law": bits_c == N * r, "decode_bits": bits_d == bits_c}
            if len(layouts) == 2:
                s1 = streams[r].clone()
                bits_p = run_pass("compress", r, "permuted")
                same = bool(torch.equal(s1, streams[r]))
                del s1
                o1 = out.clone()
                run_pass("decompress", r, "permuted")
                same_out = bool(torch.equal(o1.view(torch.int64), torch.flip(out.permute(2, 1, 0), dims=[2])
                                            .contiguous().view(torch.int64)))
                del o1
                checks[f"rate{r}"].update({"layouts_equal": same and bits_p == bits_c,
                                           "layouts_decode_equal": same_out})
    torch.cuda.synchronize()

    for _ in range(args.warmup):
        for p in passes:
            run_pass(*p)
    torch.cuda.synchronize()

    # per-pass kernel timing with events on the launch stream
    ev = {p: [] for p in passes}
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for p in passes:
            a = torch.cuda.Event(enable_timing=True)
            b = torch.cuda.Event(enable_timing=True)
            a.record(st)
            run_pass(*p)
            b.record(st)
            ev[p].append((a, b))
    st.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_step = (t1 - t0) * 1e3 / args.steps
    per = {}
    for p, lst in ev.items():
        per[p] = sum(a.elapsed_time(b) for a, b in lst) / len(lst)
    ubytes = N * 8
    value = len(passes) * ubytes / (ms_step * 1e-3) / 1e9
    detail = {}
    for (op, r, lay), ms in per.items():
        cbytes = N * r // 8
        detail[f"{op}_rate{r}_{lay}"] = {
            "ms": round(ms, 4),
            "GBps_uncompressed": round(ubytes / (ms * 1e-3) / 1e9, 1),
            "hbm_frac": round((ubytes + cbytes) / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        }
    # the dominant pass: the one that takes longest per step
    dom = max(per, key=lambda p: per[p])
    dom_ms = per[dom]
    dom_name = f"{dom[0]}_rate{dom[1]}_{dom[2]}"
    alg_bytes = ubytes + N * dom[1] // 8
    achieved = alg_bytes / (dom_ms * 1e-3) / 1e9
    traffic = measured_traffic(dom_name)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "pass": dom_name, "kernel": pass_kernels(*dom),
                "algorithmic_bytes_per_launch": alg_bytes, "mean_launch_ms": round(dom_ms, 4),
                "traffic_over_algorithmic": round(traffic / alg_bytes, 4) if traffic else None,
                "min_frac_over_passes": round(min(d["hbm_frac"] for d in detail.values()), 4)}
    for (op, r, lay), d in zip(per, detail.values()):
        d["kernels"] = pass_kernels(op, r, lay)
        t = measured_traffic(f"{op}_rate{r}_{lay}")
        d["traffic_over_algorithmic"] = round(t / (ubytes + N * r // 8), 4) if t else None
    return value, ms_step, roofline, detail, checks, f


def batch_fields(torch, count, seed):
    """1024 (count) arrays: dims U{1,2,3}, each axis log-uniform in [5, 300],
    type U{f32, f64, i32, i64}, mode U{rate 8, rate 16, precision 20,
    accuracy 1e-4}; smooth seeded fields (integers scaled to |f| < 2^28 or
    2^60).  Returns a list of (ztype, n, mode, tensor)."""
    import numpy as np
    rng = np.random.default_rng(seed)
    types = [4, 3, 1, 2]   # zfp_type_double, float, int32, int64
    out = []
    for k in range(count):
        dims = int(rng.integers(1, 4))
        n = [int(round(math.exp(rng.uniform(math.log(5), math.log(300))))) for _ in range(dims)]
        ztype = types[int(rng.integers(0, 4))]
        mode = ["rate8", "rate16", "prec20", "acc1e-4"][int(rng.integers(0, 4))]
        shape = list(reversed(n))
        f = torch.zeros(shape, device="cuda:0", dtype=torch.float64)
        for _ in range(4):
            term = torch.full(shape, float(rng.uniform(0, 2 * math.pi)), device="cuda:0", dtype=torch.float64)
            for d in range(dims):
                c = torch.arange(n[d], device="cuda:0", dtype=torch.float64) * (2 * math.pi * int(rng.integers(1, 7)) / n[d])
                view = [1] * dims
                view[dims - 1 - d] = n[d]
                term = term + c.view(view)
            f += float(rng.uniform(0.1, 0.25)) * torch.sin(term)
        f += 1e-3 * torch.randn(shape, device="cuda:0", dtype=torch.float64)
        if ztype == 4:
            t = f
        elif ztype == 3:
            t = f.float()
        elif ztype == 1:
            t = torch.round(f * (1 << 28)).to(torch.int32)
        else:
            t = torch.round(f * (1 << 60)).to(torch.int64)
        out.append((ztype, n, mode, t.contiguous()))
    return out


def batch_bench(args):
    """Batched codec over 1024 mixed arrays (one graph replay per direction)
    against a loop of one zfp_hip_encode / zfp_hip_decode call per array;
    both write bit-identical streams and arrays."""
    import torch
    import zfp_amd as Z

    fields = batch_fields(torch, args.batch_arrays, 20251203)
    ubytes = sum(t.numel() * t.element_size() for _, _, _, t in fields)
    jobs_c, jobs_d, per, outs = [], [], [], []
    modes = {"rate8": ("rate", 8), "rate16": ("rate", 16), "prec20": ("precision", 20),
             "acc1e-4": ("accuracy", 1e-4)}
    for ztype, n, mode, t in fields:
        dims = len(n)
        p = Z.stream_params(*modes[mode], ztype, dims)
        nw = Z.max_stream_words(ztype, n, p)
        nb = math.prod((v + 3) // 4 for v in n)
        w_batch = torch.zeros(nw, dtype=torch.int64, device="cuda:0")
        w_loop = torch.zeros(nw, dtype=torch.int64, device="cuda:0")
        var = p[0] != p[1]
        idx = torch.zeros(nb + 1, dtype=torch.int64, device="cuda:0") if var else None
        o_batch = torch.empty_like(t)
        o_loop = torch.empty_like(t)
        jobs_c.append(Z.make_job(ztype, n, None, t.data_ptr(), p, w_batch.data_ptr(), 0, nw,
                                 idx.data_ptr() if var else None))
        jobs_d.append(Z.make_job(ztype, n, None, o_batch.data_ptr(), p, w_batch.data_ptr(), 0, nw,
                                 idx.data_ptr() if var else None))
        per.append((ztype, n, p, t, w_loop, o_loop))
        outs.append((w_batch, o_batch, idx))
    torch.cuda.synchronize()

    bc = Z.Batch(jobs_c, Z.ZFP_HIP_BATCH_COMPRESS)
    bits_b = bc.run()   # tunes the launch shape, then runs
    bd = Z.Batch(jobs_d, Z.ZFP_HIP_BATCH_DECOMPRESS)
    bd.run()

    def timed(fn, reps):
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            ts.append(time.perf_counter() - t0)
        return ts

    def loop_c():
        for ztype, n, p, t, w, _ in per:
            Z.encode(ztype, n, None, t.data_ptr(), p, w.data_ptr())

    def loop_d():
        for ztype, n, p, _, w, o in per:
            Z.decode(ztype, n, None, o.data_ptr(), p, w.data_ptr())

    tb_c = timed(bc.run, args.batch_reps)
    tb_d = timed(bd.run, args.batch_reps)
    loop_c()   # warm-up (and registers the variable-rate indexes for loop_d)
    loop_d()
    tl_c = timed(loop_c, max(2, args.batch_reps // 3))
    tl_d = timed(loop_d, max(2, args.batch_reps // 3))
    torch.cuda.synchronize()
    # bit-exactness: the batch and the per-array calls agree on every word
    # and every decompressed value (each of them is parity-tested against
    # the oracle at small sizes)
    same_streams = same_arrays = True
    for k, (ztype, n, p, t, w_loop, o_loop) in enumerate(per):
        w_b, o_b, _ = outs[k]
        nw = (bits_b[k] + 63) // 64
        same_streams &= bool(torch.equal(w_b[:nw], w_loop[:nw]))
        same_arrays &= bool(torch.equal(o_b.view(torch.uint8), o_loop.view(torch.uint8)))
    shape = bc.shape()
    bc.close()
    bd.close()

    def pct(ts, q):
        s = sorted(ts)
        return s[min(len(s) - 1, int(round(q * (len(s) - 1))))] * 1e3

    med = lambda ts: pct(ts, 0.5) / 1e3   # noqa: E731
    return {
        "workload": f"{len(fields)} arrays, dims U{{1,2,3}}, axes log-uniform in [5,300], "
                    "f32/f64/i32/i64, rate 8/16, precision 20, accuracy 1e-4",
        "uncompressed_GB": round(ubytes / 1e9, 3),
        "batched_compress_GBps": round(ubytes / med(tb_c) / 1e9, 1),
        "per_call_compress_GBps": round(ubytes / med(tl_c) / 1e9, 1),
        "compress_speedup": round(med(tl_c) / med(tb_c), 2),
        "batched_decompress_GBps": round(ubytes / med(tb_d) / 1e9, 1),
        "per_call_decompress_GBps": round(ubytes / med(tl_d) / 1e9, 1),
        "decompress_speedup": round(med(tl_d) / med(tb_d), 2),
        "batch_ms_p50": round(pct(tb_c, 0.5), 3), "batch_ms_p99": round(pct(tb_c, 0.99), 3),
        "per_call_loop_ms_p50": round(pct(tl_c, 0.5), 3),
        "shape": {"wg": shape[0], "waves_per_cu": shape[1], "blocks_per_task": shape[2]},
        "bit_exact_vs_per_call": {"streams": same_streams, "arrays": same_arrays},
    }


def vr_fields(torch):
    """2D float32 8191x8193 and 4D float64 96x96x96x24: smooth seeded fields
    with a 10% box of U(-1,1) noise (x fastest in memory)."""
    g = torch.Generator(device="cuda:0").manual_seed(20251205)
    out = {}
    ny, nx = 8193, 8191
    cx = torch.arange(nx, device="cuda:0", dtype=torch.float64) * (2 * math.pi / nx)
    cy = torch.arange(ny, device="cuda:0", dtype=torch.float64) * (2 * math.pi / ny)
    f = torch.sin(3 * cx).view(1, nx) * torch.cos(2 * cy).view(ny, 1) + 0.5 * torch.sin(5 * cx + 4 * cy.view(ny, 1))
    by, bx = int(ny * 0.316), int(nx * 0.316)   # ~10% of the area
    f[:by, :bx] = 2 * torch.rand((by, bx), device="cuda:0", dtype=torch.float64, generator=g) - 1
    out["2d_f32_8191x8193"] = (3, [nx, ny], f.float().contiguous())
    n4 = [96, 96, 96, 24]
    c = [torch.arange(v, device="cuda:0", dtype=torch.float64) * (2 * math.pi / v) for v in n4]
    f = (torch.sin(2 * c[0]).view(1, 1, 1, 96) * torch.cos(c[1]).view(1, 1, 96, 1)
         + torch.sin(3 * c[2]).view(1, 96, 1, 1) * torch.cos(2 * c[3]).view(24, 1, 1, 1))
    f[:, :, :30, :30] = 2 * torch.rand((24, 96, 30, 30), device="cuda:0", dtype=torch.float64, generator=g) - 1
    out["4d_f64_96x96x96x24"] = (4, n4, f.contiguous())
    return out


def vr_bench(args):
    """Variable-rate modes (precision 16/28, accuracy 1e-3/1e-7) on the 2D
    float32 and 4D float64 arrays: each compression codes every block once
    (2D) or runs a length pass, device scan and write pass (4D);
    decompression from the recorded block index.  Timed
    per call (wall clock, after a warm-up); checks: decompression consumes
    exactly the bits compression wrote, and the batched kernels (a one-job
    batch, 2D only) write the same stream as the per-field path."""
    import torch
    import zfp_amd as Z

    res = {}
    modes = [("precision", 16), ("precision", 28), ("accuracy", 1e-3), ("accuracy", 1e-7)]
    for name, (ztype, n, f) in vr_fields(torch).items():
        dims = len(n)
        ubytes = f.numel() * f.element_size()
        for mode, val in modes:
            if args.vr_only and f"{name}_{mode}{val}" not in args.vr_only.split(","):
                continue
            p = Z.stream_params(mode, val, ztype, dims)
            nw = Z.max_stream_words(ztype, n, p)
            w = torch.zeros(nw, dtype=torch.int64, device="cuda:0")
            o = torch.empty_like(f)
            bits = Z.encode(ztype, n, None, f.data_ptr(), p, w.data_ptr())
            dbits = Z.decode(ztype, n, None, o.data_ptr(), p, w.data_ptr())
            tc, td = [], []
            for _ in range(3):
                t0 = time.perf_counter()
                Z.encode(ztype, n, None, f.data_ptr(), p, w.data_ptr())
                t1 = time.perf_counter()
                Z.decode(ztype, n, None, o.data_ptr(), p, w.data_ptr())
                t2 = time.perf_counter()
                tc.append(t1 - t0)
                td.append(t2 - t1)
            entry = {"bits_per_value": round(bits / f.numel(), 3),
                     "compress_GBps": round(ubytes / min(tc) / 1e9, 1),
                     "decompress_GBps": round(ubytes / min(td) / 1e9, 1),
                     "decode_bits_equal": dbits == bits}
            if dims <= 3:
                nb = math.prod((v + 3) // 4 for v in n)
                wb = torch.zeros_like(w)
                idx = torch.zeros(nb + 1, dtype=torch.int64, device="cuda:0")
                b = Z.Batch([Z.make_job(ztype, n, None, f.data_ptr(), p, wb.data_ptr(), 0, nw, idx.data_ptr())])
                bb = b.run()[0]
                b.close()
                k = (bits + 63) // 64
                entry["batch_stream_equal"] = bb == bits and bool(torch.equal(wb[:k], w[:k]))
                del wb, idx
            res[f"{name}_{mode}{val}"] = entry
            del w, o
        del f
    torch.cuda.synchronize()
    return res


def serve_pool(torch, count, rng):
    """The serving workload's resident fields: 3D float64 / float32 in turn,
    edge log-uniform in [32, 384], fixed rate U{4..32} or accuracy
    10^U(-6,-2), a smooth seeded field plus N(0, 1e-3) noise.  Returns dicts
    with ztype, n, p (codec parameters), f (device tensor), nw (stream
    capacity in words), nb (blocks) and var (variable length)."""
    import zfp_amd as Z
    pool = []
    for k in range(count):
        e = int(round(math.exp(rng.uniform(math.log(32), math.log(384)))))
        n = [e, e, e]
        ztype = 4 if k % 2 == 0 else 3
        if rng.random() < 0.5:
            p = Z.stream_params("rate", int(rng.integers(4, 33)), ztype, 3)
        else:
            p = Z.stream_params("accuracy", float(10.0 ** rng.uniform(-6, -2)), ztype, 3)
        c = torch.arange(e, device="cuda:0", dtype=torch.float64) * (2 * math.pi / e)
        f = (torch.sin(3 * c).view(1, 1, e) * torch.cos(2 * c).view(1, e, 1) + torch.sin(c + 1).view(e, 1, 1)
             + 1e-3 * torch.randn((e, e, e), device="cuda:0", dtype=torch.float64))
        f = f if ztype == 4 else f.float()
        pool.append(dict(ztype=ztype, n=n, p=p, f=f, nw=Z.max_stream_words(ztype, n, p),
                         nb=(e // 4 + (e % 4 > 0)) ** 3, var=p[0] != p[1]))
    return pool


def serve_picks(rng, count, npool):
    """The request trace: (pool field, "c" compress or "d" decompress), 50/50."""
    return [(int(rng.integers(0, npool)), "c" if rng.random() < 0.5 else "d") for _ in range(count)]


def serve_bench(args):
    """Serving loop: a pool of resident 3D float32/float64 fields (volume
    log-uniform 32^3..384^3; fixed rate U{4..32} or accuracy 10^U(-6,-2)),
    and R requests (50/50 compress/decompress of a random pool field) that
    arrive as a Poisson process at 0.3/0.7/0.95 of the capacity measured
    with all requests queued at once.  Latency is submission to completion,
    per request; every result is compared with the same job run alone."""
    import numpy as np
    import torch
    import zfp_amd as Z

    rng = np.random.default_rng(20251204)
    pool = []
    for P in serve_pool(torch, args.serve_fields, rng):
        ztype, n, p, f, nw, nb, var = (P[k] for k in ("ztype", "n", "p", "f", "nw", "nb", "var"))
        # reference results: the job alone, one call each
        w = torch.zeros(nw, dtype=torch.int64, device="cuda:0")
        bits = Z.encode(ztype, n, None, f.data_ptr(), p, w.data_ptr())
        o = torch.empty_like(f)
        Z.decode(ztype, n, None, o.data_ptr(), p, w.data_ptr())
        idx = None
        if var:
            # the index of the stream, from a one-job batch (same kernels)
            idx = torch.zeros(nb + 1, dtype=torch.int64, device="cuda:0")
            wi = torch.zeros_like(w)
            b1 = Z.Batch([Z.make_job(ztype, n, None, f.data_ptr(), p, wi.data_ptr(), 0, nw, idx.data_ptr())])
            b1.run()
            b1.close()
            del wi
        pool.append(dict(ztype=ztype, n=n, p=p, f=f, nw=nw, nb=nb, var=var, w=w, bits=bits, o=o, idx=idx,
                         ubytes=f.numel() * f.element_size()))
    torch.cuda.synchronize()

    R = args.serve_requests
    picks = serve_picks(rng, R, len(pool))
    # per-request output buffers, allocated once
    outs = []
    for fi, kind in picks:
        P = pool[fi]
        if kind == "c":
            outs.append((torch.zeros(P["nw"], dtype=torch.int64, device="cuda:0"),
                         torch.zeros(P["nb"] + 1, dtype=torch.int64, device="cuda:0") if P["var"] else None))
        else:
            outs.append((torch.empty_like(P["f"]), None))
    torch.cuda.synchronize()

    def job(i):
        fi, kind = picks[i]
        P = pool[fi]
        buf, idx = outs[i]
        if kind == "c":
            return (Z.make_job(P["ztype"], P["n"], None, P["f"].data_ptr(), P["p"], buf.data_ptr(), 0, P["nw"],
                               idx.data_ptr() if idx is not None else None), Z.ZFP_HIP_BATCH_COMPRESS)
        return (Z.make_job(P["ztype"], P["n"], None, buf.data_ptr(), P["p"], P["w"].data_ptr(), 0, P["nw"],
                           P["idx"].data_ptr() if P["var"] else None), Z.ZFP_HIP_BATCH_DECOMPRESS)

    jobs = [job(i) for i in range(R)]
    nbytes = [pool[fi]["ubytes"] + pool[fi]["nw"] * 8 for fi, _ in picks]
    ubytes = sum(pool[fi]["ubytes"] for fi, _ in picks)

    def run(arrivals):
        sv = Z.Server(64 << 30, 0, 2)
        t0 = time.perf_counter()
        ids = []
        for i in range(R):
            while arrivals is not None and time.perf_counter() - t0 < arrivals[i]:
                pass
            ids.append(sv.submit(jobs[i][0], jobs[i][1], nbytes[i]))
        res = [sv.wait(r) for r in ids]
        t1 = time.perf_counter()
        st = sv.stats()
        sv.close()
        return res, t1 - t0, st

    run(None)   # warm-up
    res, t_all, _ = run(None)
    capacity = ubytes / t_all   # bytes/s with everything queued at once
    mean_req = ubytes / R
    loads = {}
    exact = True
    for L in (0.3, 0.7, 0.95):
        lam = L * capacity / mean_req
        arr = np.cumsum(rng.exponential(1.0 / lam, size=R))
        res, t_run, st = run(arr)

        def pct(lat):
            lat = sorted(lat)
            n = len(lat)
            return (round(lat[n // 2], 3), round(lat[min(n - 1, int(0.99 * (n - 1) + 0.5))], 3)) if n else (None, None)

        p50, p99 = pct([x[1] for x in res])
        c50, c99 = pct([x[1] for x, (_, kind) in zip(res, picks) if kind == "c"])
        d50, d99 = pct([x[1] for x, (_, kind) in zip(res, picks) if kind == "d"])
        loads[str(L)] = {"offered_req_per_s": round(lam, 1), "p50_ms": p50, "p99_ms": p99,
                         "compress_p50_ms": c50, "compress_p99_ms": c99,
                         "decompress_p50_ms": d50, "decompress_p99_ms": d99,
                         "GBps": round(ubytes / t_run / 1e9, 1), "batches": st["batches"],
                         "max_batch": st["max_batch"], "max_inflight": st["max_inflight"]}
        # every result equals the job run alone
        for i, (fi, kind) in enumerate(picks):
            P = pool[fi]
            bits = res[i][0]
            if bits != P["bits"]:
                exact = False
            elif kind == "c":
                nw = (bits + 63) // 64
                exact &= bool(torch.equal(outs[i][0][:nw], P["w"][:nw]))
            else:
                exact &= bool(torch.equal(outs[i][0].view(torch.uint8), P["o"].view(torch.uint8)))
    return {
        "workload": f"{R} requests (50/50 compress/decompress) over {len(pool)} resident 3D f32/f64 fields, "
                    "volume log-uniform 32^3..384^3, rate U{4..32} or accuracy 10^U(-6,-2), "
                    "Poisson arrivals, 64 GiB HBM budget, 2 batches in flight",
        "capacity_GBps": round(capacity / 1e9, 1),
        "loads": loads,
        "bit_exact_vs_alone": exact,
    }


def host_provenance(kind):
    """What the CPU baseline ran on and how the timed library was built:
    lscpu's model, sockets, cores and threads, the CPUs this process may
    use, and the compiler and flags of the reference build (oracle/Makefile's
    libzfp_ref.so recipe) or of the oracle."""
    import subprocess
    info = {}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core", "CPU(s)": "cpus", "CPU max MHz": "max_mhz"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                info[keys[k.strip()]] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    try:
        info["usable_cpus"] = len(os.sched_getaffinity(0))
    except AttributeError:
        info["usable_cpus"] = os.cpu_count()
    try:
        info["compiler"] = subprocess.run(["gcc", "--version"], capture_output=True, text=True,
                                          timeout=10).stdout.splitlines()[0]
    except (OSError, subprocess.SubprocessError, IndexError):
        info["compiler"] = None
    target = "_ref/libzfp_ref.so:" if kind == "reference" else "liboracle.so:"
    flags = None
    try:
        lines = open(os.path.join(ROOT, "oracle", "Makefile")).read().splitlines()
        for i, line in enumerate(lines):
            if line.startswith("$(HERE)" + target):
                rec = " ".join(x.strip().rstrip("\\") for x in lines[i + 1:i + 4] if x.startswith(("\t", " ")))
                rec = rec[rec.find("$(CC)"):]
                if kind != "reference":
                    rec = rec.replace("$(CFLAGS)", "-O2 -fPIC -std=c99 -ffp-contract=off")
                flags = " ".join(t for t in rec.split() if t.startswith("-") and not t.startswith(("-I", "-o", "-L")))
                break
    except OSError:
        pass
    info["flags"] = flags
    return info


def cpu_baseline(args, f_dev):
    """Reference serial zfp on a 256^3 slab of the benchmark field (or the
    oracle when the reference build is absent): one warm-up repetition,
    then at least five timed ones, more while the budget lasts."""
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import zfptest as T
    m = min(256, f_dev.shape[0])
    slab = f_dev[:m, :m, :m].contiguous().cpu().numpy().reshape(-1)
    n = [m, m, m]
    rates = [4, 16, 32]
    kind = "reference" if T.have_ref() else "port"
    cores = os.cpu_count() or 1
    res = {"serial_compress": [], "serial_decompress": [], "omp_compress": []}
    t_start = time.perf_counter()
    t_budget = t_start + args.cpu_seconds
    reps = -1  # the first repetition warms up and is not recorded
    while reps < 5 or (reps < 9 and time.perf_counter() < t_budget):
        for r in rates:
            p = T.stream_params_rate(r, T.T_DOUBLE, 3)
            if kind == "reference":
                api = T.ref_api()
                t0 = time.perf_counter()
                words, size = T.api_compress(api, slab, 0, n, None, T.T_DOUBLE, p)
                t1 = time.perf_counter()
                out = np.zeros_like(slab)
                T.api_decompress(api, words, out, 0, n, None, T.T_DOUBLE, p)
                t2 = time.perf_counter()
                T.api_compress(api, slab, 0, n, None, T.T_DOUBLE, p, policy=1, threads=cores)
                t3 = time.perf_counter()
                if reps >= 0:
                    res["omp_compress"].append((r, t3 - t2))
            else:
                t0 = time.perf_counter()
                words, bits = T.oracle_compress(slab, 0, n, None, T.T_DOUBLE, p)
                t1 = time.perf_counter()
                out = np.zeros_like(slab)
                T.oracle_decompress(words, out, 0, n, None, T.T_DOUBLE, p)
                t2 = time.perf_counter()
            if reps >= 0:
                res["serial_compress"].append((r, t1 - t0))
                res["serial_decompress"].append((r, t2 - t1))
        reps += 1
    nbytes = slab.nbytes

    def med(key):
        ts = sorted(t for _, t in res[key])
        return ts[len(ts) // 2] if ts else None

    sc, sd = med("serial_compress"), med("serial_decompress")
    # same definition as the GPU value: uncompressed bytes of compress+decompress passes / time
    value = 2 * nbytes / (sc + sd) / 1e9
    out = {"value": round(value, 4), "unit": "GB/s", "cores": 1, "kind": kind,
           "sample": f"3D float64 {m}^3 slab ({nbytes >> 20} MiB) of the workload field, rates 4/16/32, "
                     f"1 warm-up + {reps} timed repetitions, median of serial zfp_compress + zfp_decompress",
           "cpu_seconds": round(time.perf_counter() - t_start, 1),
           "serial_compress_GBps": round(nbytes / sc / 1e9, 4),
           "serial_decompress_GBps": round(nbytes / sd / 1e9, 4)}
    oc = med("omp_compress")
    if oc:
        out["omp_compress_GBps"] = round(nbytes / oc / 1e9, 4)
        out["omp_threads"] = cores
    out["host"] = host_provenance(kind)
    return out


def copy_calibration():
    import torch
    src = torch.ones(1 << 27, dtype=torch.float64, device="cuda:0")   # 1 GiB
    dst = torch.empty_like(src)
    for _ in range(3):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration": {"bytes_read": src.numel() * 8, "bytes_written": dst.numel() * 8,
                                           "launches": 3}}))


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              f"(sharding a field across GPUs is not built)", file=sys.stderr)
        raise SystemExit(2)
    value, ms_step, roofline, detail, checks, f = gpu_bench(args)
    cpu = None if args.no_cpu else cpu_baseline(args, f)
    del f
    batched = None if args.no_batch else batch_bench(args)
    serving = None if args.no_serve else serve_bench(args)
    variable_rate = None if args.no_vr else vr_bench(args)
    line = {
        "metric": "compress/decompress GB/s (uncompressed) and % HBM peak; batched vs per-call",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"fixed-rate 3D float64 {args.size}^3 (1 GiB at 512^3), rates 4/16/32, "
                               f"contiguous + permuted negative-stride view, compress + decompress",
                   "passes_per_step": 12},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "passes": detail,
        "parity_checks": checks,
        "batched": batched,
        "serving": serving,
        "variable_rate": variable_rate,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
