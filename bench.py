#!/usr/bin/env python3
"""Benchmark of the fused FastLanes decode + filter + aggregate hot path.

Workload (one step = one pass of the fused kernel over the whole table, inputs
resident in HBM): a 120M-row fact table -- orderdate sorted yyyymmdd uniform
over 1992..1998 as FastLanes delta, quantity U{1..50} at bw 6, discount
U{0..10} at bw 4, price U[90000,10^7) at bw 24, all in the unified transposed
layout -- and the query year==1993 AND 1<=discount<=3 AND quantity<25,
SUM(price*discount) as int64.

Prints one JSON line.  `value` is rows/s of the whole timed region; `roofline`
is the dominant kernel's encoded-byte throughput against MI355X's 8 TB/s,
timed with HIP events on the stream the kernel runs on; `cpu_baseline` is the
CPU decode + query restatement timed on this host's cores.

    python bench.py [--steps K] [--warmup W] [--rows N] [--no-cpu]
"""
import argparse
import ctypes
import json
import os
import pathlib
import statistics
import sys
import time

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
METRIC = "fused decode+query rows/s & % HBM peak; learned vs exhaustive encode GB/s"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--rows", type=int, default=120_000_000)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-secondary", action="store_true", help="skip the secondary workloads")
    ap.add_argument("--no-c5", action="store_true", help="skip the learned-selector workload")
    ap.add_argument("--only", default="", help="comma list of secondary workloads to run "
                    "(c1, c3, star, materialize, tile, c4, c5); default all")
    ap.add_argument("--c4-no-per-segment", action="store_true",
                    help="skip the batched section's one-launch-per-segment baseline (thousands of small "
                         "launches; leave it out of counter-collection runs)")
    ap.add_argument("--c4-segments", type=int, default=4096,
                    help="segments of the batched workload (fewer for counter-collection runs)")
    ap.add_argument("--c5-train-vectors", type=int, default=1 << 22,
                    help="training-set size of the learned selector (smaller for profiler runs)")
    return ap.parse_args()


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def time_kernel(fn, out, steps, warmup):
    """Time `steps` launches; returns (total_ms, [per-launch ms])."""
    import torch

    for _ in range(warmup):
        out.zero_()
        fn()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    ends = [torch.cuda.Event(enable_timing=True) for _ in range(steps)]
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0.record()
    for i in range(steps):
        out.zero_()
        starts[i].record()
        fn()
        ends[i].record()
    t1.record()
    torch.cuda.synchronize()
    return t0.elapsed_time(t1), [s.elapsed_time(e) for s, e in zip(starts, ends)]


def time_back_to_back(fn, steps, warmup):
    """Time `steps` launches issued back to back between two HIP events on
    the launching stream (no host work or other kernels between them);
    returns total_ms."""
    import torch

    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1)


def profiled_traffic(encoded_bytes):
    """HBM bytes per launch of the headline kernel from the committed PMC
    summary (FETCH_SIZE + WRITE_SIZE, calibrated on the copy kernel), when it
    was taken on this same workload (equal encoded bytes); else None."""
    f = ROOT / "profiles" / "q11_fused_120m_pmc.md"
    if not f.exists():
        return None
    kv = {}
    for line in f.read_text().splitlines():
        if ":" in line and line.strip().startswith("- `"):
            k, _, v = line.strip()[3:].partition("`:")
            kv[k.strip()] = v.strip().split()[0] if v.strip() else ""
    try:
        if int(kv["encoded_bytes"]) != encoded_bytes:
            return None
        return int(kv["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def copy_bandwidth(n_bytes=1 << 30, reps=20):
    """Read+write GB/s of the library's plain streaming copy kernel."""
    import torch

    import fastlanesgpu_amd as fls

    src = torch.ones(n_bytes // 4, dtype=torch.int32, device="cuda")
    dst = torch.empty_like(src)
    lib = fls.load()
    s = fls.current_stream()
    for _ in range(3):
        fls.call("fls_copy_bytes", src.data_ptr(), dst.data_ptr(), n_bytes, s)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        lib.fls_copy_bytes(src.data_ptr(), dst.data_ptr(), n_bytes, s)
    e1.record()
    torch.cuda.synchronize()
    ok = bool(torch.equal(src, dst))
    return 2 * n_bytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9, ok


def read_bandwidth(n_bytes=1 << 30, reps=20):
    """Read-only GB/s of the library's streaming checksum kernel (the
    achievable read bandwidth that the decode kernels are held against)."""
    import torch

    import fastlanesgpu_amd as fls

    src = torch.randint(0, 1 << 30, (n_bytes // 4,), dtype=torch.int32, device="cuda")
    out = torch.zeros(1, dtype=torch.int64, device="cuda")
    lib, s = fls.load(), fls.current_stream()
    for _ in range(3):
        fls.call("fls_read_bytes", src.data_ptr(), n_bytes, out.data_ptr(), s)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        lib.fls_read_bytes(src.data_ptr(), n_bytes, out.data_ptr(), s)
    e1.record()
    torch.cuda.synchronize()
    ok = int(out.item()) == int(src.long().sum().item())
    return n_bytes * reps / (e0.elapsed_time(e1) * 1e-3) / 1e9, ok


def bench_c1(steps, warmup, seed, cpu=True):
    """Plain bit-unpack: 2^24 values at bw 11, decode + write + checksum."""
    import numpy as np
    import torch

    import fastlanesgpu_amd as fls

    n, bw = 1 << 24, 11
    vals = fls.gen_c1(seed, n, bw)
    words = fls.pack_column(vals.cpu().numpy().view("uint32"), bw)
    d_in = torch.from_numpy(words.view("int32")).cuda()
    d_out = torch.empty(n, dtype=torch.int32, device="cuda")
    d_sum = fls.checksum_buffer()
    fn = lambda: fls.unpack_checksum(d_in, d_out, n, bw, d_sum)  # noqa: E731
    _, per = time_kernel(fn, d_sum, steps, warmup)
    ms = statistics.mean(per)
    b2b = time_back_to_back(fn, 10 * steps, warmup) / (10 * steps)
    ok = torch.equal(d_out, vals)
    bytes_moved = words.size * 4 + n * 4
    res = {"workload": "2^24 u32 uniform [0,2^11), bw 11: decode + write + checksum", "call_ms": ms,
           "values_per_s": n / (ms * 1e-3), "GBs_read_plus_write": bytes_moved / (ms * 1e-3) / 1e9,
           "kernel_ms_back_to_back": b2b, "GBs_read_plus_write_back_to_back": bytes_moved / (b2b * 1e-3) / 1e9,
           "frac_hbm_back_to_back": bytes_moved / (b2b * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "decoded_equal": bool(ok),
           "note": "one launch per call (the checksum is reduced inside it); call_ms brackets each call with "
                   "HIP events, kernel_ms_back_to_back is the mean of launches issued back to back"}
    if cpu:
        T = testlib()
        gpu_sum = int(d_sum.cpu().numpy().view(np.uint64)[0])
        plain = vals.cpu().numpy().view(np.uint32)
        res["cpu_baseline"] = cpu_leg(
            "values/s", n, lambda t: T.o_c1_encoded(words, n, bw, threads=t), 7,
            f"all {n} values (decode + write + checksum)", n, lambda: T.o_c1_encoded(words, n, bw, threads=1), 5,
            f"all {n} values",
            lambda a, b: a[1] == gpu_sum and b[1] == gpu_sum and np.array_equal(a[0], plain))
    return res


def bench_c3(steps, warmup, seed, cpu=True):
    """Fused decode + semi-joins + group-by (Q2.1 shape), 120M rows, 4 dimension tables."""
    import numpy as np
    import torch

    import fastlanesgpu_amd as fls

    n = 120_000_000
    fd = fls.gen_c3_fact(seed, n)
    fact = [fd[k].cpu().numpy().view("uint32") for k in ("datefk", "suppfk", "partfk", "revenue")]
    del fd
    enc = fls.encode_c3_table(*fact)
    del fact
    cols = [c.to_device() for c in enc]
    nbytes = sum(c.nbytes() for c in enc)
    dims = fls.gen_c3_dims(seed)
    ws = fls.q21_build(dims)
    out = torch.zeros(14000, dtype=torch.int64, device="cuda")
    fn = lambda: fls.query_q21(*cols, n, dims, ws, accumulate=True, d_out=out)  # noqa: E731
    _, per = time_kernel(fn, out, steps, warmup)
    res_c = out.clone()
    # the reference's hash tables, probed per row (Crystal primitives) and
    # through their occupancy bitmaps
    pht, keep = fls.q21_build_pht(dims, occupancy=False)
    fn2 = lambda: fls.query_q21_pht(*cols, n, pht, accumulate=True, d_out=out)  # noqa: E731
    _, per2 = time_kernel(fn2, out, max(3, steps // 5), 2)
    res_p = out.clone()
    pht_b, keep_b = fls.q21_build_pht(dims)
    fn3 = lambda: fls.query_q21_pht(*cols, n, pht_b, accumulate=True, d_out=out)  # noqa: E731
    _, per3 = time_kernel(fn3, out, steps, warmup)
    ms, ms2, ms3 = statistics.mean(per), statistics.mean(per2), statistics.mean(per3)
    out_b = torch.zeros_like(out)
    b2b = time_back_to_back(lambda: fls.query_q21(*cols, n, dims, ws, accumulate=True, d_out=out_b), 20, 2) / 20
    res = {"workload": "fused decode+probe+group-by (Q2.1 shape), 120M rows; fks at 12/16/20 bits, revenue 24",
           "encoded_bytes": nbytes, "kernel_ms_compact": ms, "rows_per_s_compact": n / (ms * 1e-3),
           "GBs_compact": nbytes / (ms * 1e-3) / 1e9, "frac_hbm_compact": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "kernel_ms_compact_back_to_back": b2b,
           "frac_hbm_compact_back_to_back": nbytes / (b2b * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "timing": "kernel_ms_*: HIP events around each launch (host work between launches included when the "
                     "host is slower than the kernel); *_back_to_back: the mean of 20 launches issued back to back",
           "kernel_ms_pht": ms2, "rows_per_s_pht": n / (ms2 * 1e-3),
           "kernel_ms_pht_bitmaps": ms3, "rows_per_s_pht_bitmaps": n / (ms3 * 1e-3),
           "frac_hbm_pht_bitmaps": nbytes / (ms3 * 1e-3) / 1e9 / HBM_PEAK_GBS,
           "groups": int((res_c[7000:] > 0).sum().item()),
           "paths_agree": bool(torch.equal(res_c, out) and torch.equal(res_c, res_p))}
    if cpu:
        T = testlib()
        dims_h = {k: v.cpu().numpy().view(np.uint32) for k, v in dims.items() if not k.startswith("n_")}
        sizes = {k: dims[k] for k in ("n_date", "n_supp", "n_part")}
        gpu = res_c.cpu().numpy()
        n_vec = (n + 1023) // 1024
        one = n_vec // 16
        full = lambda t: T.o_q21_encoded(enc, n, dims_h, sizes, threads=t)  # noqa: E731
        part = lambda t: T.o_q21_encoded(enc, n, dims_h, sizes, threads=t, v_end=one)  # noqa: E731
        ref_part = part(host_threads())
        res["cpu_baseline"] = cpu_leg(
            "rows/s", n, full, 5, f"full {n}-row table (decode 4 columns + dimension lookups + group-by)",
            one * 1024, lambda: part(1), 3, f"first {one} vectors ({one * 1024} rows, 1/16 of the table)",
            lambda a, b: np.array_equal(a[0], gpu[:7000]) and np.array_equal(a[1], gpu[7000:].view(np.uint64))
            and np.array_equal(b[0], ref_part[0]) and np.array_equal(b[1], ref_part[1]))
    return res


def bench_star(steps, warmup, seed, cpu=True):
    """Fused decode + star joins + group-by (Q3.1 and Q4.1 shapes), 120M rows."""
    import numpy as np
    import torch

    import fastlanesgpu_amd as fls

    n = 120_000_000
    fd = fls.gen_c3_fact(seed, n, with_customer=True)
    ex = fls.gen_star_extra(seed, n)
    f = [fd[k].cpu().numpy().view("uint32") for k in ("datefk", "custfk", "suppfk", "partfk", "revenue")]
    f.append(ex.pop("supplycost").cpu().numpy().view("uint32"))
    del fd
    enc = fls.encode_star_table(*f)
    del f
    cols = [c.to_device() for c in enc]
    dims = fls.gen_c3_dims(seed)
    dims["p_mfgr"] = ex.pop("p_mfgr")
    res = {"workload": "120M fact rows, dimensions 2556 dates / 600K customers / 40K suppliers / 800K parts; "
                       "keys at 12/20/16/20 bits, revenue 24, supplycost 16 (FOR)"}
    q31_cols = [cols[0], cols[1], cols[2], cols[4]]
    for name, run, build, used, n_slots in (
            ("q31", lambda ws, out: fls.query_q31(*q31_cols, n, dims, ws, accumulate=True, d_out=out),
             fls.q31_build, [0, 1, 2, 4], 4375),
            ("q41", lambda ws, out: fls.query_q41(*cols, n, dims, ws, accumulate=True, d_out=out),
             fls.q41_build, [0, 1, 2, 3, 4, 5], 175)):
        ws = build(dims)
        out = torch.zeros(2 * n_slots, dtype=torch.int64, device="cuda")
        _, per = time_kernel(lambda: run(ws, out), out, steps, warmup)
        ms = statistics.mean(per)
        out_b = torch.zeros_like(out)
        b2b = time_back_to_back(lambda: run(ws, out_b), 20, 2) / 20
        nbytes = sum(enc[i].nbytes() for i in used)
        one = torch.zeros_like(out)
        run(ws, one)
        total = torch.zeros_like(out)
        for _ in range(3):
            run(ws, total)
        res[name] = {"encoded_bytes": nbytes, "kernel_ms": ms, "rows_per_s": n / (ms * 1e-3),
                     "GBs": nbytes / (ms * 1e-3) / 1e9, "frac_hbm": nbytes / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "kernel_ms_back_to_back": b2b,
                     "frac_hbm_back_to_back": nbytes / (b2b * 1e-3) / 1e9 / HBM_PEAK_GBS,
                     "groups": int((one[n_slots:] > 0).sum().item()), "rows_qualifying": int(one[n_slots:].sum()),
                     "repeat_consistent": bool(torch.equal(total, 3 * one))}
        if cpu:
            T = testlib()
            dims_h = {k: v.cpu().numpy().view(np.uint32) for k, v in dims.items() if not k.startswith("n_")}
            sizes = dict(fls.C3_DIMS)
            gpu = one.cpu().numpy()
            n_vec = (n + 1023) // 1024
            part_v = n_vec // 16
            if name == "q31":
                q31c = [enc[i] for i in (0, 1, 2, 4)]
                run = lambda t, ve=None: T.o_q31_encoded(q31c, n, dims_h, sizes, threads=t, v_end=ve)  # noqa: E731
            else:
                run = lambda t, ve=None: T.o_q41_encoded(enc, n, dims_h, sizes, threads=t, v_end=ve)  # noqa: E731
            ref_part = run(host_threads(), part_v)
            res[name]["cpu_baseline"] = cpu_leg(
                "rows/s", n, run, 5, f"full {n}-row table", part_v * 1024, lambda: run(1, part_v), 3,
                f"first {part_v} vectors ({part_v * 1024} rows, 1/16 of the table)",
                lambda a, b: np.array_equal(a[0], gpu[:n_slots]) and np.array_equal(
                    a[1], gpu[n_slots:].view(np.uint64)) and np.array_equal(b[0], ref_part[0]))
    return res


def bench_materialize(steps, warmup, seed):
    """Decode-to-materialize: 120M-row delta orderdate to plain values in row
    order (untransposed through LDS) and in stored order; transpose_o alone."""
    import torch

    import fastlanesgpu_amd as fls

    n = 120_000_000
    od = fls.gen_c2(seed, n)[0]
    w, vm, bs = fls.delta_encode_device(od, n)
    col = fls.Column(fls.ENC_DELTA, n, w, vmeta=vm, bases=bs)
    enc_bytes = col.nbytes()
    n_vec = (n + 1023) // 1024
    res = {"workload": "120M-row orderdate column (sorted dates), FastLanes delta; output 4 B per row",
           "encoded_bytes": enc_bytes}
    out = torch.empty(n_vec * 1024, dtype=torch.int32, device="cuda")
    for name, order in (("rows", fls.ORDER_ROWS), ("stored", fls.ORDER_STORED)):
        _, per = time_kernel(lambda: fls.decode_column(col, n, order, out=out), out, steps, warmup)
        ms = statistics.mean(per)
        moved = enc_bytes + 4 * (n if order == fls.ORDER_ROWS else n_vec * 1024)
        res[name] = {"kernel_ms": ms, "rows_per_s": n / (ms * 1e-3), "GBs_read_plus_write": moved / (ms * 1e-3) / 1e9,
                     "frac_hbm": moved / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    res["rows_equal_generated"] = bool(torch.equal(fls.decode_column(col, n, fls.ORDER_ROWS, out=out), od))
    src = out.clone()
    dst = torch.empty_like(out)
    _, per = time_kernel(lambda: fls.transpose_vectors(src, n_vec, inverse=True, out=dst), dst, steps, warmup)
    ms = statistics.mean(per)
    res["transpose_o"] = {"kernel_ms": ms, "GBs_read_plus_write": 8 * n_vec * 1024 / (ms * 1e-3) / 1e9,
                          "frac_hbm": 8 * n_vec * 1024 / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS}
    return res


def bench_tile_vs_fastlanes(steps, warmup, seed):
    """Decode to memory of the same 2^26-value columns in FastLanes and in the
    tile-based GPU-FOR / GPU-DFOR / GPU-RFOR formats (the competitor baseline
    the reference carries): bit-packing at widths 3 and 11 (uniform), delta
    (the reference delta bench's column 7 i) and RLE (geometric runs, mean 8)."""
    import numpy as np
    import torch

    import fastlanesgpu_amd as fls

    n = 1 << 26
    n_vec = n // 1024
    rng = np.random.default_rng(seed)
    out = torch.empty(n, dtype=torch.int32, device="cuda")

    def timed(fn):
        _, per = time_kernel(fn, out, steps, warmup)
        return statistics.mean(per)

    def entry(ms, enc_bytes):
        return {"kernel_ms": ms, "values_per_s": n / (ms * 1e-3), "encoded_bytes": enc_bytes,
                "bits_per_value": 8 * enc_bytes / n,
                "GBs_read_plus_write": (enc_bytes + 4 * n) / (ms * 1e-3) / 1e9}

    def tile(x, kind):
        col = fls.tile_encode(x, kind)
        dcol = col.to_device()
        ms = timed(lambda: fls.tile_decode(dcol, out=out))
        ok = bool(torch.equal(out, torch.from_numpy(x.view(np.int32)).cuda()))
        return dict(entry(ms, col.nbytes()), decoded_equal=ok)

    res = {"workload": "2^26 uint32 values per column, decode to memory (4 B per value written)"}
    for bw in (3, 11):
        x = rng.integers(0, 1 << bw, n).astype(np.uint32)
        d_in = torch.from_numpy(fls.pack_column(x, bw).view(np.int32)).cuda()
        ms = timed(lambda: fls.unpack_global(d_in, out, n_vec, bw))
        ok = bool(torch.equal(out, torch.from_numpy(x.view(np.int32)).cuda()))
        res[f"bitpack_bw{bw}"] = {"fastlanes": dict(entry(ms, d_in.numel() * 4), decoded_equal=ok),
                                  "gpu_for": tile(x, "for")}
    x = (7 * np.arange(n, dtype=np.uint64)).astype(np.uint32)
    w, vm, bs = fls.delta_encode_device(torch.from_numpy(x.view(np.int32)).cuda(), n)
    col = fls.Column(fls.ENC_DELTA, n, w, vmeta=vm, bases=bs)
    ms = timed(lambda: fls.decode_column(col, n, fls.ORDER_ROWS, out=out))
    ok = bool(torch.equal(out, torch.from_numpy(x.view(np.int32)).cuda()))
    res["delta_7i"] = {"fastlanes_row_order": dict(entry(ms, col.nbytes()), decoded_equal=ok),
                       "gpu_dfor": tile(x, "dfor")}
    lens = rng.geometric(1 / 8, n)
    x = np.repeat(rng.integers(0, 1 << 16, n), lens)[:n].astype(np.uint32)
    w, vm, rm, bs = fls.rle_encode_device(torch.from_numpy(x.view(np.int32)).cuda(), n)
    col = fls.Column(fls.ENC_RLE, n, w, vmeta=vm, bases=bs, aux=rm)
    ms = timed(lambda: fls.decode_column(col, n, fls.ORDER_ROWS, out=out))
    ok = bool(torch.equal(out, torch.from_numpy(x.view(np.int32)).cuda()))
    res["rle_mean8"] = {"fastlanes_row_order": dict(entry(ms, col.nbytes()), decoded_equal=ok),
                        "gpu_rfor": tile(x, "rfor")}
    return res


def bench_c4(steps, seed, cpu=True, per_segment=True, n_seg=4096):
    """4096 uneven segments (bitpack/delta/RLE) decoded + range-filtered:
    one graph-replayed persistent launch vs one launch per segment."""
    import numpy as np
    import torch

    import fastlanesgpu_amd as fls

    t = time.perf_counter()
    b = fls.batch_c4(seed, n_seg, 10, 24)
    torch.cuda.synchronize()
    build_s = time.perf_counter() - t
    info = b.info()
    d = torch.zeros(n_seg, dtype=torch.int64, device="cuda")
    torch.cuda.synchronize()
    t = time.perf_counter()
    best, tune_best_ms = b.autotune(d, force=True)  # every candidate timed (no cache)
    torch.cuda.synchronize()
    tune_cost_ms = (time.perf_counter() - t) * 1e3
    ref = None

    def timed(fn, reps):
        nonlocal ref
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        got = d.clone()
        same = True if ref is None else bool(torch.equal(ref, got))
        ref = got if ref is None else ref
        return e0.elapsed_time(e1) / reps, same

    graph_ms, _ = timed(lambda: b.graph_launch(best, d), max(5, steps))
    run_ms, same_run = timed(lambda: b.run(best, d), max(5, steps))
    seg_ms, same_seg = timed(lambda: b.run_per_segment(best, d), 3) if per_segment else (None, True)
    nbytes = info["encoded_bytes"]
    cpu_res = None
    if cpu:
        T = testlib()
        gpu_counts = ref.cpu().numpy().view(np.uint64)
        all_ids = list(range(0, n_seg, 8))
        one_ids = list(range(0, n_seg, 64))
        segs = {s: b.segment_host(s) for s in all_ids}
        work = lambda ids: sum(int(segs[s][1]) for s in ids)  # noqa: E731
        run = lambda ids, t: T.o_c4_encoded([segs[s] for s in ids], threads=t)  # noqa: E731
        cpu_res = cpu_leg(
            "values/s", work(all_ids), lambda t: run(all_ids, t), 5,
            f"every 8th segment ({len(all_ids)} segments, {work(all_ids)} values, decode + range count)",
            work(one_ids), lambda: run(one_ids, 1), 3,
            f"every 64th segment ({len(one_ids)} segments, {work(one_ids)} values)",
            lambda a, c: np.array_equal(a, gpu_counts[all_ids]) and np.array_equal(c, gpu_counts[one_ids]))
        del segs
    b.close()
    return {"cpu_baseline": cpu_res,
            "workload": f"{n_seg} segments, lengths log-uniform by octave in [2^10,2^24), bw U{{1..32}}, "
                        "40/30/30 bitpack/delta/RLE, middle-half range filter, count per segment",
            "n_values": info["n_values"], "encoded_bytes": nbytes,
            "segments_by_encoding": [info["n_bitpack"], info["n_delta"], info["n_rle"]],
            "build_s": build_s,
            "tune": {"tile_vectors": best.tile_vectors, "block_threads": best.block_threads,
                     "schedule": ["static", "dynamic"][best.schedule],
                     "class_tiles": {f"{['bitpack', 'delta', 'rle', 'mixed'][c // 2]}_{['narrow', 'wide'][c % 2]}":
                                     1 << ((best.class_tiles >> (4 * c)) & 15)
                                     for c in range(8) if (best.class_tiles >> (4 * c)) & 15},
                     "tuning_cost_ms": tune_cost_ms, "best_candidate_ms": tune_best_ms},
            "graph_ms": graph_ms, "launch_ms": run_ms, "per_segment_ms": seg_ms,
            "values_per_s": info["n_values"] / (graph_ms * 1e-3),
            "encoded_GBs": nbytes / (graph_ms * 1e-3) / 1e9,
            "frac_hbm": nbytes / (graph_ms * 1e-3) / 1e9 / HBM_PEAK_GBS,
            "speedup_vs_per_segment": seg_ms / graph_ms if seg_ms else None,
            "counts_agree": bool(same_run and same_seg)}


def bench_c5(seed, n_rows, n_train=1 << 22, reps=3, cpu=True):
    """Learned encoding selector: train on 2^22 synthetic vectors (four
    families, exhaustive labels) with the library's own MFMA kernels, then
    encode a held-out 120M-row Q1.1 column set three ways -- learned (with the
    exhaustive fallback below confidence 0.9), exhaustive from statistics,
    exhaustive by trying all four encoders -- and run Q1.1 over the learned
    output."""
    import torch

    import fastlanesgpu_amd as fls
    from fastlanesgpu_amd import selector as sel

    def ev():
        return torch.cuda.Event(enable_timing=True)

    # ---- training set + training
    e0, e1 = ev(), ev()
    e0.record()
    feat, label, _ = sel.build_training_set(seed + 100, n_train)
    e1.record()
    torch.cuda.synchronize()
    gen_ms = e0.elapsed_time(e1)
    mlp = sel.Mlp(seed)
    epochs, batch = 3, 4096
    t = time.perf_counter()
    losses = mlp.train(feat, label, epochs=epochs, batch=batch, lr=3e-3, seed=seed)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t
    steps = epochs * ((n_train + batch - 1) // batch)
    # MLP kernel timings (forward over the whole training set; one gradient step)
    cls = torch.empty(n_train, dtype=torch.uint8, device="cuda")
    conf = torch.empty(n_train, dtype=torch.float32, device="cuda")
    fwd = lambda: fls.call("fls_mlp_forward", mlp.wb.data_ptr(), feat.data_ptr(), n_train,  # noqa: E731
                           cls.data_ptr(), conf.data_ptr(), None, fls.current_stream())
    fwd()
    e0, e1 = ev(), ev()
    e0.record()
    for _ in range(reps):
        fwd()
    e1.record()
    torch.cuda.synchronize()
    fwd_ms = e0.elapsed_time(e1) / reps
    train_acc = float((cls // 10 == label // 10).float().mean().item())
    fb, lb = feat[:batch].contiguous(), label[:batch].contiguous()
    mlp.accumulate_grad(fb, lb, batch)
    e0, e1 = ev(), ev()
    e0.record()
    for _ in range(20):
        mlp.accumulate_grad(fb, lb, batch)
    e1.record()
    torch.cuda.synchronize()
    grad_ms = e0.elapsed_time(e1) / 20
    mlp.grad.zero_()
    flop_fwd = 2 * (16 * 64 + 64 * 64 + 64 * 40)  # 15,360 per vector
    del feat, label

    # ---- held-out column set: the Q1.1 table, values resident in HBM
    cols = fls.gen_c2(seed + 7, n_rows)
    n = n_rows
    in_bytes = 4 * n * len(cols)
    learned = sel.LearnedSelector(mlp, conf_min=0.9)
    ws_l = [sel.MixedWorkspace(n) for _ in cols]
    ws_x = [sel.MixedWorkspace(n) for _ in cols]
    ws_t = [sel.MixedWorkspace(n) for _ in cols]
    tbuf = sel.trial_buffer(n)

    def run(fn):
        fn()
        torch.cuda.synchronize()
        a, b = ev(), ev()
        a.record()
        for _ in range(reps):
            fn()
        b.record()
        torch.cuda.synchronize()
        return a.elapsed_time(b) / reps

    def enc_learned():
        for c, w in zip(cols, ws_l):
            w.fallbacks.zero_()
            learned.encode(c, n, w)

    def enc_stats():
        for c, w in zip(cols, ws_x):
            sel.mixed_encode_exhaustive(c, n, w)

    def enc_trial():
        for c, w in zip(cols, ws_t):
            sel.mixed_encode_trial(c, n, w, buf=tbuf)

    ms_l, ms_x, ms_t = run(enc_learned), run(enc_stats), run(enc_trial)
    size = lambda ws: sum(int(w.total.item()) * 4 + 16 * sel.n_vectors(n) for w in ws)  # noqa: E731
    s_l, s_x, s_t = size(ws_l), size(ws_x), size(ws_t)
    nv = sel.n_vectors(n)
    fallbacks = sum(int(w.fallbacks.item()) for w in ws_l)
    per_col = []
    for c, wl, wx in zip(("orderdate", "quantity", "discount", "price"), ws_l, ws_x):
        el = torch.bincount((wl.cls[:nv] // 10).long(), minlength=4).tolist()
        per_col.append({"column": c, "learned_bytes": int(wl.total.item()) * 4 + 16 * nv,
                        "exhaustive_bytes": int(wx.total.item()) * 4 + 16 * nv, "learned_encodings": el})
    # ---- Q1.1 over the learned output vs the same query on the raw values
    mixed = [w.column(n) for w in ws_l]
    rev, cnt = fls.query_q11(*mixed, n)
    od, qt, dc, pr = (c.long() for c in cols)
    m = (od > 19930000) & (od < 19940000) & (qt < 25) & (dc >= 1) & (dc <= 3)
    ref = (int((pr * dc * m).sum().item()), int(m.sum().item()))
    del od, qt, dc, pr, m
    q_out = torch.zeros(2, dtype=torch.int64, device="cuda")
    q_ms = run(lambda: fls.query_q11(*mixed, n, d_out=q_out, sync=False))
    cpu_res = None
    if cpu:
        import numpy as np

        T = testlib()
        k_all, k_one = 2048, 256  # vectors per column
        sample = [c[:k_all * 1024].cpu().numpy().view(np.uint32) for c in cols]
        gpu_cls = [w.cls[:k_all].cpu().numpy() for w in ws_x]
        vec_all = np.concatenate(sample)
        vec_one = np.concatenate([x[:k_one * 1024] for x in sample])
        cpu_res = cpu_leg(
            "GB/s", 4 * vec_all.size / 1e9, lambda t: T.o_exhaustive_classes(vec_all, threads=t), 3,
            f"exhaustive search (all four encoders per vector) over the first {k_all} vectors of each of the "
            f"{len(cols)} held-out columns ({vec_all.size // 1024} vectors)",
            4 * vec_one.size / 1e9, lambda: T.o_exhaustive_classes(vec_one, threads=1), 3,
            f"the first {k_one} vectors of each column ({vec_one.size // 1024} vectors)",
            lambda a, b: np.array_equal(a, np.concatenate(gpu_cls))
            and np.array_equal(b, np.concatenate([g[:k_one] for g in gpu_cls])))
    return {"cpu_baseline_exhaustive": cpu_res,
            
        "workload": f"train on {n_train} vectors (uniform / Zipf 1.1 / sorted+N(0,16) / runs mean 32), "
                    f"encode held-out Q1.1 column set {len(cols)} x {n} u32",
        "train": {"gen_features_labels_ms": gen_ms, "epochs": epochs, "batch": batch, "adam_steps": steps,
                  "train_s": train_s, "loss_per_epoch": losses, "train_encoding_accuracy": train_acc,
                  "fwd_ms_all_vectors": fwd_ms, "fwd_TFLOPs": flop_fwd * n_train / (fwd_ms * 1e-3) / 1e12,
                  "grad_step_ms_batch": grad_ms,
                  "grad_TFLOPs": 3 * flop_fwd * batch / (grad_ms * 1e-3) / 1e12,
                  "mfma_peak_TFLOPs_bf16": 2500.0},
        "encode_GBs": {"learned": in_bytes / (ms_l * 1e-3) / 1e9,
                       "exhaustive_stats": in_bytes / (ms_x * 1e-3) / 1e9,
                       "exhaustive_trial": in_bytes / (ms_t * 1e-3) / 1e9},
        "encode_ms": {"learned": ms_l, "exhaustive_stats": ms_x, "exhaustive_trial": ms_t},
        "speedup_vs_trial": ms_t / ms_l, "speedup_vs_stats": ms_x / ms_l,
        "bytes": {"learned": s_l, "exhaustive_stats": s_x, "exhaustive_trial": s_t, "raw": in_bytes},
        "size_error": s_l / s_x - 1, "trial_equals_stats": s_t == s_x,
        "fallback_rate": fallbacks / (nv * len(cols)), "per_column": per_col,
        "q11_over_learned": {"revenue": rev, "count": cnt, "equal_to_raw_values": (rev, cnt) == ref,
                             "ms": q_ms, "rows_per_s": n / (q_ms * 1e-3)},
    }


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads():
    return len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)


def testlib():
    """The oracle's ctypes wrappers (tests/fls_testlib.py): the CPU baselines
    are the oracle's scalar decode + query."""
    sys.path.insert(0, str(ROOT / "tests"))
    import fls_testlib as T

    return T


def median_run(fn, runs):
    times, res = [], None
    for _ in range(runs):
        t = time.perf_counter()
        res = fn()
        times.append(time.perf_counter() - t)
    return statistics.median(times), res


def cpu_leg(unit, work_all, run_all, runs_all, sample_all, work_one, run_one, runs_one, sample_one, matches):
    """Oracle baseline of one workload on all host threads and on one thread.
    run_all(threads) / run_one() return results; matches(res_all, res_one)
    says whether they equal the GPU's."""
    threads = host_threads()
    run_all(threads)  # warm-up
    t_all, r_all = median_run(lambda: run_all(threads), runs_all)
    t_one, r_one = median_run(run_one, runs_one)
    return {"value": work_all / t_all, "unit": unit, "cores": threads, "kind": "port",
            "sample": f"{sample_all}, median of {runs_all} runs on {threads} threads after one warm-up",
            "single_core_value": work_one / t_one,
            "single_core_sample": f"{sample_one}, median of {runs_one} runs on 1 thread",
            "result_matches_gpu": bool(matches(r_all, r_one))}


def cpu_baseline(enc, n_rows, gpu_result):
    """Oracle decode + query on the host cores: all threads (one warm-up, the
    median of 7 runs over the full table), then one thread (median of 5 runs
    over the full table) -- about 10 s of CPU work in all."""
    sys.path.insert(0, str(ROOT / "tests"))
    import fls_testlib as T

    threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    n_vec = (n_rows + 1023) // 1024

    def median_time(runs, nthreads, first=0, count=None):
        times, res = [], None
        for _ in range(runs):
            t = time.perf_counter()
            res = T.o_q11_encoded(enc, n_rows, first, n_vec if count is None else count, threads=nthreads)
            times.append(time.perf_counter() - t)
        return statistics.median(times), res

    T.o_q11_encoded(enc, n_rows, 0, min(n_vec, 1024), threads=threads)  # warm-up
    t_all, res = median_time(7, threads)
    t_one, res_one = median_time(5, 1)
    return {
        "value": n_rows / t_all,
        "unit": "rows/s",
        "cores": threads,
        "kind": "port",
        "sample": f"full {n_rows}-row table, median of 7 runs on {threads} threads after one warm-up "
                  f"(scalar per-lane decode + predicate/sum loop)",
        "cpu_model": cpu_model(),
        "single_core_rows_per_s": n_rows / t_one,
        "single_core_sample": f"full {n_rows}-row table, median of 5 runs on 1 thread",
        "result_matches_gpu": res == gpu_result and res_one == gpu_result,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "the fused query shards by 1024-row vector ranges; this benchmark measures one GPU",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch

    import fastlanesgpu_amd as fls

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    n = args.rows
    log(f"[bench] generating {n} rows on the GPU")
    dcols = fls.gen_c2(args.seed, n)
    torch.cuda.synchronize()
    cols = [c.cpu().numpy().view(np.uint32) for c in dcols]
    del dcols
    log("[bench] encoding (host encoders)")
    t = time.perf_counter()
    enc = fls.encode_c2_table(*cols)
    enc_s = time.perf_counter() - t
    raw_bytes = 4 * 4 * n
    del cols
    dev = [c.to_device() for c in enc]
    col_bytes = [c.nbytes() for c in enc]
    total_bytes = sum(col_bytes)
    torch.cuda.synchronize()

    out = torch.zeros(2, dtype=torch.int64, device="cuda")
    full = lambda: fls.query_q11(*dev, n, accumulate=True, d_out=out, sync=False)  # noqa: E731
    pred = lambda: fls.query_q11(*dev, n, accumulate=True, pred_load=True, d_out=out, sync=False)  # noqa: E731

    # one launch for the result, then K launches back to back (the kernel adds
    # into `out`, so after K more launches it must hold K + 1 times the result)
    out.zero_()
    full()
    result = tuple(int(x) for x in out.cpu().tolist())
    log(f"[bench] timing full-scan kernel: {args.steps} back-to-back launches after {args.warmup} warm-up")
    out.zero_()
    total_ms = time_back_to_back(full, args.steps, args.warmup)
    repeat_ok = tuple(int(x) for x in out.cpu().tolist()) == tuple((args.steps + args.warmup) * r for r in result)
    k_ms = total_ms / args.steps
    _, per = time_kernel(full, out, max(5, args.steps // 5), 2)  # single launches, each between two events
    log("[bench] timing predicated-load variant")
    out.zero_()
    pred()
    result_p = tuple(int(x) for x in out.cpu().tolist())
    total_ms_p = time_back_to_back(pred, args.steps, args.warmup)
    k_ms_p = total_ms_p / args.steps

    achieved = total_bytes / (k_ms * 1e-3) / 1e9
    traffic = profiled_traffic(total_bytes)
    log("[bench] copy-kernel bandwidth")
    copy_gbs, copy_ok = copy_bandwidth()
    read_gbs, read_ok = read_bandwidth()
    line = {
        "metric": METRIC,
        "value": n * args.steps / (total_ms * 1e-3),
        "unit": "rows/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": total_ms / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": "fused decode+filter+sum (Q1.1 shape), 120M-row fact table: orderdate FastLanes delta, "
                        "quantity bw6, discount bw4, price bw24, unified transposed layout",
            "rows": n,
            "encoded_bytes": total_bytes,
            "encoded_bytes_per_column": dict(zip(["orderdate", "quantity", "discount", "price"], col_bytes)),
            "bytes_per_row": total_bytes / n,
            "raw_bytes": raw_bytes,
            "parallelism": "single GPU",
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": "profiles/q11_fused_120m_pmc.md (FETCH_SIZE x2.000 + WRITE_SIZE x1.000, "
                              "factors from the copy kernel)" if traffic else None,
            "kernel": "q11_fast_kernel<6,4,24,full scan>",
            "kernel_ms": k_ms,
            "kernel_ms_single_launch": statistics.mean(per),
            "timing": "achieved = encoded bytes / (back-to-back total / launches), HIP events on the "
                      "launching stream; kernel_ms_single_launch brackets each launch alone",
            "copy_kernel_GBs": copy_gbs,
            "copy_kernel_ok": copy_ok,
            "frac_of_copy": achieved / copy_gbs,
            "read_only_GBs": read_gbs,
            "read_only_ok": read_ok,
            "frac_of_read_only": achieved / read_gbs,
        },
        "result": {"revenue": result[0], "rows": result[1], "repeat_launches_consistent": repeat_ok},
        "pred_load": {
            "rows_per_s": n * args.steps / (total_ms_p * 1e-3),
            "kernel_ms": k_ms_p,
            "encoded_byte_rate_GBs": total_bytes / (k_ms_p * 1e-3) / 1e9,
            "note": "skips later columns of waves with no qualifying row; reads fewer bytes than the full scan",
            "result_matches": result_p == result,
        },
        "host_encode_s": enc_s,
    }
    if not args.no_cpu:
        log("[bench] CPU baseline")
        line["cpu_baseline"] = cpu_baseline(enc, n, result)
    del dev, enc
    if not args.no_secondary:
        only = {x.strip() for x in args.only.split(",") if x.strip()}
        want = lambda name: not only or name in only  # noqa: E731
        line["secondary"] = {}
        if want("c1"):
            log("[bench] secondary: plain unpack")
            line["secondary"]["c1_unpack"] = bench_c1(args.steps, args.warmup, args.seed, cpu=not args.no_cpu)
        if want("c3"):
            log("[bench] secondary: probe + group-by")
            line["secondary"]["c3_probe_groupby"] = bench_c3(max(5, args.steps // 2), 2, args.seed,
                                                             cpu=not args.no_cpu)
        if want("star"):
            log("[bench] secondary: star joins")
            line["secondary"]["star_joins"] = bench_star(max(5, args.steps // 2), 2, args.seed, cpu=not args.no_cpu)
        if want("materialize"):
            log("[bench] secondary: materialize")
            line["secondary"]["materialize"] = bench_materialize(max(5, args.steps // 2), 2, args.seed)
        if want("tile"):
            log("[bench] secondary: tile-based formats vs FastLanes")
            line["secondary"]["tile_vs_fastlanes"] = bench_tile_vs_fastlanes(max(5, args.steps // 2), 2, args.seed)
        if want("c4"):
            log("[bench] secondary: batched segments")
            line["secondary"]["c4_batched"] = bench_c4(max(5, args.steps // 5), args.seed, cpu=not args.no_cpu,
                                                       per_segment=not args.c4_no_per_segment,
                                                       n_seg=args.c4_segments)
        if want("c5") and not args.no_c5:
            log("[bench] secondary: learned encoding selector")
            line["secondary"]["c5_learned_selector"] = bench_c5(args.seed, n, n_train=args.c5_train_vectors,
                                                                cpu=not args.no_cpu)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
