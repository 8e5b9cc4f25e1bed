"""Batched decode + range filter over uneven segments (one persistent launch,
per-segment baseline, hipGraph replay, autotuner) against the oracle."""
import os

import numpy as np
import pytest

import fastlanesgpu_amd as fls
import fls_testlib as T

pytestmark = pytest.mark.gpu


def counts_of(t):
    return T.from_dev(t, np.uint64)


def test_segment_generator_matches_oracle():
    for s in (0, 1, 2, 3, 5, 8, 13):
        enc, bw, want = T.o_c4_values(31, s, 10, 16)
        assert fls.c4_segment_params(31, s, 10, 16)[:2] == (enc, bw)
        got = fls.gen_c4_segment(31, s, 10, 16)
        np.testing.assert_array_equal(T.from_dev(got, np.uint32), want)


@pytest.mark.parametrize("tile", [1, 8, 64, 512])
def test_small_batch_all_paths(tile):
    torch = T.torch_cuda()
    n_seg, lmin, lmax = 96, 10, 16
    want, lo, hi = T.o_c4_counts(7, n_seg, lmin, lmax)
    b = fls.batch_c4(7, n_seg, lmin, lmax)
    info = b.info()
    assert info["n_segments"] == n_seg and info["n_bitpack"] + info["n_delta"] + info["n_rle"] == n_seg
    assert min(info["n_bitpack"], info["n_delta"], info["n_rle"]) > 0
    for s in range(n_seg):
        si = b.segment(s)
        assert (si["lo"], si["hi"]) == (lo[s], hi[s])
    d = torch.zeros(n_seg, dtype=torch.int64, device="cuda")
    for block in (256, 512):
        for sched in (fls.SCHED_STATIC, fls.SCHED_DYNAMIC):
            t = fls.tune(tile, block, sched)
            d.fill_(-1)
            b.run(t, d)
            np.testing.assert_array_equal(counts_of(d), want)
            d.fill_(-1)
            b.run_per_segment(t, d)
            np.testing.assert_array_equal(counts_of(d), want)


def test_graph_replays():
    torch = T.torch_cuda()
    n_seg = 40
    want, _, _ = T.o_c4_counts(9, n_seg, 10, 17)
    b = fls.batch_c4(9, n_seg, 10, 17)
    d = torch.zeros(n_seg, dtype=torch.int64, device="cuda")
    t = fls.tune(32, 256, fls.SCHED_DYNAMIC)
    for rep in range(4):
        d.fill_(-7)  # eager work between replays; the graph must re-zero
        torch.cuda.synchronize()
        b.graph_launch(t, d)
        np.testing.assert_array_equal(counts_of(d), want, err_msg=f"replay {rep}")
    t2 = fls.tune(8, 512, fls.SCHED_STATIC)  # new shape: re-captured
    b.graph_launch(t2, d)
    np.testing.assert_array_equal(counts_of(d), want)


def test_autotune_and_cache(tmp_path):
    torch = T.torch_cuda()
    n_seg = 48
    want, _, _ = T.o_c4_counts(11, n_seg, 10, 16)
    b = fls.batch_c4(11, n_seg, 10, 16)
    d = torch.zeros(n_seg, dtype=torch.int64, device="cuda")
    cache = str(tmp_path / "tune.txt")
    best, ms = b.autotune(d, cache_path=cache)
    assert ms > 0 and best.tile_vectors > 0 and os.path.exists(cache)
    best2, ms2 = b.autotune(d, cache_path=cache)
    assert ms2 == -1.0 and (best2.tile_vectors, best2.block_threads, best2.schedule, best2.class_tiles) == \
        (best.tile_vectors, best.block_threads, best.schedule, best.class_tiles)
    b.run(best, d)
    np.testing.assert_array_equal(counts_of(d), want)


def test_caller_described_segments():
    torch = T.torch_cuda()
    rng = np.random.default_rng(3)
    segs, want = [], []
    specs = [("bitpack", 5000, 9), ("delta", 70_000, 0), ("rle", 33_333, 14), ("bitpack", 1, 3), ("rle", 1024, 1),
             ("mixed", 50_000, 13), ("rle_aux", 9000, 20), ("mixed", 700, 31), ("bitpack_t", 3 * 1024 + 777, 13),
             ("bitpack_for", 20_000 + 3, 10), ("rle_long", 40_000 + 9, 12), ("delta_steps", 64 * 1024, 0),
             ("bitpack", 8 * 1024, 0)]
    from fastlanesgpu_amd import selector as sel

    keep = []
    for kind, n, bw in specs:
        if kind == "mixed":
            x = np.repeat(rng.integers(0, 2**bw, n // 5 + 1), 5)[:n].astype(np.uint32)
            x[: n // 3] = np.sort(x[: n // 3])
            ws = sel.MixedWorkspace(n)
            nv = sel.n_vectors(n)
            ws.cls[:nv] = torch.from_numpy(rng.integers(0, 40, nv).astype(np.uint8)).cuda()
            d_x = T.to_dev(x)
            sel.mixed_encode_classes(d_x, n, ws)
            col = ws.column(n)
            keep += [ws, d_x]
            rmeta = None
        elif kind == "rle_aux":
            starts = rng.random(n) < 0.125
            starts[0] = True
            x = rng.integers(0, 2**bw, n).astype(np.uint32)[np.cumsum(starts) - 1]
            w, vm, rm, bs = fls.rle_encode_device(T.to_dev(x), n)
            col = fls.Column(fls.ENC_RLE, n, w, vmeta=vm, bases=bs, aux=rm)
            rmeta = None
        elif kind == "bitpack_t":  # bit-packed after transpose_i: its partial vector is masked in that order
            x = rng.integers(0, 2**bw, n).astype(np.uint32)
            col = fls.Column(fls.ENC_BITPACK, n, fls.pack_column_device(T.to_dev(x), n, bw, transposed=True), bw=bw,
                             transposed=True)
            rmeta = None
        elif kind == "bitpack":
            x = rng.integers(0, 2**bw, n).astype(np.uint32)
            col = fls.Column(fls.ENC_BITPACK, n, fls.pack_column_device(T.to_dev(x), n, bw), bw=bw)
            rmeta = None
        elif kind == "bitpack_for":  # frame of reference: values = packed + base
            base = 5_000_000
            x = (rng.integers(0, 2**bw, n) + base).astype(np.uint32)
            col = fls.Column(fls.ENC_BITPACK, n, fls.pack_column_device(T.to_dev(x), n, bw, base=base), bw=bw,
                             base=base)
            rmeta = None
        elif kind == "rle_long":  # runs of mean 300: many lanes lie inside one run (index width 0)
            starts = rng.random(n) < 1 / 300
            starts[0] = True
            x = rng.integers(0, 2**bw, n).astype(np.uint32)[np.cumsum(starts) - 1]
            w, vm, rm, bs = fls.rle_encode_device(T.to_dev(x), n)
            col = fls.Column(fls.ENC_RLE, n, w, vmeta=vm, bases=bs, aux=rm)
            rmeta = None
        elif kind == "delta_steps":  # sorted, with a step size that changes every 1000 rows: widths vary per vector
            steps = (rng.integers(0, 2, n) << rng.integers(0, 20, n // 1000 + 1).repeat(1000)[:n]).astype(np.uint64)
            x = np.cumsum(steps).astype(np.uint32)
            w, vm, bs = fls.delta_encode_device(T.to_dev(x), n)
            col = fls.Column(fls.ENC_DELTA, n, w, vmeta=vm, bases=bs)
            assert len(set((T.from_dev(vm, np.uint64) & 0xFF).tolist())) > 3
            rmeta = None
        elif kind == "delta":
            x = np.sort(rng.integers(0, 2**30, n)).astype(np.uint32)
            w, vm, bs = fls.delta_encode_device(T.to_dev(x), n)
            col = fls.Column(fls.ENC_DELTA, n, w, vmeta=vm, bases=bs)
            rmeta = None
        else:
            starts = rng.random(n) < 0.125
            starts[0] = True
            x = rng.integers(0, 2**bw, n).astype(np.uint32)[np.cumsum(starts) - 1]
            w, vm, rm, bs = fls.rle_encode_device(T.to_dev(x), n)
            col = fls.Column(fls.ENC_RLE, n, w, vmeta=vm, bases=bs)
            rmeta = rm
        lo, hi = int(np.percentile(x, 20)), int(np.percentile(x, 70))
        segs.append({"col": col, "rmeta": rmeta, "n_values": n, "lo": lo, "hi": hi})
        want.append(int(((x >= lo) & (x <= hi)).sum()))
    b = fls.batch_from_segments(segs)
    d = torch.zeros(len(segs), dtype=torch.int64, device="cuda")
    for tile, block, classes in ((4, 256, None), (1, 512, None), (64, 256, None), (512, 512, None),
                                 (8, 256, {0: 2, 1: 64, 2: 16, 4: 4, 5: 256, 6: 32})):
        d.fill_(-1)
        b.run(fls.tune(tile, block, fls.SCHED_DYNAMIC, class_tiles=classes), d)
        np.testing.assert_array_equal(counts_of(d), np.array(want, np.uint64), err_msg=f"tile {tile} block {block}")
        d.fill_(-1)
        b.run_per_segment(fls.tune(tile, block, fls.SCHED_STATIC, class_tiles=classes), d)
        np.testing.assert_array_equal(counts_of(d), np.array(want, np.uint64), err_msg=f"per segment, tile {tile}")


def test_full_batch_4096_segments():
    torch = T.torch_cuda()
    n_seg = 4096
    b = fls.batch_c4(1, n_seg, 10, 24)
    info = b.info()
    assert info["n_values"] > 5_000_000_000
    d = torch.zeros(n_seg, dtype=torch.int64, device="cuda")
    b.run(fls.tune(64, 256, fls.SCHED_DYNAMIC), d)
    got = counts_of(d).copy()
    want, lo, hi = T.o_c4_counts(1, n_seg, 10, 24, threads=min(16, os.cpu_count() or 1))
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("bw", [1, 2, 4, 8, 3, 12])
def test_narrow_bitpack_ranges(bw):
    """Bit-packed segments (FOR base or none) under ranges that start below
    the base, end above the largest value, hold one value, every value, none
    above the values, and lo > hi (no value): the batched counts equal
    numpy's lo <= x <= hi on the values."""
    torch = T.torch_cuda()
    rng = np.random.default_rng(bw)
    n = 20 * 1024 + 333
    segs, want, keep = [], [], []
    for base in (0, 1000):
        x = (rng.integers(0, 2**bw, n) + base).astype(np.uint32)
        col = fls.Column(fls.ENC_BITPACK, n, fls.pack_column_device(T.to_dev(x), n, bw, base=base), bw=bw,
                         base=base)
        keep.append(col)
        top = base + 2**bw - 1
        for lo, hi in ((base + 2**bw // 4, base + 3 * 2**bw // 4), (0, base + 2**bw // 2), (0, 0xFFFFFFFF),
                       (base + 2**bw // 2, 0xFFFFFFFF), (base, base), (top, top), (top + 1, top + 100),
                       (base + 1, base), (5, 3), (0, max(base, 1) - 1)):
            segs.append({"col": col, "rmeta": None, "n_values": n, "lo": lo, "hi": hi})
            want.append(int(((x >= lo) & (x <= hi)).sum()))
    b = fls.batch_from_segments(segs)
    d = torch.zeros(len(segs), dtype=torch.int64, device="cuda")
    for tile, block in ((4, 256), (64, 512)):
        d.fill_(-1)
        b.run(fls.tune(tile, block, fls.SCHED_DYNAMIC), d)
        np.testing.assert_array_equal(counts_of(d), np.array(want, np.uint64), err_msg=f"bw {bw} tile {tile}")


def test_empty_ranges_every_encoding():
    """hi < lo counts no value whatever the segment's encoding (the fast paths,
    the one-lane path of partial and MIXED vectors, and the delta path with
    widths that change inside a tile)."""
    torch = T.torch_cuda()
    rng = np.random.default_rng(11)
    n = 9 * 1024 + 77
    segs, want, keep = [], [], []
    x_sorted = np.sort(rng.integers(0, 2**20, n)).astype(np.uint32)
    w, vm, bs = fls.delta_encode_device(T.to_dev(x_sorted), n)
    keep.append(fls.Column(fls.ENC_DELTA, n, w, vmeta=vm, bases=bs))
    starts = rng.random(n) < 0.125
    starts[0] = True
    x_rle = rng.integers(0, 2**12, n).astype(np.uint32)[np.cumsum(starts) - 1]
    w2, vm2, rm2, bs2 = fls.rle_encode_device(T.to_dev(x_rle), n)
    keep.append(fls.Column(fls.ENC_RLE, n, w2, vmeta=vm2, bases=bs2, aux=rm2))
    x_bp = rng.integers(0, 2**9, n).astype(np.uint32)
    keep.append(fls.Column(fls.ENC_BITPACK, n, fls.pack_column_device(T.to_dev(x_bp), n, 9), bw=9))
    for col, x in zip(keep, (x_sorted, x_rle, x_bp)):
        mid = int(np.median(x))
        for lo, hi in ((mid + 1, mid), (0xFFFFFFFF, 0), (mid, mid)):
            segs.append({"col": col, "rmeta": None, "n_values": n, "lo": lo, "hi": hi})
            want.append(int(((x >= lo) & (x <= hi)).sum()))
    b = fls.batch_from_segments(segs)
    d = torch.zeros(len(segs), dtype=torch.int64, device="cuda")
    for tile in (4, 64):
        d.fill_(-1)
        b.run(fls.tune(tile, 256, fls.SCHED_DYNAMIC), d)
        np.testing.assert_array_equal(counts_of(d), np.array(want, np.uint64), err_msg=f"tile {tile}")
        d.fill_(-1)
        b.run_per_segment(fls.tune(tile, 256, fls.SCHED_STATIC), d)
        np.testing.assert_array_equal(counts_of(d), np.array(want, np.uint64), err_msg=f"per segment, tile {tile}")
