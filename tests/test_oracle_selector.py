"""CPU checks of the selector oracle (oracle/selector.c): the per-vector mixed
encoders by round trip through the pinned decoders and by closed-form sizes
computed independently in numpy, the features by closed-form cases, the
training workload by its distribution, and the MLP restatement by finite
differences."""
import numpy as np
import pytest

from fls_testlib import (OrcRle, o_delta_decode, o_exhaustive_class, o_features, o_mix_encode, o_mix_sizes,
                         o_mlp_forward, o_mlp_grad, o_perm, o_train_vector, o_uniform_py, o_unpack, oracle)


def bitlen(x):
    return int(x).bit_length()


def vectors():
    rng = np.random.default_rng(7)
    out = {
        "zeros": np.zeros(1024, np.uint32),
        "const": np.full(1024, 0xDEADBEEF, np.uint32),
        "ramp": np.arange(1024, dtype=np.uint32),
        "ramp_big": (np.arange(1024, dtype=np.uint64) * 3 + 4_000_000_000).astype(np.uint32),
        "max": np.full(1024, 0xFFFFFFFF, np.uint32),
        "desc": np.arange(1024, 0, -1).astype(np.uint32),
        "runs8": np.repeat(rng.integers(0, 1 << 20, 128), 8).astype(np.uint32),
        "runs_odd": np.repeat(rng.integers(0, 1 << 9, 342), 3)[:1024].astype(np.uint32),
        "one_change": np.r_[np.zeros(500, np.uint32), np.full(524, 9, np.uint32)],
        "wrap_delta": np.r_[np.full(512, 0xFFFFFFF0, np.uint32), np.arange(512, dtype=np.uint32)],
    }
    for bw in (1, 5, 11, 17, 31, 32):
        out[f"u{bw}"] = rng.integers(0, 1 << bw, 1024, dtype=np.uint64).astype(np.uint32)
    out["for24"] = (rng.integers(0, 1 << 12, 1024) + 9_000_000).astype(np.uint32)
    out["sorted_noise"] = np.maximum(np.arange(1024) * 40 + rng.integers(-16, 17, 1024), 0).astype(np.uint32)
    return out


VECS = vectors()


def decode_mixed(words, enc, width, aux):
    """Decode a payload with the pinned decoders, back to original order."""
    if enc in (0, 1):
        t = o_unpack(words, width)
        if enc == 1:
            t = (t + np.uint32(aux & 0xFFFFFFFF)).astype(np.uint32)
    elif enc == 2:
        t = o_delta_decode(words[32:], width, words[:32])
    else:
        h = OrcRle()
        h.n_runs, h.vbase, h.vbw, h.ibw = (aux >> 32) & 0xFFFF, aux & 0xFFFFFFFF, width, (aux >> 48) & 1
        for l in range(32):
            h.bases[l] = int(words[l])
        idx = np.zeros(1024, np.uint32)
        idx[:32 * h.ibw] = words[32:32 + 32 * h.ibw]
        vals = np.zeros(1024, np.uint32)
        rest = words[32 + 32 * h.ibw:]
        vals[:rest.size] = rest
        t = np.zeros(1024, np.uint32)
        import ctypes
        oracle().orc_rle_decode_u32(ctypes.byref(h), vals, idx, t)
    return o_perm("orc_transpose_o_u32", t)


@pytest.mark.parametrize("name", sorted(VECS))
@pytest.mark.parametrize("enc", [0, 1, 2, 3])
def test_mix_encode_roundtrip(name, enc):
    x = VECS[name]
    words, width, aux = o_mix_encode(x, enc)
    np.testing.assert_array_equal(decode_mixed(words, enc, width, aux), x)


def closed_form_sizes(x):
    """Sizes from first principles: a FastLanes delta chain / RLE index chain
    is a block of 32 consecutive rows (rows 32c..32c+31)."""
    x = x.astype(np.uint64)
    blocks = x.reshape(32, 32)
    d = (blocks[:, 1:] - blocks[:, :-1]) % (1 << 32)
    w_delta = bitlen(int(d.max())) if d.size else 0
    runs = 1 + int(np.count_nonzero(x[1:] != x[:-1]))
    inner = int(np.count_nonzero(blocks[:, 1:] != blocks[:, :-1])) > 0
    w_for = bitlen(int(x.max() - x.min()))
    return np.array([128 * bitlen(int(x.max())) + 16, 128 * w_for + 16, 128 + 128 * w_delta + 16,
                     128 + 128 * inner + 4 * ((runs * w_for + 31) // 32) + 16], np.uint32)


@pytest.mark.parametrize("name", sorted(VECS))
def test_mix_sizes_closed_form(name):
    s, _ = o_mix_sizes(VECS[name])
    np.testing.assert_array_equal(s, closed_form_sizes(VECS[name]))


def test_exhaustive_class_cases():
    assert o_exhaustive_class(VECS["zeros"]) == 0           # bit-pack at width 0 (ties go to bit-pack)
    assert o_exhaustive_class(VECS["const"]) == 10          # FOR at width 0
    assert o_exhaustive_class(VECS["ramp_big"]) == 20 + 1   # delta, width 2 (step 3): bucket 1-2
    assert o_exhaustive_class(VECS["runs8"]) // 10 == 3     # RLE
    assert o_exhaustive_class(VECS["for24"]) == 10 + 5      # FOR, width 12 -> bucket 9-12
    assert o_exhaustive_class(VECS["u11"]) == 5             # bit-pack, width 11


def test_exhaustive_is_argmin():
    for name, x in VECS.items():
        s, w = o_mix_sizes(x)
        e = int(np.argmin(s))  # first minimum = lowest encoding number
        top = (0, 2, 4, 6, 8, 12, 16, 20, 24, 32)
        bucket = next(b for b, t in enumerate(top) if w[e] <= t)
        assert o_exhaustive_class(x) == 10 * e + bucket, name


def test_features_constant():
    f = o_features(np.full(1024, 5, np.uint32))
    expect = [3 / 32, 0, 0, 1, 1, 0, 1, 0, 0, 1, 3 / 32, 0, 3 / 32, 3 / 32, 0, 1]
    np.testing.assert_array_equal(f, np.array(expect, np.float32))


def test_features_ramp():
    x = np.arange(1024, dtype=np.uint32)
    f = o_features(x)
    sample = np.concatenate([x[32 * L:32 * L + 32] for L in (0, 10, 20, 31)])
    wsum = sum(bitlen(v) for v in sample)
    expect = [10 / 32, 10 / 32, 1 / 32, 0, 1, 1 / 32, 1 / 128, 1 / 32, 10 / 32, 0, 0, 0, wsum / 4096, 10 / 32, 1, 1]
    np.testing.assert_array_equal(f, np.array(expect, np.float32))


def test_features_sample_only():
    """Only lines 0, 10, 20, 31 (32 rows each) are read: changing any other row leaves the features."""
    rng = np.random.default_rng(3)
    x = rng.integers(0, 1 << 20, 1024).astype(np.uint32)
    y = x.copy()
    mask = ~np.isin(np.arange(1024) // 32, [0, 10, 20, 31])
    y[mask] = rng.integers(0, 1 << 30, int(mask.sum())).astype(np.uint32)
    np.testing.assert_array_equal(o_features(x), o_features(y))


def test_train_vector_families():
    from fastlanesgpu_amd.selector import zipf_cdf

    cdf = zipf_cdf()
    seed = 11
    run_lengths, rank1, zipf_n = [], 0, 0
    fams = set()
    for v in range(160):
        x = o_train_vector(seed, v, cdf).astype(np.int64)
        fam = o_uniform_py(seed, 300001, 8 * v, 4)
        fams.add(fam)
        if fam == 0:
            bw = 1 + o_uniform_py(seed, 300001, 8 * v + 1, 32)
            base = o_uniform_py(seed, 300001, 8 * v + 3, 1 << 31) if o_uniform_py(seed, 300001, 8 * v + 2, 2) else 0
            assert np.all(((x - base) % (1 << 32)) < (1 << bw))
        elif fam == 1:
            stride = 1 << o_uniform_py(seed, 300001, 8 * v + 1, 12)
            base = o_uniform_py(seed, 300001, 8 * v + 2, 1 << 20)
            r = (x - base)
            assert np.all(r % stride == 0)
            ranks = r // stride
            assert ranks.min() >= 1 and ranks.max() <= 4096
            rank1 += int(np.count_nonzero(ranks == 1))
            zipf_n += ranks.size
        elif fam == 2:
            step = o_uniform_py(seed, 300001, 8 * v + 1, 256)
            start = o_uniform_py(seed, 300001, 8 * v + 2, 1 << 24)
            resid = x - (start + np.arange(1024) * step)
            assert np.all(np.abs(resid[x > 0]) <= 16 * 6)  # Irwin-Hall bound: 6 sigma
            assert 8 < resid.std() < 24                       # sigma 16
        else:
            change = np.flatnonzero(np.diff(x) != 0)
            run_lengths.append(1024 / (len(change) + 1))
    assert fams == {0, 1, 2, 3}
    assert 0.12 < rank1 / zipf_n < 0.20   # P(rank 1) of Zipf 1.1 over 4096 ranks = 0.160
    assert 20 < np.mean(run_lengths) < 45  # geometric mean 32 (equal neighbouring runs merge)


# ----------------------------------------------------------------------- MLP
def rand_problem(n=24, seed=5):
    from fastlanesgpu_amd.selector import Mlp

    rng = np.random.default_rng(seed)
    w = Mlp.init_weights(seed).astype(np.float64)
    w += rng.normal(0, 0.05, w.size)  # non-zero biases too
    feat = rng.random((n, 16)).astype(np.float32)
    label = rng.integers(0, 40, n).astype(np.uint8)
    return w, feat, label


def test_mlp_forward_matches_numpy():
    w, feat, _ = rand_problem()
    W1, b1 = w[:1024].reshape(16, 64), w[1024:1088]
    W2, b2 = w[1088:5184].reshape(64, 64), w[5184:5248]
    W3, b3 = w[5248:7808].reshape(64, 40), w[7808:7848]
    h1 = np.maximum(feat.astype(np.float64) @ W1 + b1, 0)
    h2 = np.maximum(h1 @ W2 + b2, 0)
    np.testing.assert_allclose(o_mlp_forward(w, feat, False), h2 @ W3 + b3, rtol=1e-12, atol=1e-12)


def test_mlp_grad_finite_differences():
    w, feat, label = rand_problem()
    loss, g = o_mlp_grad(w, feat, label, False)
    n = feat.shape[0]
    rng = np.random.default_rng(1)
    groups = [(0, 1024), (1024, 1088), (1088, 5184), (5184, 5248), (5248, 7808), (7808, 7848)]
    idx = np.concatenate([rng.integers(a, b, 8) for a, b in groups])
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (o_mlp_grad(wp, feat, label, False)[0] - o_mlp_grad(wm, feat, label, False)[0]) / (2 * h) / n
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(fd)), (i, fd, g[i])
    # the loss is the summed softmax cross-entropy of the forward logits
    z = o_mlp_forward(w, feat, False)
    lse = np.log(np.exp(z - z.max(1, keepdims=True)).sum(1)) + z.max(1)
    assert loss == pytest.approx(float((lse - z[np.arange(n), label]).sum()), rel=1e-12)


def test_bf16_rounding():
    from fastlanesgpu_amd.selector import Mlp  # noqa: F401  (package import for path)

    o = oracle()
    from fls_testlib import _sel
    _sel()
    assert o.orc_bf16(1.0) == 1.0
    assert o.orc_bf16(1.0 + 2 ** -8) == 1.0                 # tie -> even
    assert o.orc_bf16(1.0 + 3 * 2 ** -8) == 1.0 + 2 ** -6   # tie -> even (up)
    assert o.orc_bf16(1.0 + 2 ** -8 + 2 ** -12) == 1.0 + 2 ** -7
    assert o.orc_bf16(-2.5) == -2.5
    w, feat, label = rand_problem()
    a = o_mlp_forward(w, feat, True)
    b = o_mlp_forward(w, feat, False)
    assert np.max(np.abs(a - b)) < 0.05 * np.max(np.abs(b))
