"""GPU parity of untranspose (transpose_i / transpose_o on the device) and of
decode-to-materialize (any encoding -> plain values, stored or row order)
against the oracle's permutations and the values that were encoded."""
import numpy as np
import pytest

import fastlanesgpu_amd as fls
import fls_testlib as T

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(2027)


@pytest.mark.parametrize("n_vec", [1, 3, 257])
def test_transpose_matches_oracle(n_vec):
    x = RNG.integers(0, 2**32, n_vec * 1024, dtype=np.uint64).astype(np.uint32)
    fwd = T.from_dev(fls.transpose_vectors(T.to_dev(x), n_vec))
    inv = T.from_dev(fls.transpose_vectors(T.to_dev(x), n_vec, inverse=True))
    for v in range(n_vec):
        s = slice(1024 * v, 1024 * (v + 1))
        np.testing.assert_array_equal(fwd[s], T.o_perm("orc_transpose_i_u32", x[s]))
        np.testing.assert_array_equal(inv[s], T.o_perm("orc_transpose_o_u32", x[s]))
    back = T.from_dev(fls.transpose_vectors(T.to_dev(fwd), n_vec, inverse=True))
    np.testing.assert_array_equal(back, x)


def column_values(n, seed):
    """Sorted runs with noise: every encoding has something to work with."""
    r = np.random.default_rng(seed)
    base = np.sort(r.integers(0, 5_000_000, n)).astype(np.uint32)
    runs = np.repeat(r.integers(0, 1 << 20, n // 37 + 1), 37)[:n].astype(np.uint32)
    return np.where(r.random(n) < 0.5, base, runs).astype(np.uint32)


KINDS = ["bitpack", "bitpack_t", "delta", "rle", "for", "exhaustive"]


def encode(x, kind):
    if kind == "bitpack":
        return fls.bitpack_column(x).to_device()
    return T.encode_any(x, kind)


@pytest.mark.parametrize("kind", KINDS)
@pytest.mark.parametrize("n", [1, 1000, 1024, 5_000 + 7, 300_000 + 11])
def test_decode_rows_is_the_encoded_column(kind, n):
    x = column_values(n, n)
    col = encode(x, kind)
    got = T.from_dev(fls.decode_column(col, n, fls.ORDER_ROWS))
    assert got.size == n
    np.testing.assert_array_equal(got, x)


@pytest.mark.parametrize("kind", ["bitpack_t", "delta", "rle"])
@pytest.mark.parametrize("n", [5_000 + 7, 4 * 1024 + 1, 4 * 1024 + 130])
def test_decode_rows_into_an_unaligned_output(kind, n):
    # row order into an output that is not 16-byte aligned takes the
    # one-word stores; the sizes end the last vector inside a 4-row group
    torch = T.torch_cuda()
    x = column_values(n, n + 3)
    col = encode(x, kind)
    buf = torch.full((n + 8,), -1, dtype=torch.int32, device="cuda")
    fls.decode_column(col, n, fls.ORDER_ROWS, out=buf[1:n + 1])
    got = T.from_dev(buf)
    np.testing.assert_array_equal(got[1:n + 1], x)
    assert got[0] == 0xFFFFFFFF and np.all(got[n + 1:] == 0xFFFFFFFF)
    aligned = T.from_dev(fls.decode_column(col, n, fls.ORDER_ROWS))
    np.testing.assert_array_equal(aligned, x)


@pytest.mark.parametrize("kind", ["bitpack_t", "delta", "rle"])
def test_decode_stored_order_into_an_unaligned_output(kind):
    # stored order into an output that is not 16-byte aligned takes the
    # one-word stores; the aligned output takes the LDS stage and 16-byte
    # stores: both hold the same words
    torch = T.torch_cuda()
    n = 3 * 1024 + 300
    x = column_values(n, n + 5)
    col = encode(x, kind)
    aligned = T.from_dev(fls.decode_column(col, n, fls.ORDER_STORED))
    buf = torch.full((4 * 1024 + 8,), -1, dtype=torch.int32, device="cuda")
    fls.decode_column(col, n, fls.ORDER_STORED, out=buf[1:4 * 1024 + 1])
    got = T.from_dev(buf)
    np.testing.assert_array_equal(got[1:4 * 1024 + 1], aligned)
    assert got[0] == 0xFFFFFFFF and np.all(got[4 * 1024 + 1:] == 0xFFFFFFFF)


@pytest.mark.parametrize("kind", KINDS)
def test_decode_stored_order(kind):
    n = 9 * 1024 + 500
    x = column_values(n, 77)
    col = encode(x, kind)
    got = T.from_dev(fls.decode_column(col, n, fls.ORDER_STORED))
    assert got.size == 10 * 1024
    if kind == "bitpack":
        np.testing.assert_array_equal(got[:n], x)
        return
    # transposed storage: transpose_o of each stored vector gives the rows back
    rows = np.concatenate([T.o_perm("orc_transpose_o_u32", got[1024 * v:1024 * (v + 1)]) for v in range(10)])
    np.testing.assert_array_equal(rows[:n], x)


def test_full_size_delta_date_column():
    # 120M-row orderdate of the Q1.1 table, delta-encoded on the GPU: the
    # row-order materialization equals the generated column on the device
    torch = T.torch_cuda()
    n = 120_000_000
    od = fls.gen_c2(5, n)[0]
    w, vm, bs = fls.delta_encode_device(od, n)
    col = fls.Column(fls.ENC_DELTA, n, w, vmeta=vm, bases=bs)
    out = fls.decode_column(col, n, fls.ORDER_ROWS)
    assert bool(torch.equal(out, od[:n]))
    st = fls.decode_column(col, n, fls.ORDER_STORED)
    back = fls.transpose_vectors(st, (n + 1023) // 1024, inverse=True)
    assert bool(torch.equal(back[:n], od[:n]))


def test_errors():
    x = column_values(3000, 1)
    col = encode(x, "delta")
    with pytest.raises(fls.FlsError):
        fls.decode_column(col, 3000, order=7)
    with pytest.raises(fls.FlsError):  # in place is not a permutation
        d = T.to_dev(np.arange(1024, dtype=np.uint32))
        fls.transpose_vectors(d, 1, out=d)
    with pytest.raises(fls.FlsError):  # vectors are moved 16 bytes at a time
        d = T.to_dev(np.arange(2048, dtype=np.uint32))
        fls.transpose_vectors(d[1:1025], 1)
