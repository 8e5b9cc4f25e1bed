"""GPU parity of the decode kernels (C-ABI calls into libfls.so) vs the oracle."""
import numpy as np
import pytest

import fastlanesgpu_amd as fls
import fls_testlib as T

pytestmark = pytest.mark.gpu
RNG = np.random.default_rng(2024)


@pytest.mark.parametrize("variant", ["normal", "fused"])
def test_unpack_reference_kat_all_widths(variant):
    # cuda_{normal,fused}_t32_1024_uf1_unpack_test.cu:37-366 on the GPU path
    g = T.golden()
    for bw in range(33):
        d_in = T.to_dev(np.ascontiguousarray(g[f"packed_{variant}"][bw][: max(1, 32 * bw)]))
        d_out = T.dev_empty(1024)
        fls.unpack_global(d_in, d_out, 1, bw)
        np.testing.assert_array_equal(T.from_dev(d_out), g[f"rand_{variant}"][bw], err_msg=f"bw={bw}")


def test_unpack_many_vectors_every_width():
    n_vec = 37
    for bw in range(33):
        x = RNG.integers(0, 2**bw, n_vec * 1024, dtype=np.uint64).astype(np.uint32)
        words = fls.pack_column(x, bw)
        # spot-check the packed words of two vectors against the oracle
        for v in (0, n_vec - 1):
            np.testing.assert_array_equal(words[v * 32 * bw:(v + 1) * 32 * bw], T.o_pack(x[v * 1024:(v + 1) * 1024], bw))
        d_in = T.to_dev(words if words.size else np.zeros(1, np.uint32))
        d_out = T.dev_empty(n_vec * 1024)
        fls.unpack_global(d_in, d_out, n_vec, bw)
        np.testing.assert_array_equal(T.from_dev(d_out), x, err_msg=f"bw={bw}")


@pytest.mark.parametrize("bw", [3, 11, 32])
def test_unpack_into_an_unaligned_output(bw):
    # 16-byte stores need a 16-byte aligned output; any other output takes the
    # one-word stores, with the same values and checksum
    torch = T.torch_cuda()
    n_vec = 5
    x = RNG.integers(0, 2**bw, n_vec * 1024, dtype=np.uint64).astype(np.uint32)
    d_in = T.to_dev(fls.pack_column(x, bw))
    buf = torch.full((n_vec * 1024 + 8,), -1, dtype=torch.int32, device="cuda")
    fls.unpack_global(d_in, buf[1:], n_vec, bw)
    got = T.from_dev(buf)
    np.testing.assert_array_equal(got[1:n_vec * 1024 + 1], x)
    assert got[0] == 0xFFFFFFFF and np.all(got[n_vec * 1024 + 1:] == 0xFFFFFFFF)
    d_sum = fls.checksum_buffer()
    fls.unpack_checksum(d_in, buf[3:], n_vec * 1024, bw, d_sum)
    np.testing.assert_array_equal(T.from_dev(buf)[3:n_vec * 1024 + 3], x)
    assert int(T.from_dev(d_sum, np.uint64)[0]) == T.o_checksum(x)


def test_c1_full_size_values_and_checksum():
    # plain bit-unpack workload: 2^24 values uniform in [0, 2^11), bw 11
    n, bw, seed = 1 << 24, 11, 42
    x = T.o_gen_c1(seed, n, bw)
    words = fls.pack_column(x, bw)
    d_in = T.to_dev(words)
    d_out = T.dev_empty(n)
    d_sum = fls.checksum_buffer()
    fls.unpack_checksum(d_in, d_out, n, bw, d_sum)
    np.testing.assert_array_equal(T.from_dev(d_out), x)
    want = T.o_checksum(x)
    assert int(T.from_dev(d_sum, np.uint64)[0]) == want
    # checksum-only mode (no materialisation) gives the same number
    fls.unpack_checksum(d_in, None, n, bw, d_sum)
    assert int(T.from_dev(d_sum, np.uint64)[0]) == want


@pytest.mark.parametrize("n", [1, 1000, 1025, 5 * 1024 + 77])
@pytest.mark.parametrize("bw", [0, 1, 7, 11, 24, 32])
def test_checksum_ragged(n, bw):
    x = RNG.integers(0, 2**bw, n, dtype=np.uint64).astype(np.uint32)
    words = fls.pack_column(x, bw)
    d_in = T.to_dev(words if words.size else np.zeros(1, np.uint32))
    n_vec = (n + 1023) // 1024
    d_out = T.dev_empty(n_vec * 1024)
    d_sum = fls.checksum_buffer()
    fls.unpack_checksum(d_in, d_out, n, bw, d_sum)
    np.testing.assert_array_equal(T.from_dev(d_out)[:n], x)
    assert int(T.from_dev(d_sum, np.uint64)[0]) == T.o_checksum(x)


def test_checksum_buffer_reused_across_sizes():
    # one buffer, calls of shrinking and growing grids (fewer blocks than the
    # reduction's 64 groups, exactly 64, and many more): every result is
    # right and the reduction state is zero again after each call
    d_sum = fls.checksum_buffer()
    for n, bw in ((3 * 8 * 1024 + 5, 9), (64 * 8 * 1024, 13), (1, 4), (700 * 8 * 1024 + 999, 11), (8 * 1024, 32)):
        x = RNG.integers(0, 2**bw, n, dtype=np.uint64).astype(np.uint32)
        d_in = T.to_dev(fls.pack_column(x, bw))
        n_vec = (n + 1023) // 1024
        fls.unpack_checksum(d_in, T.dev_empty(n_vec * 1024), n, bw, d_sum)
        got = T.from_dev(d_sum, np.uint64)
        assert int(got[0]) == T.o_checksum(x), (n, bw)
        assert not got[1:].any(), (n, bw)
    with pytest.raises(ValueError, match="checksum_buffer"):
        fls.unpack_checksum(d_in, None, n, bw, T.dev_empty(1, "u64"))


@pytest.mark.parametrize("kind", ["sorted", "random", "dates", "constant"])
def test_delta_decode(kind):
    n = 11 * 1024 + 300
    if kind == "sorted":
        x = np.sort(RNG.integers(0, 2**31, n)).astype(np.uint32)
    elif kind == "random":
        x = RNG.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    elif kind == "dates":
        x = T.o_gen_c2(5, n)[0]
    else:
        x = np.full(n, 123456, np.uint32)
    words, vmeta, bases = fls.delta_encode_column(x)
    n_vec = vmeta.size
    d_out = T.dev_empty(n_vec * 1024)
    fls.delta_decode(T.to_dev(words if words.size else np.zeros(1, np.uint32)), T.to_dev(vmeta), T.to_dev(bases),
                     n_vec, d_out)
    padded = np.full(n_vec * 1024, x[-1], np.uint32)
    padded[:n] = x
    np.testing.assert_array_equal(T.from_dev(d_out), T.transpose_vectors(padded))
    # the first vector once more through the oracle's own delta decode
    bw = int(vmeta[0]) & 0xFF
    np.testing.assert_array_equal(T.from_dev(d_out)[:1024], T.o_delta_decode(words[:32 * bw], bw, bases[:32]))


def test_copy_kernel():
    torch = T.torch_cuda()
    src = torch.arange(1 << 20, dtype=torch.int32, device="cuda")
    dst = torch.zeros_like(src)
    fls.call("fls_copy_bytes", src.data_ptr(), dst.data_ptr(), src.numel() * 4, fls.current_stream())
    torch.cuda.synchronize()
    assert torch.equal(src, dst)
    assert fls.load().fls_copy_bytes(src.data_ptr(), dst.data_ptr(), 10, None) == -6


@pytest.mark.parametrize("n_words", [4, 1 << 20, (1 << 22) + 12])
def test_copy_and_read_kernels_sizes(n_words):
    """The unrolled copy and read-only kernels at sizes that exercise the
    four-in-flight main loop and the remainder loop."""
    torch = T.torch_cuda()
    g = torch.Generator(device="cpu").manual_seed(n_words)
    src = torch.randint(-(1 << 31), (1 << 31) - 1, (n_words,), dtype=torch.int32, generator=g).cuda()
    dst = torch.zeros_like(src)
    fls.call("fls_copy_bytes", src.data_ptr(), dst.data_ptr(), n_words * 4, fls.current_stream())
    out = torch.full((1,), 12345, dtype=torch.int64, device="cuda")
    fls.call("fls_read_bytes", src.data_ptr(), n_words * 4, out.data_ptr(), fls.current_stream())
    torch.cuda.synchronize()
    assert torch.equal(src, dst)
    expect = int(src.cpu().numpy().view(np.uint32).astype(np.uint64).sum()) % (1 << 64)
    assert int(out.cpu().numpy().view(np.uint64)[0]) == expect
