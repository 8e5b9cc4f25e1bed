"""fls_query_q21_noalloc (the allocation-free Q2.1 entry point) through the C
ABI: two calls in a row, the second on a smaller table with smaller
dimensions, each equal to fls_q21_build + fls_query_q21 and to the oracle;
dimensions beyond the static scratch are refused before any work."""
import ctypes

import numpy as np
import pytest

import fastlanesgpu_amd as fls
import fls_testlib as T

pytestmark = pytest.mark.gpu


def dev_dims(dims_np, sizes):
    torch = T.torch_cuda()
    d = {k: torch.from_numpy(v.view(np.int32).copy()).cuda() for k, v in dims_np.items()}
    d.update(sizes)
    return d


def one_call(seed, n, sizes):
    fact, dims = T.o_gen_c3(seed, n, **sizes)
    fact[1] %= sizes["n_supp"]
    fact[2] %= sizes["n_part"]
    enc = fls.encode_c3_table(*fact, n_date=sizes["n_date"], n_supp=sizes["n_supp"], n_part=sizes["n_part"])
    cols = [c.to_device() for c in enc]
    d = dev_dims(dims, {k: sizes[k] for k in ("n_date", "n_supp", "n_part")})
    want = fls.query_q21(*cols, n, d, fls.q21_build(d)).cpu().numpy()
    got = fls.query_q21_noalloc(*cols, n, d).cpu().numpy()
    np.testing.assert_array_equal(got, want)
    sums, counts = T.o_q21(fact, dims, n)
    np.testing.assert_array_equal(got[:7000], sums)
    np.testing.assert_array_equal(got[7000:].view(np.uint64), counts)
    assert counts.sum() > 0


def test_noalloc_equals_build_and_query_on_two_calls():
    # full dimension sizes first (1.7 MB of tables), then a smaller table
    # whose smaller tables sit at other offsets of the same scratch array
    one_call(3, 2_000_000 + 11, dict(n_date=2556, n_supp=40000, n_part=800000, n_cust=10))
    one_call(4, 70_000 + 5, dict(n_date=2556, n_supp=400, n_part=8000, n_cust=10))


def test_noalloc_refuses_dimensions_beyond_the_scratch():
    torch = T.torch_cuda()
    n_part = 3_000_000  # 375 KB of bitmap + 6 MB of brands > 4 MiB
    assert fls.load().fls_q21_workspace_bytes(40000, n_part, 2556) > 4 << 20  # FLS_Q21_STATIC_SCRATCH_BYTES
    z = lambda k: torch.zeros(k, dtype=torch.int32, device="cuda")  # noqa: E731
    d = {"s_region": z(40000), "n_supp": 40000, "p_category": z(n_part), "p_brand": z(n_part), "n_part": n_part,
         "d_year": z(2556), "n_date": 2556}
    cols = [fls.bitpack_column(np.zeros(1024, np.uint32), 1).to_device() for _ in range(4)]
    out = torch.full((14000,), 7, dtype=torch.int64, device="cuda")
    with pytest.raises(fls.FlsError, match="static scratch"):
        fls.query_q21_noalloc(*cols, 1024, d, d_out=out)
    torch.cuda.synchronize()
    assert bool((out == 7).all())  # nothing was queued
