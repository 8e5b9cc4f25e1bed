"""Write the golden fixtures for the FastLanes bit layout and delta transform.

Run in the container that holds the reference checkout, after `make -C oracle`
built oracle/_ref/libflsref.so from the reference's own host encoders
(fastlanes/src/pack.cpp, transpose.cpp, unrsum.cpp):

    python tests/golden/make_golden.py

Inputs are data the reference's own tests hold:
  * helper::rand_arr_<bw>_b<bw>_w32_arr[1024], bw = 0..32, from the generated
    unpack test helpers of both variants
    (fastlanes/generated/cuda/{normal,fused}_t32_uf1/*_unpack_helper.hpp), the
    known-answer inputs of cuda_*_unpack_test.cu:37-366;
  * the constant-data arrays of fastlanes/src/bitpack_register.cu:131 (value 5
    masked to bw) and fastlanes/example/fastlanes_bench_delta.cu:132 (i % 1024).
Outputs (expected values) come from the compiled reference encoders.

Writes tests/golden/fls_layout_golden.npz with:
  rand_normal[33,1024], rand_fused[33,1024]   golden inputs
  packed_normal[33,1024], packed_fused[33,1024]  reference pack() output
                                              (first 32*bw words valid, rest 0)
  transpose_i_perm[1024]                      ref transpose_i applied to arange
  transpose_o_ref[1024]                       ref transpose_o applied to arange
  unrsum_in[4,1024], unrsum_out[4,1024]        ref unrsum on seeded inputs
  delta_i1024_packed[1024], delta_i1024_bases[32], delta_i1024_bw  the
                                              fastlanes_bench_delta.cu encode of i%1024
  rsum_order[32]                              chain offsets read by d_rsum_32
"""
import ctypes
import pathlib
import re

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parents[2]
REF = pathlib.Path("/root/reference")
HELPERS = {
    "normal": REF / "fastlanes/generated/cuda/normal_t32_uf1/cuda_normal_t32_1024_uf1_unpack_helper.hpp",
    "fused": REF / "fastlanes/generated/cuda/fused_t32_uf1/cuda_fused_t32_1024_uf1_unpack_helper.hpp",
}


def parse_helper(path):
    text = path.read_text()
    out = np.zeros((33, 1024), np.uint32)
    for m in re.finditer(r"rand_arr_(\d+)_b(\d+)_w32_arr\[1024\]\s*=\s*\{(.*?)\};", text, re.S):
        bw = int(m.group(1))
        vals = [int(t.rstrip("UL")) for t in re.findall(r"\d+UL|\d+", m.group(3))]
        assert len(vals) == 1024, (path, bw, len(vals))
        out[bw] = np.array(vals, dtype=np.uint64).astype(np.uint32)
    return out


def main():
    lib = ctypes.CDLL(str(ROOT / "oracle/_ref/libflsref.so"))
    P = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C")
    lib.ref_pack_u32.argtypes = [P, P, ctypes.c_uint]
    for n in ("ref_transpose_i_u32", "ref_transpose_o_u32", "ref_unrsum_u32"):
        getattr(lib, n).argtypes = [P, P]

    data = {}
    for variant, path in HELPERS.items():
        rand = parse_helper(path)
        packed = np.zeros((33, 1024), np.uint32)
        for bw in range(33):
            buf = np.zeros(1024, np.uint32)
            lib.ref_pack_u32(np.ascontiguousarray(rand[bw]), buf, bw)
            packed[bw] = buf
        data[f"rand_{variant}"] = rand
        data[f"packed_{variant}"] = packed

    ar = np.arange(1024, dtype=np.uint32)
    ti = np.zeros(1024, np.uint32)
    lib.ref_transpose_i_u32(ar, ti)
    to = np.zeros(1024, np.uint32)
    lib.ref_transpose_o_u32(ar, to)
    data["transpose_i_perm"] = ti
    data["transpose_o_ref"] = to

    rng = np.random.default_rng(20240611)
    ins = [
        rng.integers(0, 2**32, 1024, dtype=np.uint64).astype(np.uint32),
        np.sort(rng.integers(0, 2**20, 1024)).astype(np.uint32),
        ar.copy(),
        np.full(1024, 5, np.uint32),
    ]
    uin = np.stack(ins)
    uout = np.zeros_like(uin)
    for k in range(len(ins)):
        lib.ref_unrsum_u32(np.ascontiguousarray(uin[k]), uout[k])
    data["unrsum_in"] = uin
    data["unrsum_out"] = uout

    # fastlanes_bench_delta.cu:131-155 with bw >= 10: value i % 1024, then
    # transpose_i, unrsum, bases = first 32 transposed values, pack.
    t = np.zeros(1024, np.uint32)
    lib.ref_transpose_i_u32(ar, t)
    d = np.zeros(1024, np.uint32)
    lib.ref_unrsum_u32(t, d)
    bw = int(d.max()).bit_length()
    packed = np.zeros(1024, np.uint32)
    lib.ref_pack_u32(d, packed, bw)
    data["delta_i1024_packed"] = packed
    data["delta_i1024_bases"] = t[:32].copy()
    data["delta_i1024_bw"] = np.array([bw], np.uint32)

    # The position order of d_rsum_32: the offsets it reads after the lane
    # index, in program order (fastlanes/src/include/fls_gen/rsum/rsum.cuh:8-106).
    rsum_src = (REF / "fastlanes/src/include/fls_gen/rsum/rsum.cuh").read_text()
    order = [int(m) for m in re.findall(r"r_0\s*=\s*\*\(in \+ \(0 \* 32\) \+ \(trd_idx \* 1\) \+ (\d+)\)", rsum_src)]
    assert len(order) == 32, order
    data["rsum_order"] = np.array(order, np.uint32)

    np.savez_compressed(ROOT / "tests/golden/fls_layout_golden.npz", **data)
    print("wrote", ROOT / "tests/golden/fls_layout_golden.npz")


if __name__ == "__main__":
    main()
