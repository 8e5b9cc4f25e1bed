"""The C++ drop-in header (include/fls/fastlanes.hpp) compiles with the host
compiler, links libfls.so and reproduces the reference pack() golden words."""
import shutil
import subprocess

import numpy as np
import pytest

import fls_testlib as T

PROGRAM = r"""
#include <cstdio>
#include <vector>
#include "fls/fastlanes.hpp"
int main(int argc, char** argv) {
    // reads 1024 values from stdin, packs them at bw=argv[1] the way the
    // reference drivers do (fls_q11.cu:207-223), prints the packed words
    unsigned bw = std::stoi(argv[1]);
    std::vector<uint32_t> in(1024), out(1024), t(1024), back(1024);
    for (auto& x : in) if (std::scanf("%u", &x) != 1) return 2;
    generated::pack::fallback::scalar::pack(in.data(), out.data(), bw);
    generated::transpose::fallback::scalar::transpose_i(in.data(), t.data());
    generated::transpose::fallback::scalar::transpose_o(t.data(), back.data());
    if (back != in) return 3;
    for (unsigned i = 0; i < 32 * bw; ++i) std::printf("%u\n", out[i]);
    try { generated::pack::fallback::scalar::pack(in.data(), out.data(), 40); return 4; }
    catch (const std::runtime_error&) {}
    return 0;
}
"""


def test_cpp_dropin_header(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no host C++ compiler")
    src = tmp_path / "dropin.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "dropin"
    libdir = T.ROOT / "fastlanesgpu_amd"
    subprocess.run(["g++", "-std=c++17", "-O1", f"-I{T.ROOT / 'include'}", str(src), f"-L{libdir}", "-lfls",
                    f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True)
    g = T.golden()
    for bw in (0, 3, 11, 24, 32):
        inp = "\n".join(str(int(x)) for x in g["rand_normal"][bw])
        res = subprocess.run([str(exe), str(bw)], input=inp, capture_output=True, text=True)
        assert res.returncode == 0, res
        words = np.array([int(x) for x in res.stdout.split()], dtype=np.uint64).astype(np.uint32)
        np.testing.assert_array_equal(words, g["packed_normal"][bw][: 32 * bw])
