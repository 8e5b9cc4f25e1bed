"""The run-time-divisor remainder of the group-by slots (fls_fastdiv.hpp)
against the hardware's own `/` and `%`, on the CPU: divisors 1, powers of two
and their neighbours, the group-table sizes of the query shapes, the largest
32-bit values and random ones, each over edge and random dividends, plus every
dividend below 2^20 for the table sizes."""
import pathlib
import subprocess

import pytest

ROOT = pathlib.Path(__file__).resolve().parents[1]

PROGRAM = r"""
#include <stdio.h>
#include <stdint.h>
#include "fls_fastdiv.hpp"

static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint32_t rnd() { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return (uint32_t)(s >> 11); }

static int check(uint32_t x, uint32_t d) {
    const fls::FastDiv f = fls::make_fastdiv(d);
    if (fls::fast_div(x, f) != x / d || fls::fast_mod(x, f) != x % d) {
        printf("FAIL x=%u d=%u\n", x, d);
        return 1;
    }
    return 0;
}

int main() {
    uint32_t ds[4096];
    int nd = 0;
    const uint32_t fixed[] = {1u, 3u, 5u, 7u, 25u, 175u, 625u, 4375u, 7000u, 8192u, 65535u, 65536u, 65537u,
                              0x7FFFFFFFu, 0x80000000u, 0x80000001u, 0xFFFFFFFEu, 0xFFFFFFFFu};
    for (uint32_t v : fixed) ds[nd++] = v;
    for (int k = 1; k < 32; ++k) {
        ds[nd++] = 1u << k;
        ds[nd++] = (1u << k) - 1u;
        ds[nd++] = (1u << k) + 1u;
    }
    while (nd < 2048) {
        const uint32_t r = rnd();
        ds[nd++] = (r >> (rnd() & 31u)) | 1u;
    }
    int bad = 0;
    for (int i = 0; i < nd; ++i) {
        const uint32_t d = ds[i];
        const uint32_t xs[] = {0u, 1u, d - 1u, d, d + 1u, 2u * d - 1u, 2u * d, 0x7FFFFFFFu, 0x80000000u,
                               0xFFFFFFFEu, 0xFFFFFFFFu, 0xFFFFFFFFu - d, (0xFFFFFFFFu / d) * d,
                               (0xFFFFFFFFu / d) * d - 1u};
        for (uint32_t x : xs) bad += check(x, d);
        for (int k = 0; k < 4000; ++k) bad += check(rnd(), d);
    }
    const uint32_t tables[] = {7u, 175u, 4375u, 7000u};
    for (uint32_t d : tables)
        for (uint32_t x = 0; x < (1u << 20); ++x) bad += check(x, d);
    printf("%s\n", bad ? "FAILED" : "OK");
    return bad != 0;
}
"""


def test_fastdiv_matches_hardware_division(tmp_path):
    src = tmp_path / "fastdiv_test.cpp"
    exe = tmp_path / "fastdiv_test"
    src.write_text(PROGRAM)
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'fastlanesgpu_amd' / 'csrc'}", str(src), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "OK", r.stdout[-2000:]


def test_fastdiv_check_catches_a_wrong_magic(tmp_path):
    # the same checker fails when the magic number is off by one
    src = tmp_path / "bad.cpp"
    exe = tmp_path / "bad"
    src.write_text(PROGRAM.replace("fls::make_fastdiv(d);",
                                   "[&]{ auto g = fls::make_fastdiv(d); g.m -= (d > 2 && (d & (d - 1))); "
                                   "return g; }();"))
    subprocess.run(["g++", "-O2", "-std=c++17", f"-I{ROOT / 'fastlanesgpu_amd' / 'csrc'}", str(src), "-o", str(exe)],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    if r.returncode == 0:
        pytest.fail("a wrong magic number went unnoticed")
