"""The hot path driven from C++ through the C ABI alone, as the reference's
query driver does it (fls_q11.cu:137-223: host encode, copy to the device,
one fused launch, copy the sum back, check against a scalar loop): no Python
between the caller and libfls.so.  Built with hipcc where the test runs."""
import shutil
import subprocess

import pytest

import fls_testlib as T

pytestmark = pytest.mark.gpu

PROGRAM = r"""
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdint>
#include <vector>
#include "fls/fls.h"

static uint64_t st = 0x2545F4914F6CDD1Dull;
static uint32_t rnd(uint32_t m) { st ^= st << 13; st ^= st >> 7; st ^= st << 17; return (uint32_t)(st % m); }

#define CHECK(x) do { int rc_ = (x); if (rc_) { std::printf("FAIL %s: %s\n", #x, fls_status_string(rc_)); return 1; } } while (0)

static int to_dev(const std::vector<uint32_t>& h, uint32_t** d) {
    if (hipMalloc(d, h.size() * 4 + 4) != hipSuccess) return 1;
    return hipMemcpy(*d, h.data(), h.size() * 4, hipMemcpyHostToDevice) != hipSuccess;
}

int main() {
    const uint64_t n = 3 * 1024 * 1024 + 777;  // 3,146,505 rows, a partial last vector
    std::vector<uint32_t> od(n), q(n), d(n), p(n);
    for (uint64_t i = 0; i < n; ++i) {
        od[i] = 19920101 + 10000 * rnd(7) + 100 * rnd(12) + rnd(28);
        q[i] = 1 + rnd(50);
        d[i] = rnd(11);
        p[i] = 90000 + rnd(9910000);
    }
    // the reference's CPU check: a scalar loop over the raw columns
    int64_t want = 0;
    uint64_t want_n = 0;
    for (uint64_t i = 0; i < n; ++i)
        if (od[i] > 19930000 && od[i] < 19940000 && q[i] < 25 && d[i] >= 1 && d[i] <= 3) {
            want += (int64_t)p[i] * d[i];
            ++want_n;
        }
    // host encode: FOR-free bit-packing at 32 (orderdate), 6, 4 and 24 bits
    const uint32_t bws[4] = {32, 6, 4, 24};
    const std::vector<uint32_t>* cols[4] = {&od, &q, &d, &p};
    fls_column c[4];
    for (int k = 0; k < 4; ++k) {
        std::vector<uint32_t> packed(fls_packed_words(n, bws[k]));
        CHECK(fls_pack_column_u32(cols[k]->data(), n, bws[k], 0, packed.data()));
        uint32_t* dw = nullptr;
        if (to_dev(packed, &dw)) return 2;
        c[k] = fls_column{FLS_ENC_BITPACK, bws[k], 0, 0, packed.size(), dw, nullptr, nullptr, nullptr};
    }
    int64_t* d_rev = nullptr;
    uint64_t* d_cnt = nullptr;
    if (hipMalloc(&d_rev, 8) != hipSuccess || hipMalloc(&d_cnt, 8) != hipSuccess) return 2;
    fls_q11_params prm = {19930000, 19940000, 25, 1, 3};
    CHECK(fls_query_q11(&c[0], &c[1], &c[2], &c[3], n, &prm, 0, d_rev, d_cnt, nullptr));
    int64_t got = 0;
    uint64_t got_n = 0;
    if (hipMemcpy(&got, d_rev, 8, hipMemcpyDeviceToHost) != hipSuccess) return 2;
    if (hipMemcpy(&got_n, d_cnt, 8, hipMemcpyDeviceToHost) != hipSuccess) return 2;
    std::printf("q11 %lld %llu want %lld %llu\n", (long long)got, (unsigned long long)got_n, (long long)want,
                (unsigned long long)want_n);
    if (got != want || got_n != want_n) return 3;
    // decode-to-materialize the quantity column back to plain values
    uint32_t* d_plain = nullptr;
    if (hipMalloc(&d_plain, n * 4) != hipSuccess) return 2;
    CHECK(fls_decode_column_u32(&c[1], n, FLS_ORDER_ROWS, d_plain, nullptr));
    std::vector<uint32_t> back(n);
    if (hipMemcpy(back.data(), d_plain, n * 4, hipMemcpyDeviceToHost) != hipSuccess) return 2;
    if (back != q) return 4;
    std::printf("OK\n");
    return 0;
}
"""


def test_cpp_driver_q11_through_the_c_abi(tmp_path):
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    src = tmp_path / "driver.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "driver"
    libdir = T.ROOT / "fastlanesgpu_amd"
    subprocess.run([hipcc, "-std=c++17", "-O1", f"-I{T.ROOT / 'include'}", str(src), f"-L{libdir}", "-lfls",
                    f"-Wl,-rpath,{libdir}", "-o", str(exe)], check=True, timeout=180)
    res = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0 and res.stdout.strip().endswith("OK"), res.stdout + res.stderr
