"""bc::div_rcp, bc::div_wide and bc::div_f24 (bioneuroncore_amd/csrc/bc_libm.h) return the IEEE
quotient bit for bit: random float and double operands over wide exponent ranges (subnormal operands and
quotients near the underflow and overflow thresholds included), divisors with all-ones and
near-power-of-two significands, and zero/subnormal/infinite edge cases."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = r'''
#include "bc_libm.h"
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
int main() {
    std::mt19937_64 g(2024);
    long badf = 0, badd = 0;
    for (long i = 0; i < 40000000L; i++) {
        uint32_t ea = (uint32_t)(g() % 80 + 87), eb = (uint32_t)(g() % 80 + 87);
        uint32_t ma = g() & 0x7FFFFF, mb = (i % 3 == 0) ? 0x7FFFFF : (i % 3 == 1 ? (g() & 0xF) : (g() & 0x7FFFFF));
        uint32_t ba = (ea << 23) | ma | ((g() & 1u) << 31), bb = (eb << 23) | mb | ((g() & 1u) << 31);
        float a, b; memcpy(&a, &ba, 4); memcpy(&b, &bb, 4);
        float r = a / b, q = bc::div_rcp(a, b, bc::rcp_prep(b));
        if (memcmp(&r, &q, 4)) badf++;
        uint64_t Ea = g() % 600 + 723, Eb = g() % 600 + 723;
        uint64_t Ma = g() & 0xFFFFFFFFFFFFFull, Mb = (i % 3 == 0) ? 0xFFFFFFFFFFFFFull : (g() & 0xFFFFFFFFFFFFFull);
        uint64_t da = (Ea << 52) | Ma | ((g() & 1ull) << 63), db = (Eb << 52) | Mb | ((g() & 1ull) << 63);
        double A, B; memcpy(&A, &da, 8); memcpy(&B, &db, 8);
        double R = A / B, Q = bc::div_rcp(A, B, bc::rcp_prep(B));
        if (memcmp(&R, &Q, 8)) badd++;
    }
    long badw = 0;
    for (long i = 0; i < 40000000L; i++) {
        // exponents 0..254 (subnormal through huge), so quotients span underflow and overflow
        uint32_t ea = (uint32_t)(g() % 255), eb = (uint32_t)(i % 4 == 0 ? g() % 255 : g() % 60 + 97);
        uint32_t ma = g() & 0x7FFFFF, mb = (i % 3 == 0) ? 0x7FFFFF : (i % 3 == 1 ? (g() & 0xF) : (g() & 0x7FFFFF));
        uint32_t ba = (ea << 23) | ma | ((g() & 1u) << 31), bb = (eb << 23) | mb | ((g() & 1u) << 31);
        float a, b; memcpy(&a, &ba, 4); memcpy(&b, &bb, 4);
        float r = a / b, q = bc::div_wide(a, b, 1.0 / (double)b);
        if (memcmp(&r, &q, 4) && !(std::isnan(r) && std::isnan(q))) badw++;
    }
    long bade = 0;
    const float ea[] = {0.0f, -0.0f, 1e-40f, -1e-42f, INFINITY, -INFINITY, 3e38f, 1e-37f, 5.0f};
    const float eb[] = {1.0f, -3.0f, 1e-30f, 7e30f, 1e-38f, 3e-39f, 0.0f, -0.0f, INFINITY, 1e-45f};
    for (float a : ea) for (float b : eb) {
        float r = a / b, q = bc::div_rcp(a, b, bc::rcp_prep(b));
        if (memcmp(&r, &q, 4) && !(std::isnan(r) && std::isnan(q))) bade++;
        float w = bc::div_wide(a, b, 1.0 / (double)b);
        if (memcmp(&r, &w, 4) && !(std::isnan(r) && std::isnan(w))) bade++;
    }
    // div_f24: double numerators of any exponent, divisors that are widened floats
    long bad24 = 0;
    for (long i = 0; i < 40000000L; i++) {
        uint64_t Ea = (i % 5 == 0) ? g() % 2047 : g() % 400 + 823, Ma = g() & 0xFFFFFFFFFFFFFull;
        if (i % 7 == 0) Ma = (g() & 0xFFFFFFull) << 29;   // numerators that are floats too
        uint64_t da = (Ea << 52) | Ma | ((g() & 1ull) << 63);
        uint32_t eb = (uint32_t)(i % 4 == 0 ? g() % 255 : g() % 60 + 97);
        uint32_t mb = (i % 3 == 0) ? 0x7FFFFF : (i % 3 == 1 ? (g() & 0xF) : (g() & 0x7FFFFF));
        uint32_t bb = (eb << 23) | mb | ((g() & 1u) << 31);
        double A; float bf; memcpy(&A, &da, 8); memcpy(&bf, &bb, 4);
        const double B = bf;
        double R = A / B, Q = bc::div_f24(A, B, 1.0 / B);
        if (memcmp(&R, &Q, 8) && !(std::isnan(R) && std::isnan(Q))) bad24++;
    }
    const double ea24[] = {0.0, -0.0, 5e-324, -1e-310, INFINITY, -INFINITY, 1.7e308, 2.3e-308, 5.0};
    const float eb24[] = {1.0f, -3.0f, 1e-30f, 7e30f, 1e-38f, 3e-39f, 0.0f, -0.0f, INFINITY, 1e-45f};
    for (double a : ea24) for (float b : eb24) {
        double R = a / (double)b, Q = bc::div_f24(a, (double)b, 1.0 / (double)b);
        if (memcmp(&R, &Q, 8) && !(std::isnan(R) && std::isnan(Q))) bad24++;
    }
    printf("%ld %ld %ld %ld %ld\n", badf, badd, bade, badw, bad24);
}
'''


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"), reason="hipcc absent")
def test_div_rcp_is_ieee_division(tmp_path):
    src = tmp_path / "d.cpp"
    src.write_text(SRC)
    exe = tmp_path / "d"
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    subprocess.run([hipcc, "-O2", "-ffp-contract=off", "-I", os.path.join(ROOT, "bioneuroncore_amd", "csrc"),
                    "-o", str(exe), str(src)], check=True, capture_output=True)
    badf, badd, bade, badw, bad24 = map(int, subprocess.run([str(exe)], capture_output=True, text=True,
                                                            check=True).stdout.split())
    assert (badf, badd, bade, badw, bad24) == (0, 0, 0, 0, 0)
