"""The device exp (bioneuroncore_amd/csrc/bc_libm.h) returns the bits of the host glibc ::exp
the reference calls, and the device x^4 matches glibc pow(x, 4) after rounding to float.
The header is compiled for the host with hipcc and compared on millions of inputs."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = r'''
#include "bc_libm.h"
#include <cmath>
#include <cstdio>
#include <random>
int main() {
    std::mt19937_64 g(11);
    long bad = 0, n = 0;
    for (long i = 0; i < 3000000; i++) {
        double x;
        switch (i % 5) {
            case 0: x = std::uniform_real_distribution<double>(-745, 709.7)(g); break;
            case 1: x = (double)(float)std::uniform_real_distribution<double>(-60, 60)(g); break;
            case 2: x = std::uniform_real_distribution<double>(-1, 1)(g); break;
            case 3: x = std::ldexp(std::uniform_real_distribution<double>(-1, 1)(g), -(int)(g() % 60)); break;
            default: x = -(double)(float)std::uniform_real_distribution<double>(0, 1)(g); break;
        }
        n++;
        if (bc::as_u64(bc::exp(x)) != bc::as_u64(std::exp(x))) bad++;
    }
    long bad4 = 0;
    for (long i = 0; i < 2000000; i++) {
        float x = (float)std::uniform_real_distribution<double>(0, 200)(g);
        float a = (float)(1.0 / (1.0 + std::pow((double)x, 4.0)));
        float b = (float)(1.0 / (1.0 + (double)x * (double)x * ((double)x * (double)x)));
        if (a != b) bad4++;
    }
    printf("%ld %ld %ld\n", bad, n, bad4);
    return 0;
}
'''


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"), reason="hipcc absent")
def test_exp_bits_match_glibc(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(SRC)
    exe = tmp_path / "t"
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    subprocess.run([hipcc, "-O2", "-ffp-contract=off", "-I", os.path.join(ROOT, "bioneuroncore_amd", "csrc"),
                    "-o", str(exe), str(src)], check=True, capture_output=True)
    bad, n, bad4 = map(int, subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split())
    assert n == 3000000
    assert bad == 0
    assert bad4 == 0
