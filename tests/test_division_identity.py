"""The correctly rounded quotients of the WENO5 kernels (csrc/fs_device.h,
fs_div_const and fs_div_recip) against IEEE division on the host.

The kernels take the reference's x/3, (x*7)/6 and the weight
normalisation a_k/sum by Markstein's correction from a correctly rounded
reciprocal (one correction for the constant divisors 3 and 6, whose RN(1/d)
has relative error 2^-54; two for a variable divisor).  The theorem says the
result is RN(x/d); this checks it bit for bit on random operands over the
exponent range the kernels see, on multiples of the divisor and on signed
zeros, with the host's fused multiply-add."""
import os
import subprocess
import tempfile

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static uint64_t s = 0x9E3779B97F4A7C15ULL;
static uint64_t nx(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double rnd(int emin, int emax, int sign) {
    uint64_t m = nx() & ((1ULL << 52) - 1);
    int e = emin + (int)(nx() % (uint64_t)(emax - emin + 1));
    uint64_t b = ((uint64_t)(e + 1023) << 52) | m | (sign ? (nx() & 1) << 63 : 0);
    double x; memcpy(&x, &b, 8); return x;
}
static double div_const(double x, double d, double y) {
    double q0 = x * y;
    double r = fma(-d, q0, x);
    return copysign(fma(r, y, q0), x);
}
static double div_recip(double a, double b, double y) {
    double q0 = a * y;
    double q1 = fma(fma(-b, q0, a), y, q0);
    return fma(fma(-b, q1, a), y, q1);
}
static long bad = 0;
static void check(double want, double got) { if (memcmp(&want, &got, 8)) bad++; }
int main(void) {
    const long n = 4000000;
    for (long i = 0; i < n; i++) {
        double x = rnd(-900, 900, 1);
        if (i % 4 == 1) x = 3.0 * (double)(nx() % 1000000000);
        if (i % 4 == 2) x = rnd(-3, 3, 1) * 7.0;
        check(x / 3, div_const(x, 3, (double)1 / 3));
        check(x / 6, div_const(x, 6, (double)1 / 6));
        check((x * 11) / 6, div_const(x * 11, 6, (double)1 / 6));
        double a1 = rnd(-40, 40, 0), a2 = rnd(-40, 40, 0), a3 = rnd(-40, 40, 0);
        double S = a1 + a2 + a3, y = (double)1 / S;
        check(a1 / S, div_recip(a1, S, y));
        check(a2 / S, div_recip(a2, S, y));
        check(a3 / S, div_recip(a3, S, y));
    }
    double z[2] = {0.0, -0.0};
    for (int k = 0; k < 2; k++) {
        check(z[k] / 3, div_const(z[k], 3, (double)1 / 3));
        check(z[k] / 6, div_const(z[k], 6, (double)1 / 6));
        check(0.0 / 7.5, div_recip(0.0, 7.5, (double)1 / 7.5));
    }
    printf("%ld\n", bad);
    return 0;
}
"""


def test_markstein_quotients_equal_ieee_division():
    with tempfile.TemporaryDirectory() as d:
        c, exe = os.path.join(d, "q.c"), os.path.join(d, "q")
        with open(c, "w") as f:
            f.write(SRC)
        try:
            subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, c, "-lm"], check=True,
                           capture_output=True)
        except (OSError, subprocess.CalledProcessError) as e:
            pytest.skip(f"no host C compiler: {e}")
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    assert int(out.strip()) == 0
