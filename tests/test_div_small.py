"""halton_qrng divides by its base b (a prime below 2^9) at every digit.
The kernel replaces the division with q0 = RN(x y), y = RN(1/b), and one
correction q1 = RN(q0 + r0 y), r0 = fma(-q0, b, x) (halton.hip, div_small).
For odd b, x/b is never closer than ulp/(2b) to a rounding midpoint (x = m b
for a midpoint m would need an odd 54-bit significand times b inside 53
bits), while q0 + r0 y is within about 1.5 2^-53 ulp of x/b, so q1 is the
IEEE quotient for b < 2^50.  The same sequence, compiled for the host with
gcc, must match division for every odd divisor below 4096 over the kernel's
domain [0, 2b) and over many binades."""
import os
import shutil
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
static double div_small(double x, double b, double y) {
    double q0 = x * y, r0 = fma(-q0, b, x);
    return fma(r0, y, q0);
}
static uint64_t st = 987654321;
static uint64_t next(void) {
    uint64_t z = (st += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static long bad = 0, n_checked = 0;
static void check(double x, double b, double y) {
    double a = div_small(x, b, y), c = x / b;
    n_checked++;
    if (memcmp(&a, &c, 8)) { if (bad < 5) printf("mismatch %a / %g: %a vs %a\n", x, b, a, c); bad++; }
}
int main(int argc, char **argv) {
    long per = atol(argv[1]);
    for (int ib = 3; ib < 4096; ib += 2) {
        const double b = ib, y = 1.0 / b;
        for (long i = 0; i < per; i++) {
            /* the kernel's domain: u in [0, 1) plus a digit below b */
            double u = (double)(next() >> 11) * 0x1p-53;
            double x = u + (double)(next() % (uint64_t)ib);
            check(x, b, y);
            /* a random double of a random binade in [2^-300, 2^10) */
            uint64_t m = next() & ((1ull << 52) - 1);
            int e = (int)(next() % 310) - 300;
            uint64_t bits = ((uint64_t)(e + 1023) << 52) | m;
            memcpy(&x, &bits, 8);
            check(x, b, y);
        }
        for (int e = -300; e <= 12; e++) {  /* powers of two, their neighbours, multiples of b near them */
            double p = ldexp(1.0, e);
            check(p, b, y); check(nextafter(p, 0), b, y); check(nextafter(p, INFINITY), b, y);
            check(p * b, b, y); check(nextafter(p * b, 0), b, y); check(nextafter(p * b, INFINITY), b, y);
        }
        check(0.0, b, y);
    }
    printf("%ld mismatches in %ld\n", bad, n_checked);
    return bad != 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_one_correction_division_by_small_odd_integers_is_ieee(tmp_path):
    src = tmp_path / "div_small.c"
    exe = tmp_path / "div_small"
    src.write_text(SRC)
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-fno-fast-math", str(src), "-o", str(exe), "-lm"], check=True)
    out = subprocess.run([str(exe), "2000"], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("0 mismatches"), out.stdout
