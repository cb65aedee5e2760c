"""The FP64 transforms divide by sqrt(2) with a product and two fma
corrections instead of a division (transforms.hip, div_sqrt2).  The same
sequence, compiled for the host with gcc, must round exactly as IEEE
division does: random significands over many binades, both signs, values next
to powers of two, and the boundaries of the guarded range."""
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
static const double S = 1.4142135623730951, Y = 0x1.6a09e667f3bccp-1;
static double div_sqrt2(double x) {
    double ax = fabs(x);
    if (!(ax >= 0x1p-960 && ax <= 0x1p1000)) return x / S;
    double q0 = x * Y, r0 = fma(-q0, S, x), q1 = fma(r0, Y, q0), r1 = fma(-q1, S, x);
    return fma(r1, Y, q1);
}
static uint64_t st = 12345;
static uint64_t next(void) {
    uint64_t z = (st += 0x9e3779b97f4a7c15ull);
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
static long bad = 0;
static void check(double x) {
    double a = div_sqrt2(x), b = x / S;
    if (memcmp(&a, &b, 8)) { if (bad < 5) printf("mismatch %a: %a vs %a\n", x, a, b); bad++; }
}
int main(int argc, char **argv) {
    long n = atol(argv[1]);
    if (Y != 1.0 / S) { printf("bad reciprocal\n"); return 2; }
    for (long i = 0; i < n; i++) {
        uint64_t r = next(), m = next() & ((1ull << 52) - 1);
        int e = (int)(r % 2046) - 1022;              /* every binade */
        uint64_t bits = ((uint64_t)(e + 1023) << 52) | m | ((r >> 63) << 63);
        double x;
        memcpy(&x, &bits, 8);
        check(x);
    }
    for (int e = -1074; e <= 1023; e++) {          /* powers of two and neighbours */
        double p = ldexp(1.0, e);
        check(p); check(-p); check(nextafter(p, 0)); check(nextafter(p, INFINITY));
    }
    check(0.0); check(-0.0); check(INFINITY); check(-INFINITY);
    check(0x1p-960); check(nextafter(0x1p-960, 0)); check(0x1p1000); check(nextafter(0x1p1000, INFINITY));
    printf("%ld mismatches\n", bad);
    return bad != 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_div_sqrt2_rounds_as_division(tmp_path):
    c = tmp_path / "d.c"
    c.write_text(SRC)
    exe = tmp_path / "d"
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", str(exe), str(c), "-lm"], check=True)
    out = subprocess.run([str(exe), "20000000"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout
    assert "0 mismatches" in out.stdout
