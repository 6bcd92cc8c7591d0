"""The division the kd-formed smoothers use for a face, q / dx^2 from z = RN(1 / dx^2) by
y0 = RN(q z), y1 = RN(y0 + RN(q - b y0) z), y2 = RN(y1 + RN(q - b y1) z) with fused multiply-adds
(Markstein's theorem: Muller et al., Handbook of Floating-Point Arithmetic, 2010, section 5.3),
against IEEE division, bit for bit (CPU; no GPU).

The same sequence is compiled here with gcc (FP contraction off, fma() from libm) and run over
random quotients across 400 binades, dx^2 of non-power-of-two spacings, faces of random kd pairs,
and quotients that sit next to rounding ties (q = b m for short m).  The product alone, and the
corrections with the remainder's sign flipped, are checked to differ somewhere, so the test tells
them apart.  (One correction already matched on every sample; the second is what the theorem needs
when the first product is not faithful, which random samples rarely reach.)  On the
GPU the kernels' faces are compared with the stored-record path's (true divisions) bit for bit in
tests/test_gpu_kd_records.py.
"""
import ctypes
import shutil
import subprocess

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <string.h>
static uint64_t s;
static uint64_t xr(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double rnd(void) {
    uint64_t m = xr() & ((1ull << 52) - 1), e = 1023 - 200 + (xr() % 400), b = (e << 52) | m;
    double d; memcpy(&d, &b, 8); return d;
}
static double div_by(double q, double b, double z, int steps, int flip) {
    double y = q * z, r;
    for (int k = 0; k < steps; k++) { r = flip ? fma(b, y, -q) : fma(-b, y, q); y = fma(r, z, y); }
    return y;
}
long mismatches(long n, int steps, int flip, uint64_t seed) {
    long bad = 0;
    s = seed;
    for (long k = 0; k < n; k++) {
        double b, q;
        if (k % 4 == 0) { double dx = 1.0 + (xr() % 100000) * 0.037; b = dx * dx; } else b = rnd();
        if (k % 3 == 0) { double ka = rnd(), kb = rnd(); q = 0.5 * (ka + kb) * (1e-3 + (xr() % 1000)); }
        else q = rnd();
        if (k % 7 == 0) q = b * (double)(xr() % 4096 + 1) / 64.0;
        double want = q / b, got = div_by(q, b, 1.0 / b, steps, flip);
        if (memcmp(&want, &got, 8)) bad++;
    }
    return bad;
}
"""


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc is not installed")
    d = tmp_path_factory.mktemp("divchk")
    (d / "divchk.c").write_text(SRC)
    so = d / "divchk.so"
    subprocess.run([gcc, "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-o", str(so),
                    str(d / "divchk.c"), "-lm"], check=True)
    L = ctypes.CDLL(str(so))
    L.mismatches.argtypes = [ctypes.c_long, ctypes.c_int, ctypes.c_int, ctypes.c_uint64]
    L.mismatches.restype = ctypes.c_long
    return L


def test_two_corrections_give_the_correctly_rounded_quotient(lib):
    assert lib.mismatches(20_000_000, 2, 0, 88172645463325252) == 0


def test_the_checks_catch_a_missing_or_wrong_correction(lib):
    assert lib.mismatches(2_000_000, 0, 0, 88172645463325252) > 0
    assert lib.mismatches(2_000_000, 2, 1, 88172645463325252) > 0
