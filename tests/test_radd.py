"""The closed-form repeated rounded addition (score_amd/csrc/gm_radd.h) against the
plain loop f = f + inc, bit for bit, over random and adversarial cases: ties at
every binade, increments below half an ulp, starts at 0 and at 1.0, and the
first-reaching-1.0 count that decides a front cell's completion step, sums that
land exactly on a binade top or on 1.0; increments that are zero or negative fall
back to single additions."""
import os
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

HARNESS = r"""
#include <stdio.h>
#include <stdlib.h>
#include <math.h>
#include "gm_radd.h"

static uint64_t s = 88172645463325252ull;
static uint64_t nx(void) { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static double ur(void) { return (double)(nx() >> 11) * (1.0 / 9007199254740992.0); }

static int check(double f0, double inc, uint64_t n) {
    /* n additions */
    double a = f0;
    for (uint64_t i = 0; i < n; i++) a = a + inc;
    double b = gm_radd_n(f0, inc, n);
    if (gm_dbits(a) != gm_dbits(b)) { printf("N %a %a %llu: %a vs %a\n", f0, inc, (unsigned long long)n, a, b); return 1; }
    /* first reaching 1.0 within n */
    double c = f0; uint64_t jn = 0;
    for (uint64_t i = 1; i <= n; i++) { c = c + inc; if (c >= 1.0) { jn = i; break; } }
    double d = f0;
    uint64_t jd = gm_radd_until_one(d, inc, n);
    if (jn != jd || (jn && gm_dbits(c) != gm_dbits(d))) {
        printf("U %a %a %llu: %llu %a vs %llu %a\n", f0, inc, (unsigned long long)n, (unsigned long long)jn, c,
               (unsigned long long)jd, d);
        return 1;
    }
    /* the count alone, certified from the real sums where it can be */
    uint64_t jc = gm_radd_count_to_one(f0, inc, n);
    if (jc != jn) {
        printf("C %a %a %llu: %llu vs %llu\n", f0, inc, (unsigned long long)n, (unsigned long long)jn,
               (unsigned long long)jc);
        return 1;
    }
    return 0;
}

int main(int argc, char **argv) {
    long cases = atol(argv[1]);
    int bad = 0;
    for (long k = 0; k < cases && bad < 5; k++) {
        double f0, inc;
        int kind = (int)(nx() % 11);
        if (kind == 0) { f0 = ur() * 0.01; inc = pow(10.0, -5.0 * ur()) * 0.1; }
        else if (kind == 1) { f0 = 0.0; inc = ur() * 0.1; }
        else if (kind == 2) { f0 = 1.0; inc = ur() * 0.1; }
        else if (kind == 3) {            /* exact ties in the binade of f0 */
            f0 = ldexp(1.0 + ur(), -(int)(nx() % 12) - 1);
            int e; frexp(f0, &e);
            double u = ldexp(1.0, e - 1 - 52);
            double q = floor(ur() * 4096.0);
            inc = q * u + 0.5 * u;
            if (inc == 0.0) inc = 0.5 * u;
        } else if (kind == 4) {          /* increments at and below half an ulp */
            f0 = ldexp(1.0 + ur(), -(int)(nx() % 6) - 1);
            int e; frexp(f0, &e);
            double u = ldexp(1.0, e - 1 - 52);
            inc = u * (0.25 + 0.5 * ur());
        } else if (kind == 5) {          /* increments that are not positive, starts at and above 1.0 */
            f0 = (nx() & 1) ? 1.0 + ur() * 0.01 : ur();
            inc = (nx() % 4 == 0) ? 0.0 : -ur() * 0.01;
        } else if (kind == 6) {          /* exact sums: runs that end on a binade top and on 1.0 itself */
            int p = 8 + (int)(nx() % 30);
            f0 = ldexp((double)(nx() % (1ull << (p - 2))), -p);
            inc = ldexp((double)(1 + nx() % 7), -p);
        } else if (kind == 7) {          /* a run whose last sum rounds up onto the binade top, 1.0 included */
            int e = -(int)(nx() % 12);
            double u = ldexp(1.0, e - 53);
            double q = (double)(1 + nx() % (1u << 20));
            inc = q * u + u * (0.5 + 0.5 * ur());
            double F0 = 9007199254740991.0 - q - (double)(nx() % 50) * (q + 1.0);
            f0 = F0 * u;
        } else if (kind == 9) {          /* real sums that reach 1.0 within rounding of an integer count */
            f0 = (nx() & 1) ? 0.0 : ur() * 0.5;
            inc = (1.0 - f0) / (double)(1 + nx() % 2500);
            if (nx() & 1) inc = gm_dfrom(gm_dbits(inc) + (nx() % 5) - 2);
        } else if (kind == 10) {         /* the slowest cells: counts around and beyond the cap */
            f0 = ur() * 0.9;
            inc = (1.0 - f0) / (1500.0 + 3000.0 * ur());
        } else {                         /* carries and increments as the growth produces them */
            f0 = ur() * 0.0068 * ur();
            inc = 0.1 * ur() * (0.1 + 0.9 * ur()) * (0.6494 + 0.3506 * ur());
        }
        uint64_t n = 1 + nx() % 3000;
        bad += check(f0, inc, n);
    }
    printf("bad %d\n", bad);
    return bad != 0;
}
"""


def test_closed_form_matches_loop():
    cc = shutil.which("g++") or shutil.which("c++")
    if not cc:
        pytest.skip("no C++ compiler")
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "h.cpp")
        exe = os.path.join(d, "h")
        open(src, "w").write(HARNESS)
        subprocess.run([cc, "-O2", "-std=c++17", "-ffp-contract=off", "-I", os.path.join(ROOT, "score_amd", "csrc"),
                        src, "-o", exe, "-lm"], check=True)
        out = subprocess.run([exe, "200000"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout[-2000:]
    assert "bad 0" in out.stdout
