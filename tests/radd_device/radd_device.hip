/*
 * The device build of gm_radd.h against plain device additions, bit for bit.
 *
 * Every thread takes one (f0, inc, n) case, adds inc to f0 n times with the
 * hardware addition, and compares the value and the first step that reaches 1.0
 * with gm_radd_n, gm_radd_until_one and gm_radd_count_to_one; a second kernel
 * compares gm_floor_div
 * with 64-bit integer division.  The cases are drawn as tests/test_radd.py
 * draws them for the host build.  Prints "bad <count>" and exits non-zero on a
 * mismatch.
 *
 * usage: radd_device <cases> <divisions>
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <vector>

#include "gm_radd.h"

#define CHECK(x)                                                                      \
    do {                                                                              \
        hipError_t e_ = (x);                                                          \
        if (e_ != hipSuccess) {                                                       \
            fprintf(stderr, "%s:%d %s: %s\n", __FILE__, __LINE__, #x, hipGetErrorString(e_)); \
            exit(3);                                                                  \
        }                                                                             \
    } while (0)

struct Case {
    double f0, inc;
    uint64_t n;
};

__global__ void k_radd(const Case *cs, int ncase, unsigned *bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= ncase) return;
    const Case c = cs[i];
    double a = c.f0, first = 0.0;
    uint64_t jn = 0;
    for (uint64_t k = 1; k <= c.n; k++) {
        a = a + c.inc;
        if (jn == 0 && a >= 1.0) {
            jn = k;
            first = a;
        }
    }
    const double b = gm_radd_n(c.f0, c.inc, c.n);
    double d = c.f0;
    const uint64_t jd = gm_radd_until_one(d, c.inc, c.n);
    bool ok = gm_dbits(a) == gm_dbits(b) && jn == jd;
    if (ok && jn) ok = gm_dbits(first) == gm_dbits(d);
    /* the count alone, certified from the real sums where it can be */
    if (ok) ok = gm_radd_count_to_one(c.f0, c.inc, c.n) == jn;
    if (!ok) atomicAdd(bad, 1u);
}

__global__ void k_floor_div(const uint64_t *num, const uint64_t *den, int n, unsigned *bad) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double m = gm_floor_div((double)num[i], (double)den[i]);
    if ((uint64_t)m != num[i] / den[i]) atomicAdd(bad, 1u);
}

static uint64_t s = 88172645463325252ull;
static uint64_t nx() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
}
static double ur() { return (double)(nx() >> 11) * (1.0 / 9007199254740992.0); }

static Case draw() {
    double f0, inc;
    const int kind = (int)(nx() % 11);
    if (kind == 0) {
        f0 = ur() * 0.01;
        inc = pow(10.0, -5.0 * ur()) * 0.1;
    } else if (kind == 1) {
        f0 = 0.0;
        inc = ur() * 0.1;
    } else if (kind == 2) {
        f0 = 1.0;
        inc = ur() * 0.1;
    } else if (kind == 3) { /* exact ties in the binade of f0 */
        f0 = ldexp(1.0 + ur(), -(int)(nx() % 12) - 1);
        int e;
        frexp(f0, &e);
        const double u = ldexp(1.0, e - 1 - 52);
        const double q = floor(ur() * 4096.0);
        inc = q * u + 0.5 * u;
    } else if (kind == 4) { /* increments at and below half an ulp */
        f0 = ldexp(1.0 + ur(), -(int)(nx() % 6) - 1);
        int e;
        frexp(f0, &e);
        inc = ldexp(1.0, e - 1 - 52) * (0.25 + 0.5 * ur());
    } else if (kind == 5) { /* increments that are not positive, starts at and above 1.0 */
        f0 = (nx() & 1) ? 1.0 + ur() * 0.01 : ur();
        inc = (nx() % 4 == 0) ? 0.0 : -ur() * 0.01;
    } else if (kind == 6) { /* runs that end on a binade top and on 1.0 itself */
        const int p = 8 + (int)(nx() % 30);
        f0 = ldexp((double)(nx() % (1ull << (p - 2))), -p);
        inc = ldexp((double)(1 + nx() % 7), -p);
    } else if (kind == 7) { /* a last sum that rounds up onto the binade top */
        const int e = -(int)(nx() % 12);
        const double u = ldexp(1.0, e - 53);
        const double q = (double)(1 + nx() % (1u << 20));
        inc = q * u + u * (0.5 + 0.5 * ur());
        f0 = (9007199254740991.0 - q - (double)(nx() % 50) * (q + 1.0)) * u;
    } else if (kind == 9) { /* real sums that reach 1.0 within rounding of an integer count */
        f0 = (nx() & 1) ? 0.0 : ur() * 0.5;
        inc = (1.0 - f0) / (double)(1 + nx() % 2500);
        if (nx() & 1) inc = gm_dfrom(gm_dbits(inc) + (nx() % 5) - 2);
    } else if (kind == 10) { /* the slowest cells: counts around and beyond the cap */
        f0 = ur() * 0.9;
        inc = (1.0 - f0) / (1500.0 + 3000.0 * ur());
    } else { /* carries and increments as the growth produces them */
        f0 = ur() * 0.0068 * ur();
        inc = 0.1 * ur() * (0.1 + 0.9 * ur()) * (0.6494 + 0.3506 * ur());
    }
    return Case{f0, inc, 1 + nx() % 3000};
}

int main(int argc, char **argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: radd_device <cases> <divisions>\n");
        return 2;
    }
    const int ncase = atoi(argv[1]), ndiv = atoi(argv[2]);
    if (ncase < 1 || ndiv < 1) {
        fprintf(stderr, "usage: radd_device <cases> <divisions>\n");
        return 2;
    }
    std::vector<Case> cs(ncase);
    for (auto &c : cs) c = draw();
    /* quotients of integer-valued doubles: numerators below 2^53, divisors in
     * [1, 2^52], a seventh of them on or just below a multiple */
    std::vector<uint64_t> num(ndiv), den(ndiv);
    for (int i = 0; i < ndiv; i++) {
        const int eD = (int)(nx() % 53), eN = (int)(nx() % 54);
        uint64_t D = (nx() & ((1ull << eD) - 1)) | (1ull << eD);
        if (D > (1ull << 52)) D = 1ull << 52;
        uint64_t N = eN ? (nx() & ((1ull << eN) - 1)) : 0;
        if (i % 7 == 0) N = (N / D) * D + ((i % 3 == 0 || N < D) ? 0 : D - 1);
        if (N >= (1ull << 53)) N = (1ull << 53) - 1;
        num[i] = N;
        den[i] = D;
    }
    Case *dc;
    uint64_t *dn, *dd;
    unsigned *dbad, hbad[2] = {0, 0};
    CHECK(hipMalloc(&dc, sizeof(Case) * ncase));
    CHECK(hipMalloc(&dn, sizeof(uint64_t) * ndiv));
    CHECK(hipMalloc(&dd, sizeof(uint64_t) * ndiv));
    CHECK(hipMalloc(&dbad, sizeof(hbad)));
    CHECK(hipMemcpy(dc, cs.data(), sizeof(Case) * ncase, hipMemcpyHostToDevice));
    CHECK(hipMemcpy(dn, num.data(), sizeof(uint64_t) * ndiv, hipMemcpyHostToDevice));
    CHECK(hipMemcpy(dd, den.data(), sizeof(uint64_t) * ndiv, hipMemcpyHostToDevice));
    CHECK(hipMemcpy(dbad, hbad, sizeof(hbad), hipMemcpyHostToDevice));
    k_radd<<<(ncase + 255) / 256, 256>>>(dc, ncase, dbad);
    CHECK(hipGetLastError());
    k_floor_div<<<(ndiv + 255) / 256, 256>>>(dn, dd, ndiv, dbad + 1);
    CHECK(hipGetLastError());
    CHECK(hipMemcpy(hbad, dbad, sizeof(hbad), hipMemcpyDeviceToHost));
    printf("cases %d divisions %d\nbad radd %u floor_div %u\nbad %u\n", ncase, ndiv, hbad[0], hbad[1],
           hbad[0] + hbad[1]);
    CHECK(hipFree(dc));
    CHECK(hipFree(dn));
    CHECK(hipFree(dd));
    CHECK(hipFree(dbad));
    return hbad[0] + hbad[1] ? 1 : 0;
}
