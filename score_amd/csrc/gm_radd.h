/*
 * gm_radd.h -- repeated rounded addition f <- fl(f + inc) in few operations.
 *
 * SCORE advances every front cell by rxFrac = rxFrac + inc once per step
 * (SCORE_Kernel.cpp:4092).  With inc fixed, the value after n steps and the
 * first step at which it reaches 1.0 are needed exactly.  Inside one binade
 * [2^e, 2^(e+1)) every double is a multiple of u = 2^(e-52); writing
 * inc = Q*u + R with 0 <= R < u, an addition whose exact sum stays in the binade
 * adds Q*u when 2R < u and (Q+1)*u when 2R > u.  In the tie 2R = u the result is
 * the even neighbour: once f/u is even the step is Q*u for even Q and (Q+1)*u
 * for odd Q.  So a run of additions within a binade is one exact multiply-add;
 * the addition that leaves the binade, and the one tie step that makes f/u
 * even, are done with the hardware addition.  One round below does both, so a
 * walk from f to 1.0 takes about one round per binade crossed.
 *
 * Shared by host and device code; checked bit for bit against the plain loop in
 * tests/test_radd.py.
 */
#ifndef SCORE_GM_RADD_H
#define SCORE_GM_RADD_H

#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define GM_HD __host__ __device__ __forceinline__
#else
#define GM_HD static inline
#endif

GM_HD uint64_t gm_dbits(double x) {
    uint64_t b;
    memcpy(&b, &x, sizeof(b));
    return b;
}

GM_HD double gm_dfrom(uint64_t b) {
    double x;
    memcpy(&x, &b, sizeof(x));
    return x;
}

/* Additions of inc that provably stay inside the binade of f, as a count m and
 * the exact per-addition step `step` (a multiple of the binade's ulp).  m == 0
 * means the next addition must be done by the hardware (it leaves the binade,
 * resolves a tie, or f or inc is outside the closed form's range).  Returns
 * false when every further addition rounds back to f. */
GM_HD bool gm_radd_run(double f, double inc, uint64_t &m, double &step) {
    m = 0;
    step = 0.0;
    const uint64_t be = (gm_dbits(f) >> 52) & 0x7FFu;
    /* zero, negative, tiny or huge values and increments that are not positive */
    if (!(f > 0.0) || be <= 53u || be >= 0x7FEu || !(inc > 0.0)) return true;
    const double top = gm_dfrom((be + 1u) << 52);    /* 2^(e+1) */
    if (!((top - f) > inc)) return true;             /* the sum leaves the binade */
    const double u = gm_dfrom((be - 52u) << 52);     /* ulp of the binade of f */
    /* inc < 2^e here, so inc/u < 2^52; scaling by the power of two 1/u is exact */
    const double inv_u = gm_dfrom((uint64_t)(2098u - be) << 52);
    const double qd = __builtin_floor(inc * inv_u);
    const uint64_t Q = (uint64_t)qd;
    const double R = inc - qd * u;
    const uint64_t F = (uint64_t)(f * inv_u);
    uint64_t D;
    if (R + R < u) D = Q;
    else if (R + R > u) D = Q + 1;
    else {
        if (F & 1u) return true; /* tie with odd f/u: one hardware step makes it even */
        D = Q + (Q & 1u);
    }
    if (D == 0) return false; /* the increment is absorbed by rounding */
    /* steps whose exact sum stays below 2^(e+1): F + i*D + Q + 1 <= 2^53 - 1 */
    const uint64_t lim = (1ull << 53) - 2u;
    if (F + Q >= lim) return true;
    /* m = floor(num/D); num and D are exact doubles and the rounded quotient is
     * below 2^53, so truncating it gives m or m+1 */
    const uint64_t num = lim - F - Q;
    uint64_t k = (uint64_t)((double)num / (double)D);
    if (k * D > num) k--;
    m = k;
    step = (double)D * u;
    return true;
}

/* One round of at most `budget` additions: the in-binade run, then one hardware
 * addition.  Returns the number done, or 0 if f can never change again.  The run
 * never reaches 2^(e+1), so a value below 1.0 can only reach 1.0 in the hardware
 * addition that ends the round. */
GM_HD uint64_t gm_radd_round(double &f, double inc, uint64_t budget) {
    if (budget == 0) return 0;
    uint64_t m;
    double step;
    if (!gm_radd_run(f, inc, m, step)) return 0;
    if (m >= budget) {
        f = f + (double)budget * step;
        return budget;
    }
    if (m) f = f + (double)m * step;
    const double g = f + inc;
    if (m == 0 && gm_dbits(g) == gm_dbits(f)) return 0;
    f = g;
    return m + 1;
}

/* f after n additions of inc */
GM_HD double gm_radd_n(double f, double inc, uint64_t n) {
    while (n) {
        const uint64_t k = gm_radd_round(f, inc, n);
        if (k == 0) break;
        n -= k;
    }
    return f;
}

/* Additions until f first reaches 1.0, at most `cap`.  Returns the count (>= 1)
 * and leaves the reached value in f, or returns 0 when 1.0 is not reached
 * within cap additions. */
GM_HD uint64_t gm_radd_until_one(double &f, double inc, uint64_t cap) {
    if (cap == 0) return 0;
    if (f >= 1.0) { /* already full (a nucleus): the first addition decides */
        const double g = f + inc;
        if (g >= 1.0) {
            f = g;
            return 1;
        }
        return 0; /* inc < 0: the value only decreases from here */
    }
    if (!(inc > 0.0)) return 0;
    uint64_t done = 0;
    while (done < cap) {
        const uint64_t k = gm_radd_round(f, inc, cap - done);
        if (k == 0) return 0;
        done += k;
        if (f >= 1.0) return done;
    }
    return 0;
}

#endif
