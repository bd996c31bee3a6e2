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
 * walk from f to 1.0 takes about one round per binade crossed.  All counts are
 * integer-valued doubles below 2^53, so every step is exact and no 64-bit
 * integer arithmetic is needed.
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

/* parity of an integer-valued double below 2^53 */
GM_HD bool gm_odd(double v) { return v - 2.0 * __builtin_floor(v * 0.5) != 0.0; }

/* floor(num / D) for integer-valued doubles 0 <= num < 2^53 and 1 <= D < 2^53.
 * On the device the division sequence is replaced by the hardware reciprocal
 * refined by two Newton steps; the estimate is then within one of the quotient,
 * and the remainder num - m*D, an integer of magnitude below 2D and so exact in
 * one fma, moves it to the floor.  The host divides. */
#if defined(__HIP_DEVICE_COMPILE__)
GM_HD double gm_floor_div(double num, double D) {
    const double r0 = __builtin_amdgcn_rcp(D);
    const double r1 = __builtin_fma(r0, __builtin_fma(-D, r0, 1.0), r0);
    const double r2 = __builtin_fma(r1, __builtin_fma(-D, r1, 1.0), r1);
    double m = __builtin_floor(num * r2);
    double r = __builtin_fma(-m, D, num);
    while (r < 0.0) {
        m -= 1.0;
        r += D;
    }
    while (r >= D) {
        m += 1.0;
        r -= D;
    }
    return m;
}
#else
GM_HD double gm_floor_div(double num, double D) {
    /* the rounded quotient lies below 2^53, so its floor is the true floor or
     * one more */
    double m = __builtin_floor(num / D);
    if (m * D > num) m -= 1.0;
    return m;
}
#endif

/* One round of at most `budget` additions: the run that provably stays inside
 * the binade of f, as one exact multiply-add, then one hardware addition (which
 * leaves the binade, or makes f/u even in a tie).  Counts are integer-valued
 * doubles (exact below 2^53).  Returns the number of additions done, or 0 if f
 * can never change again.  A run may end exactly on 2^(e+1) (the rounded sum
 * of its last addition); the round then ends there, so a value below 1.0
 * reaches 1.0 only as the last addition of a round. */
GM_HD double gm_radd_round(double &f, double inc, double budget) {
    const uint64_t be = (gm_dbits(f) >> 52) & 0x7FFu;
    /* zero, negative, tiny or huge values and increments that are not positive
     * take plain hardware steps */
    const bool plain = !(f > 0.0) || be <= 53u || be >= 0x7FEu || !(inc > 0.0);
    const uint64_t bs = plain ? 1023u : be; /* keeps the bit patterns below valid */
    const double top = gm_dfrom((bs + 1u) << 52);                /* 2^(e+1) */
    /* the exact sum f + inc stays in the binade; then inc < 2^e and inc/u < 2^52 */
    const bool inside = !plain && (top - f) > inc;
    if (!inside) {
        const double g = f + inc;
        if (gm_dbits(g) == gm_dbits(f)) return 0.0;
        f = g;
        return 1.0;
    }
    const double u = gm_dfrom((bs - 52u) << 52);                 /* ulp of the binade */
    const double inv_u = gm_dfrom((uint64_t)(2098u - bs) << 52); /* exact power of two */
    const double qd = __builtin_floor(inc * inv_u);
    const double R = inc - qd * u;
    const double F = f * inv_u;
    const double twoR = R + R;
    /* D*u is the exact change of one addition inside the binade */
    double D;
    if (twoR == u) { /* a tie (rare): rounding to even decides */
        /* F = f/u lies in [2^52, 2^53), so its parity is the last mantissa bit of f */
        if (gm_dbits(f) & 1u) { /* one hardware step makes f/u even */
            f = f + inc;
            return 1.0;
        }
        D = qd + (gm_odd(qd) ? 1.0 : 0.0);
    } else {
        D = twoR < u ? qd : qd + 1.0;
    }
    if (D == 0.0) return 0.0; /* the increment is absorbed by rounding */
    /* addition i of the run keeps its exact sum below 2^(e+1) when
     * F + (i-1)*D + Q + 1 <= 2^53, so the run takes floor(num/D) + 1 additions
     * with num = 2^53 - 1 - F - Q >= 0 (exact). */
    const double num = (9007199254740991.0 - F) - qd;
    const double m = gm_floor_div(num, D) + 1.0;
    const double step = D * u;
    if (m >= budget) {
        f = f + budget * step;
        return budget;
    }
    f = f + m * step;
    if (f == top) return m;
    f = f + inc;
    return m + 1.0;
}

/* Additions done one by one before the closed form takes over: a round costs
 * about as many instructions as a dozen plain additions.  Short replays are
 * added up directly, and so are the first additions of a value below 1/4,
 * whose binades a round crosses a few additions at a time (a new cell starts
 * near 0 with an increment of up to 0.1). */
#define GM_RADD_PLAIN 16u
#define GM_RADD_LOW 0.25

/* f after n additions of inc */
GM_HD double gm_radd_n(double f, double inc, uint64_t n) {
    if (n <= GM_RADD_PLAIN) {
        for (uint32_t i = 0; i < (uint32_t)n; i++) f = f + inc;
        return f;
    }
    uint32_t i = 0;
    for (; i < GM_RADD_PLAIN && f < GM_RADD_LOW; i++) f = f + inc;
    double left = (double)(n - i);
    while (left > 0.0) {
        const double k = gm_radd_round(f, inc, left);
        if (k == 0.0) break;
        left -= k;
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
    double g = f;
    const uint64_t plim = cap < GM_RADD_PLAIN ? cap : GM_RADD_PLAIN;
    uint64_t lim = 0;
    while (lim < plim && g < GM_RADD_LOW) {
        g = g + inc;
        lim++;
        if (g >= 1.0) {
            f = g;
            return lim;
        }
    }
    if (lim == cap) return 0;
    const double capd = (double)(cap - lim);
    double done = 0.0;
    while (done < capd) {
        const double k = gm_radd_round(g, inc, capd - done);
        if (k == 0.0) return 0;
        done += k;
        if (g >= 1.0) {
            f = g;
            return (uint64_t)done + lim;
        }
    }
    return 0;
}

/* an estimate of num / den (den > 0), exact or nearly: the device refines the
 * hardware reciprocal once */
#if defined(__HIP_DEVICE_COMPILE__)
GM_HD double gm_quot_estimate(double num, double den) {
    const double r0 = __builtin_amdgcn_rcp(den);
    return num * __builtin_fma(r0, __builtin_fma(-den, r0, 1.0), r0);
}
#else
GM_HD double gm_quot_estimate(double num, double den) { return num / den; }
#endif

/* The count gm_radd_until_one returns, without the value reached.  While the
 * additions stay below 1.0, each rounds by at most half an ulp of a value below
 * 1.0, 2^-54, so the n-th one lies within n * 2^-54 of the real sum
 * S(n) = f + n*inc.  Hence, for the estimate n = ceil((1 - f) / inc):
 * S(n-1) + (n-1) * 2^-54 < 1 puts every addition up to the (n-1)-th below 1.0
 * (the exact sum of the next one, below 1 - 2^-54, rounds below 1.0), and
 * S(n) - (n-1) * 2^-54 >= 1 puts the exact sum of the n-th at or above 1.0, so
 * n is the count.  The sums are formed here from non-negative terms below 2 in
 * four roundings of at most 2^-53 each, which the margin 2^-50 covers.  When
 * either test fails (1.0 lies within about n * 2^-54 of a real sum, or the
 * estimate is off by one), or for a negative start or an increment of 1/2 or
 * more, the rounds decide. */
GM_HD uint64_t gm_radd_count_to_one(double f, double inc, uint64_t cap) {
    if (cap == 0) return 0;
    if (f >= 1.0) return f + inc >= 1.0 ? 1 : 0;
    if (!(inc > 0.0)) return 0;
    if (f >= 0.0 && inc < 0.5) {
        const double e = 0x1p-54, M = 0x1p-50, capd = (double)cap;
        double n = __builtin_ceil(gm_quot_estimate(1.0 - f, inc));
        if (!(n >= 1.0)) n = 1.0;
        if (n > capd) {
            /* the cap-th addition still lies below 1.0 */
            if (((f + capd * inc) + capd * e) + M < 1.0) return 0;
        } else {
            const double m = n - 1.0;
            if (((f + m * inc) + m * e) + M < 1.0 && ((f + n * inc) - m * e) - M >= 1.0) return (uint64_t)n;
        }
    }
    double g = f;
    return gm_radd_until_one(g, inc, cap);
}

#endif
