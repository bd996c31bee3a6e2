/*
 * gm_device.h -- device helpers shared by the growth machine's kernels
 * (gm_kernels.hip: one memory region; gm_regions.hip: several).
 *
 * Everything here has internal linkage and is compiled into each file that
 * includes it.  FP64 throughout, compiled with -ffp-contract=off, so every
 * product and sum is rounded exactly as in the reference's build without FMA.
 */
#ifndef SCORE_GM_DEVICE_H
#define SCORE_GM_DEVICE_H

#include "gm_internal.h"
#include "gm_radd.h"

#define GM_RH 2
#define GM_RHO_RX (1.0e11)
#define GM_DOUBLE_ACCURACY (5.0e-16)

/* velocity shape factors per infector code (SCORE_Kernel.cpp:4051-4053) */
static __constant__ double c_shape[27] = {0.6494, 0.7957, 0.6494, 0.7957, 1.0000, 0.7957, 0.6494, 0.7957, 0.6494,
                                   0.7957, 1.0000, 0.7957, 1.0000, 1.0000, 0.7957, 1.0000, 0.7957,
                                   0.6494, 0.7957, 0.6494, 0.7957, 1.0000, 0.7957, 0.6494, 0.7957, 0.6494, 1.0};

/* the 26 attempts by rank = call order of SCORE_Kernel.cpp:4611-4638:
 * offsets, infector code and overshoot class (1 face 0.0682, 2 edge 0.0542, 3 corner 0.0484) */
static __constant__ int8_t c_dx[26] = {-1, 0, 1, -1, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1};
static __constant__ int8_t c_dy[26] = {-1, -1, -1, 0, 0, 1, 1, 1, 1, 1, 1, -1, -1, -1, 0, 0, 0, -1, -1, -1, 0, 0, 0, 1, 1, 1};
static __constant__ int8_t c_dz[26] = {0, 0, 0, 0, 0, 0, 0, 0, -1, -1, -1, -1, -1, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
static __constant__ uint8_t c_code[26] = {16, 15, 14, 13, 12, 11, 10, 9, 19, 18, 17, 25, 24, 23, 22, 21, 20,
                                   8, 7, 6, 5, 4, 3, 2, 1, 0};
static __constant__ uint8_t c_cls[26] = {2, 1, 2, 1, 1, 2, 1, 2, 3, 2, 3, 3, 2, 3, 2, 1, 2, 3, 2, 3, 2, 1, 2, 3, 2, 3};

static __device__ __forceinline__ uint32_t lanemask_lt_count(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
}

/* lane j's value of v (j uniform); unsigned, so the 64-bit form does not
 * sign-extend the low word */
static __device__ __forceinline__ uint32_t rdlane(uint32_t v, uint32_t j) {
    return (uint32_t)__builtin_amdgcn_readlane((int)v, (int)j);
}
static __device__ __forceinline__ uint64_t rdlane64(uint64_t v, uint32_t j) {
    return ((uint64_t)rdlane((uint32_t)(v >> 32), j) << 32) | (uint64_t)rdlane((uint32_t)v, j);
}

static __device__ __forceinline__ bool is_pending(double v) {
    return ((unsigned long long)__double_as_longlong(v) & GM_PENDING_MASK) == GM_PENDING_TAG;
}

static __device__ __forceinline__ double pending_box(uint32_t h) {
    return __longlong_as_double((long long)(GM_PENDING_TAG | (unsigned long long)h));
}

static __device__ __forceinline__ uint32_t hash_slot(unsigned long long k, uint32_t cap) {
    k ^= k >> 33;
    k *= 0xff51afd7ed558ccdull;
    k ^= k >> 33;
    k *= 0xc4ceb9fe1a85ec53ull;
    k ^= k >> 33;
    return (uint32_t)k & (cap - 1u);
}

/* largest |component| over the six Grimmer candidates of q_def^-1 * q_rx
 * (SCORE_Math.cpp:230-239, :558-608); only + - * fabs and compares, so the
 * value is bit-identical to the host's. */
static __device__ double grimmer_omega(const double *p, const double *q, double fac) {
    double m0 = q[0], m1 = -1.0 * q[1], m2 = -1.0 * q[2], m3 = -1.0 * q[3];
    double a = m0 * p[0] - m1 * p[1] - m2 * p[2] - m3 * p[3];
    double b = p[0] * m1 + m0 * p[1] + (m2 * p[3] - m3 * p[2]);
    double c = p[0] * m2 + m0 * p[2] + (-(m1 * p[3]) + m3 * p[1]);
    double d = p[0] * m3 + m0 * p[3] + (m1 * p[2] - m2 * p[1]);
    double v[24];
    v[0] = (a + b) * fac; v[1] = (a - b) * fac; v[2] = (c + d) * fac; v[3] = (c - d) * fac;
    v[4] = (a + c) * fac; v[5] = (a - c) * fac; v[6] = (b + d) * fac; v[7] = (b - d) * fac;
    v[8] = (a + d) * fac; v[9] = (a - d) * fac; v[10] = (b + c) * fac; v[11] = (b - c) * fac;
    v[12] = (a + b + c + d) * 0.5; v[13] = (a + b - c - d) * 0.5; v[14] = (a - b + c - d) * 0.5; v[15] = (a - b - c + d) * 0.5;
    v[16] = (a + b + c - d) * 0.5; v[17] = (a + b - c + d) * 0.5; v[18] = (a - b + c + d) * 0.5; v[19] = (a - b - c - d) * 0.5;
    v[20] = a; v[21] = b; v[22] = c; v[23] = d;
    double omega = 0.0;
#pragma unroll
    for (int i = 0; i < 24; i++)
        if (fabs(v[i]) > omega) omega = fabs(v[i]);
    return omega;
}

/* Mobility weight of (nucleus rxg, deformed grain g).  Rollett-Holm pairs whose
 * disorientation is far above the LAGB/HAGB transition have P == 1.0 exactly in
 * glibc arithmetic (threshold set and checked by the host); every other pair is
 * looked up in the cache, and a miss inserts a pending entry the host resolves
 * with glibc before the next sweep reads it. */
static __device__ double pair_weight(const GmDev &d, const GmPParams &pp, uint32_t rxg, uint32_t g) {
    if (pp.model == GM_RH && pp.omega_hi >= 0.0) {
        double om = grimmer_omega(&d.rxq[4 * (size_t)rxg], &d.defq[4 * (size_t)g], pp.fac);
        if (om <= pp.omega_hi) {
            d.st->pone = 1u;
            return 1.0;
        }
    }
    unsigned long long key = (unsigned long long)rxg * d.ndefg + g + 1ull;
    uint32_t h = hash_slot(key, d.hcap);
    const uint32_t nprobe = d.hcap < GM_PROBE_MAX ? d.hcap : GM_PROBE_MAX;
    for (uint32_t probe = 0; probe < nprobe; probe++) {
        unsigned long long k = __atomic_load_n(&d.hkeys[h], __ATOMIC_RELAXED);
        if (k == 0ull) {
            unsigned long long old = atomicCAS(&d.hkeys[h], 0ull, key);
            if (old == 0ull) {
                uint32_t i = atomicAdd(&d.st->npend, 1u);
                d.pend_slot[i] = h;
                d.pend_key[i] = key;
                if (i < GM_PEND_INLINE) d.st->pend_keys[i] = key;
                return pending_box(h);
            }
            k = old;
        }
        if (k == key) {
            double v = d.hvals[h];
            return is_pending(v) ? pending_box(h) : v;
        }
        h = (h + 1u) & (d.hcap - 1u);
    }
    /* no room near the key's home (every key lies within nprobe slots of it, so
     * the key is not cached): the cell waits, marked as spilled, until the host
     * has grown the cache and its lookup is repeated */
    atomicAdd(&d.st->novf, 1u);
    return pending_box(GM_OVF_SLOT);
}

static __device__ __forceinline__ bool is_spilled(double v) {
    return (unsigned long long)__double_as_longlong(v) == (GM_PENDING_TAG | (unsigned long long)GM_OVF_SLOT);
}


static unsigned fill_blocks(size_t n) {
    size_t b = (n + GM_BLOCK - 1) / GM_BLOCK;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return (unsigned)b;
}


static __device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

/* ------------------------------------------------------------ front records */

/* the record's 16-byte quarters: A = (frac0, s0, inf), B = (inc, spare),
 * C = (P, cell, defg), then rxg */
struct RecA {
    double frac0;
    uint32_t s0, inf;
};
struct RecC {
    double P;
    uint32_t cell, defg;
};
static __device__ __forceinline__ RecA rec_a(const GmRec *r) {
    const uint4 v = *reinterpret_cast<const uint4 *>(&r->frac0);
    return RecA{__hiloint2double((int)v.y, (int)v.x), v.z, v.w};
}
static __device__ __forceinline__ RecC rec_c(const GmRec *r) {
    const uint4 v = *reinterpret_cast<const uint4 *>(&r->P);
    return RecC{__hiloint2double((int)v.y, (int)v.x), v.z, v.w};
}
static __device__ __forceinline__ void rec_set_a(GmRec *r, double frac0, uint32_t s0, uint32_t inf) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(frac0);
    *reinterpret_cast<uint4 *>(&r->frac0) = make_uint4((uint32_t)b, (uint32_t)(b >> 32), s0, inf);
}
static __device__ __forceinline__ void rec_set_b(GmRec *r, double inc) {
    *reinterpret_cast<double2 *>(&r->inc) = make_double2(inc, 0.0);
}
static __device__ __forceinline__ void rec_set_c(GmRec *r, double P, uint32_t cell, uint32_t defg) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(P);
    *reinterpret_cast<uint4 *>(&r->P) = make_uint4((uint32_t)b, (uint32_t)(b >> 32), cell, defg);
}

/* a mobility weight with its pending marker replaced by the cached value; a
 * spilled lookup has no cache slot yet and stays marked */
static __device__ __forceinline__ double resolved(const GmDev &d, double P) {
    return is_pending(P) && !is_spilled(P) ? d.hvals[(uint32_t)__double_as_longlong(P)] : P;
}

/* ------------------------------------------------------------------- sweep */

/* rxFrac increment of one cell: SCORE's v = m*p; v *= shape; rxFrac += (v*dt)*inv_dx
 * (SCORE_Kernel.cpp:4077-4092, :5459-5470); returns false when p <= DOUBLE_ACCURACY
 * (the cell is not advanced) */
static __device__ __forceinline__ bool cell_increment(const GmDev &d, const GmSweepParams &sp, double P, uint32_t defg,
                                                      uint32_t inf, double &inc) {
    double m;
    if (sp.model == GM_RH) m = P * sp.mRH;
    else if (P < 0.0) m = sp.mLAGB;
    else m = (P * sp.mGS) + ((1.0 - P) * sp.mHAGB);
    double p = (sp.Gbhalfsq * (d.rho[defg] - GM_RHO_RX)) - sp.zener;
    if (!(p > GM_DOUBLE_ACCURACY)) return false;
    double v = m * p;
    v *= c_shape[inf];
    inc = (v * sp.dt) * sp.inv_dx;
    return true;
}

/* Completion step of a cell with value f before step k: the first step >= k
 * in which repeated rounded additions of its increment reach 1.0 (GM_STEP_NONE
 * when none up to the step cap does).  Sets the increment (0.0 for a cell SCORE
 * does not advance).  The value after the completing addition, whose excess
 * over 1.0 the infection carries on, is replayed from the anchor when the cell
 * completes and wins a target (parent_fin). */
static __device__ __forceinline__ uint32_t predict_values(const GmDev &d, const GmSweepArgs &A, uint32_t k, double f,
                                                          double P, uint32_t defg, uint32_t inf, double &inc) {
    uint32_t cs = GM_STEP_NONE;
    if (cell_increment(d, A.sp, P, defg, inf, inc)) {
        if (k <= A.stepcap && A.next_only) {
            /* the step's one addition; a cell it leaves below 1.0 is re-anchored
             * and predicted again before the next step (GM_STEP_FAR_NEXT) */
            cs = f + inc >= 1.0 ? k : GM_STEP_NEXT;
        } else if (k <= A.stepcap) {
            const uint64_t j = gm_radd_count_to_one(f, inc, (uint64_t)(A.stepcap - k) + 1u);
            if (j) cs = k + (uint32_t)(j - 1u);
        }
    } else {
        inc = 0.0;
    }
    return cs;
}

/* rxFrac of a parent after its completing addition in step k: n = k - s0 + 1
 * additions of its increment to the value it was anchored with before step s0 */
static __device__ __forceinline__ double parent_fin(const GmDev &d, uint32_t slot, uint32_t k) {
    const GmRec *r = d.rec + slot;
    const RecA a = rec_a(r);
    const double inc = r->inc;
    const uint32_t n = k - a.s0 + 1u;
    return n == 1u ? a.frac0 + inc : gm_radd_n(a.frac0, inc, n);
}

/* the stored code of completion step cs predicted at step kp */
static __device__ __forceinline__ gm_cs_t cs_code(const GmDev &d, uint32_t cs, uint32_t kp) {
    if (cs == GM_STEP_NONE) return (gm_cs_t)GM_CS_NEVER;
    if (cs == GM_STEP_NEXT) return (gm_cs_t)GM_CS_FAR;
    if (cs - kp >= d.horizon) return (gm_cs_t)GM_CS_FAR;
    return (gm_cs_t)(cs & (GM_CS_MOD - 1u));
}

/* Whether code cs marks an INACTIVE slot at step code kc (a step mod GM_CS_MOD):
 * the INACTIVE sentinel, or the code of a step before kc.  A cell that completes
 * keeps the code of its completing step, and from the next step on that code
 * lies in the past, so a retiring cell's slot is not written.  A live cell's
 * code names a step in [kc, kc + horizon).  A past code is rewritten to
 * INACTIVE by the next re-anchoring, which comes at the latest at the next
 * multiple of the horizon, before the code could wrap into that range
 * (horizon <= GM_CS_MOD / 2). */
static __device__ __forceinline__ bool cs_inactive(uint32_t cs, uint32_t kc, uint32_t horizon) {
    return cs == GM_CS_INACTIVE || (cs < GM_CS_MOD && ((cs - kc) & (GM_CS_MOD - 1u)) >= horizon);
}

/* Predicts slot s from its value f before step A.k (its new anchor): rewrites
 * the record's first sector (anchor, increment) and stores the completion
 * step's code. */
static __device__ __forceinline__ void predict_slot(const GmDev &d, const GmSweepArgs &A, uint32_t s, double f,
                                                    uint32_t s0, double P, uint32_t defg, uint32_t inf) {
    double inc;
    const uint32_t cs = predict_values(d, A, A.k, f, P, defg, inf, inc);
    rec_set_a(d.rec + s, f, s0, inf);
    rec_set_b(d.rec + s, inc);
    d.cstep[s] = cs_code(d, cs, A.k);
}

/* the arguments of the step the device is at */
static __device__ __forceinline__ GmSweepArgs step_args(const GmDev &d, const GmSweepArgs &A) {
    GmSweepArgs B = A;
    B.k = d.st->kcur;
    return B;
}


/* static partition of `n` items over the grid in multiples of `q` */
static __device__ __forceinline__ void part_of(uint32_t n, uint32_t q, uint32_t &lo, uint32_t &hi) {
    uint32_t per = (n + gridDim.x - 1) / gridDim.x;
    per = (per + q - 1) / q * q;
    lo = blockIdx.x * per;
    hi = lo + per;
    if (lo > n) lo = n;
    if (hi > n) hi = n;
}


/* exclusive scan, in place, of C rows of nb <= GM_MAX_BLOCKS per-block counts by
 * one 1024-thread block; tot[c] receives the totals.  Each thread owns 8
 * consecutive entries of every row, loaded and stored as two 16-byte vectors. */
static_assert(GM_MAX_BLOCKS == 8 * 1024, "scan layout");
template <int C>
static __device__ __forceinline__ void scan_block_counts(uint32_t *bcnt, uint32_t nb, uint32_t *tot) {
    __shared__ uint32_t ws[16][C];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t i0 = threadIdx.x * 8u;
    uint32_t v[C][8], sum[C], incl[C];
#pragma unroll
    for (int c = 0; c < C; c++) {
        const uint4 *p = reinterpret_cast<const uint4 *>(bcnt + c * GM_MAX_BLOCKS + i0);
        const uint4 a = p[0], b = p[1];
        v[c][0] = a.x; v[c][1] = a.y; v[c][2] = a.z; v[c][3] = a.w;
        v[c][4] = b.x; v[c][5] = b.y; v[c][6] = b.z; v[c][7] = b.w;
    }
#pragma unroll
    for (int c = 0; c < C; c++) {
        sum[c] = 0;
#pragma unroll
        for (uint32_t j = 0; j < 8; j++) {
            if (i0 + j >= nb) v[c][j] = 0;
            sum[c] += v[c][j];
        }
        incl[c] = wave_incl_scan(sum[c]);
        if (lane == 63) ws[w][c] = incl[c];
    }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < C; c++) {
        uint32_t pre = 0, t = 0;
        for (int i = 0; i < 16; i++) {
            if (i < (int)w) pre += ws[i][c];
            t += ws[i][c];
        }
        tot[c] = t;
        uint32_t run = pre + incl[c] - sum[c];
        uint32_t o[8];
#pragma unroll
        for (uint32_t j = 0; j < 8; j++) {
            o[j] = run;
            run += v[c][j];
        }
        uint4 *q = reinterpret_cast<uint4 *>(bcnt + c * GM_MAX_BLOCKS + i0);
        q[0] = make_uint4(o[0], o[1], o[2], o[3]);
        q[1] = make_uint4(o[4], o[5], o[6], o[7]);
    }
}


static uint32_t grid_for(uint32_t items, uint32_t per_block) {
    uint32_t b = (items + per_block - 1) / per_block;
    if (b > GM_MAX_BLOCKS) b = GM_MAX_BLOCKS;
    if (b < 1) b = 1;
    return b;
}

static inline hipError_t mark(hipEvent_t *ev, int i, hipStream_t s) {
    return ev ? hipEventRecord(ev[i], s) : hipSuccess;
}


/* word and bit of cell (x,y,z) in the deformed-cell bitmap */
static __device__ __forceinline__ void dmap_loc(const GmDev &d, uint32_t x, uint32_t y, uint32_t z, uint32_t &word,
                                                uint32_t &bit) {
    const uint32_t line = (x >> 4) + d.dnbx * ((y >> 3) + d.dnby * (z >> 3));
    const uint32_t b = (x & 15u) | ((y & 7u) << 4) | ((z & 7u) << 7);
    word = line * 32u + (b >> 5);
    bit = 1u << (b & 31u);
}

static __device__ __forceinline__ bool is_deformed(const GmDev &d, uint32_t x, uint32_t y, uint32_t z) {
    uint32_t w, b;
    dmap_loc(d, x, y, z, w, b);
    return (d.dmap[w] & b) != 0u;
}

static __device__ __forceinline__ void clear_deformed(const GmDev &d, uint32_t x, uint32_t y, uint32_t z) {
    uint32_t w, b;
    dmap_loc(d, x, y, z, w, b);
    atomicAnd(&d.dmap[w], ~b);
}


/* plane GM_PAR_* of the parent array */
static __device__ __forceinline__ uint4 *par_plane(const GmDev &d, uint32_t plane) {
    return d.par + (size_t)plane * d.parentCap;
}

static __device__ __forceinline__ uint32_t cell_of(const GmDev &d, uint32_t x, uint32_t y, uint32_t z) {
    return x + y * d.nx + z * d.nxy;
}


/* Parents are cut into chunks of GM_PCHUNK for the decide and commit passes, so
 * that a 64-lane wave takes a whole chunk. */
#define GM_PCHUNK 64
#ifndef GM_PG
#define GM_PG 4                       /* lanes per parent in the propose pass (4, 8 or 16) */
#endif
#define GM_PG_ITERS ((26 + GM_PG - 1) / GM_PG)
static_assert(GM_PG == 4 || GM_PG == 8 || GM_PG == 16, "lanes per parent");

/* The new record of a won attempt of rank r on target t = (tx, ty, tz) in slot
 * gslot: the carry of the parent's overshoot by the direction's class
 * (updateFullRX*, SCORE_Kernel.cpp:4568-4599), the parent's weight unless the
 * target lies in another deformed grain (:4297-4302), anchored at step k+1;
 * the cell leaves the deformed bitmap.  Its increment and completion step are
 * predicted at once under the step's scalars (a later step whose scalars differ
 * re-anchors every cell).  A weight the host has not evaluated yet (a pair met
 * for the first time in this launch window) leaves the cell DEFERRED: it cannot
 * complete before the window ends, and is predicted then.  fresh_weight: the
 * weight is looked up whatever the deformed grain, as the halo synchronisation
 * does (SCORE_Kernel.cpp:4853). */
static __device__ __forceinline__ void infect_one(const GmDev &d, const GmPParams &pp, const GmSweepArgs &A,
                                                  uint32_t k, double pfin, double pP, const uint4 &pq0, uint32_t r,
                                                  uint32_t t, uint32_t g, uint32_t tx, uint32_t ty, uint32_t tz,
                                                  uint32_t gslot, bool fresh_weight = false) {
    const double overshoot = pfin - 1.0;
    const int cls = c_cls[r];
    const double carry = cls == 1 ? overshoot * 0.0682 : (cls == 2 ? overshoot * 0.0542 : overshoot * 0.0484);
    double Pn = pP;
    if (fresh_weight || pq0.z != g) Pn = pair_weight(d, pp, pq0.w, g);
    const uint32_t inf = c_code[r];
    double inc = 0.0;
    gm_cs_t cs = (gm_cs_t)GM_CS_DEFERRED;
    if (!is_pending(Pn)) cs = cs_code(d, predict_values(d, A, k + 1u, carry, Pn, g, inf, inc), k + 1u);
    GmRec *rec = d.rec + gslot;
    rec_set_a(rec, carry, k + 1u, inf);
    rec_set_b(rec, inc);
    rec_set_c(rec, Pn, t, g);
    *reinterpret_cast<uint4 *>(&rec->rxg) = make_uint4(pq0.w, 0u, 0u, 0u);
    d.cstep[gslot] = cs;
    /* DEFERRED cells go to the deferred list with one atomic per wave: the
     * list's counter is a single address, and atomics on one address are
     * served one after another */
    const bool deferred = cs == GM_CS_DEFERRED;
    const uint64_t dm = __ballot(deferred);
    if (dm) {
        const uint32_t lead = (uint32_t)__ffsll((long long)dm) - 1u;
        uint32_t base = 0;
        if ((threadIdx.x & 63u) == lead) base = atomicAdd(&d.st->ndeferred, (uint32_t)__popcll(dm));
        base = rdlane(base, lead);
        if (deferred) {
            const uint32_t i = base + lanemask_lt_count(dm);
            if (i < d.dlistCap) d.dlist[i] = gslot;
        }
    }
    clear_deformed(d, tx, ty, tz);
}


/* Retires a batch of parents (updateFullRX*, SCORE_Kernel.cpp:4640-4653): each
 * lane's parent, q0 = (slot, cell, defg, rxg), moves a cell from its deformed
 * grain's count to its nucleus' (its slot reads as INACTIVE from the next step
 * on by its completion code, cs_inactive).  A deformed grain losing its
 * last cell is reported so the host can update rhomax (update_myrhomax,
 * SCORE_RecoveryModels.cpp:146-165).  Consecutive parents often share their
 * nucleus and deformed grain (siblings), so each run of equal ids among
 * neighbouring lanes makes one atomic.  Every lane of the wave takes part. */
static __device__ __forceinline__ void retire_batch(const GmDev &d, const uint4 &q0, bool valid) {
    GmStats *st = d.st;
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t g = q0.z, r = q0.w;
    const uint32_t gprev = __shfl_up(g, 1, 64), rprev = __shfl_up(r, 1, 64);
    const uint64_t gstart = __ballot(valid && (lane == 0 || g != gprev));
    const uint64_t rstart = __ballot(valid && (lane == 0 || r != rprev));
    const uint64_t vmask = __ballot(valid);
    if (!valid) return;
    const uint64_t above = lane == 63 ? 0ull : (~0ull << (lane + 1)); /* lanes after this one */
    if ((gstart >> lane) & 1ull) {
        /* run length: up to the next run start or the last valid lane */
        const uint64_t nxt = (gstart | ~vmask) & above;
        const uint32_t end = nxt ? (uint32_t)__ffsll((long long)nxt) - 1u : 64u;
        const uint32_t n = end - lane;
        const uint32_t old = atomicSub(&d.defcnt[g], n);
        if (old == n) {
            const uint32_t i = atomicAdd(&st->nzeroed, 1u);
            if (i < GM_ZEROED_INLINE) st->zeroed_ids[i] = g;
            if (i < d.ndefg) d.zlist[i] = g;
        } else if (old < n) {
            /* only a cell held by two records retires twice; the count wraps
             * as SCORE's uint32 does and the host re-reads them all */
            st->cwrap = 1u;
        }
    }
    if ((rstart >> lane) & 1ull) {
        const uint64_t nxt = (rstart | ~vmask) & above;
        const uint32_t end = nxt ? (uint32_t)__ffsll((long long)nxt) - 1u : 64u;
        atomicAdd(&d.rxcnt[r], end - lane);
    }
}

/* The row of the step just done, and whether the next one may run without the
 * host: not after an error, a count that wrapped, the exhaustion of the grain
 * that sets rhomax (any_top), Xcells reaching a threshold the host acts on, a
 * weight the window cannot wait for, or lists outgrowing their allocations.
 * One thread. */
static __device__ void step_finish(const GmDev &d, const GmWindow &w, const GmStepRow &r, bool any_top) {
    GmStats *st = d.st;
    const uint32_t j = st->nexec;
    st->rows[j] = r;
    st->nexec = j + 1;
    st->kcur = st->kcur + 1u;
    st->nret = st->nret + (unsigned long long)(r.nfullI + r.nfullB);
    const uint32_t npend = __hip_atomic_load(&st->npend, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    st->npend0 = npend;
    const double xc = st->xcells + (double)(r.nfullI + r.nfullB);
    st->xcells = xc;
    /* the next step's arrivals may go to either list */
    const uint32_t kn = st->kcur;
    const uint64_t narr = (kn >= 1u && kn <= d.arr_last) ? (uint64_t)d.arr_off[kn + 1u] - d.arr_off[kn] : 0ull;
    const uint64_t slots = (uint64_t)r.nI + r.nB + narr;
    const uint64_t nIn = (uint64_t)r.nI + narr < 0xFFFFFFC0ull ? (uint64_t)r.nI + narr : 0xFFFFFFC0ull;
    /* with several regions each list is swept in whole 8-tile chunks */
    const uint64_t tiles = d.nreg > 1u ? slots / 64u + (uint64_t)GM_SW_TPC * (d.nlists + 1u)
                                       : (uint64_t)gm_tiles_inside((uint32_t)nIn) + ((uint64_t)r.nB + narr + 63) / 64 + GM_SW_TPC;
    bool halt = r.err != 0u || __hip_atomic_load(&st->cwrap, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
                any_top || xc >= w.xstop;
    halt = halt || (npend != 0u && !w.defer) || npend > w.pend_limit;
    halt = halt || __hip_atomic_load(&st->novf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
    halt = halt || slots > d.parentCap || tiles > d.tileCap;
    st->halt = halt ? 1u : 0u;
}

/* whether one of the grains exhausted in this step (zlist[z0, z1)) sets rhomax;
 * every lane of the calling wave takes part */
static __device__ __forceinline__ bool top_exhausted(const GmDev &d, const GmWindow &w, uint32_t z0, uint32_t z1) {
    bool top = false;
    for (uint32_t i = z0 + (threadIdx.x & 63u); i < z1; i += 64) top |= d.zlist[i] == w.top_grain;
    return __ballot(top) != 0ull;
}

/* The end of a step, run by one wave after the commit pass: the step's row and
 * whether the next step may run; the window's progress, (steps finished << 1) |
 * halt, goes to the host's mapped word, which paces the launches (gm_host.cpp,
 * step_window). */
static __device__ void step_tail(const GmDev &d, const GmWindow &w) {
    GmStats *st = d.st;
    const uint32_t j = st->nexec;
    const uint32_t z0 = j ? st->rows[j - 1].nzeroed : 0u, z1 = st->nzeroed < d.ndefg ? st->nzeroed : d.ndefg;
    const bool any_top = top_exhausted(d, w, z0, z1);
    if (threadIdx.x != 0) return;
    GmStepRow r;
    r.nI = st->nI;
    r.nB = st->nB;
    r.nfullI = st->nfullI;
    r.nfullB = st->nfullB;
    r.totRecI = st->totRecI;
    r.totRecB = st->totRecB;
    r.recLenI = st->recLenI;
    r.recLenB = st->recLenB;
    r.winI = st->winI;
    r.winB = st->winB;
    r.npend = st->npend - st->npend0;
    r.nattempt = st->nattempt;
    r.nzeroed = z1;
    r.err = st->err;
    r.nplaced = st->nplaced;
    r.pone = st->pone;
    r.pkey = st->arr_pkey;
    st->nplaced = 0;
    st->pone = 0;
    st->arr_pkey = 0ull;
    step_finish(d, w, r, any_top);
    if (d.prog) __hip_atomic_store(d.prog, (st->nexec << 1) | st->halt, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

/* orders a wave's LDS writes before its later LDS reads (the lanes of a wave
 * run in lockstep; this keeps the compiler from reordering across it) */
static __device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

/* Classifies a lane's eight slots of a sweep chunk (2-byte codes in cs, the
 * first nocc of them occupied): bit q of the low byte is set for a slot that
 * completes at step code kc, bit 8+q for an INACTIVE slot (cs_inactive: the
 * sentinel or a past step's code).  unknown counts slots without a valid
 * prediction: a cell is predicted before the sweep (when it is infected, or by
 * k_reanchor) or is DEFERRED, NEVER or FAR, so UNKNOWN is an internal error. */
static __device__ __forceinline__ uint32_t sweep_classify(const uint4 &cs, uint32_t nocc, uint32_t kc,
                                                          uint32_t horizon, uint32_t &unknown) {
    const uint32_t wd[4] = {cs.x, cs.y, cs.z, cs.w};
    uint32_t pack = 0;
#pragma unroll
    for (int q = 0; q < 8; q++) {
        const uint32_t x = (wd[q >> 1] >> (16 * (q & 1))) & 0xFFFFu;
        const uint32_t in = (uint32_t)q < nocc ? 1u : 0u;
        pack |= ((in & (x == kc ? 1u : 0u)) << q) | ((in & (cs_inactive(x, kc, horizon) ? 1u : 0u)) << (8 + q));
        unknown += in & (x == GM_CS_UNKNOWN ? 1u : 0u);
    }
    return pack;
}

/* The chunk's per-tile masks from every lane's pack: tile t holds slots
 * 64t .. 64t+63, the eight slots of each of lanes 8t .. 8t+7, so its completed
 * mask is those lanes' low bytes in lane order and its INACTIVE mask their high
 * bytes.  The bytes go through the wave's 128 bytes of LDS (sF, sR: eight
 * 64-bit words each); lane t < 8 returns tile t's masks. */
static __device__ __forceinline__ void sweep_masks(uint32_t pack, uint64_t *sF, uint64_t *sR, uint64_t &mF,
                                                   uint64_t &mR) {
    const uint32_t lane = threadIdx.x & 63u;
    reinterpret_cast<uint8_t *>(sF)[lane] = (uint8_t)pack;
    reinterpret_cast<uint8_t *>(sR)[lane] = (uint8_t)(pack >> 8);
    wave_sync();
    mF = lane < 8u ? sF[lane] : 0ull;
    mR = lane < 8u ? sR[lane] : 0ull;
    wave_sync();
}

/* the l with a[l] <= v < a[l+1] among a[0..n], a nondecreasing and
 * a[0] <= v < a[n]: the last of equal entries, so an empty list is skipped */
static __device__ __forceinline__ uint32_t upper_index(const uint32_t *a, uint32_t n, uint32_t v) {
    uint32_t lo = 0, hi = n;
    while (hi - lo > 1u) {
        const uint32_t m = (lo + hi) >> 1;
        if (a[m] <= v) lo = m;
        else hi = m;
    }
    return lo;
}

/* Occupied slot i of the front lists in list order: with one region Inside
 * slot i, then Border slot i - nI; with several, region after region, each
 * Inside then Border (d.L.occ holds the lists' first occupied index). */
static __device__ __forceinline__ uint32_t occ_slot(const GmDev &d, uint32_t nI, uint32_t i) {
    if (d.nreg <= 1u) return i < nI ? i : d.capI + (i - nI);
    const uint32_t l = upper_index(d.L.occ, d.nlists, i);
    return d.L.base[l] + (i - d.L.occ[l]);
}

#endif
