/*
 * gm_kernels.hip -- gfx950 kernels of the growth machine.
 *
 * One growth step of SCORE (SCORE_Kernel.cpp:6635-6654 with OMP_NUM_THREADS=1)
 * is, in serial order: sweep the Inside list, sweep the Border list, let every
 * completed Inside cell infect its 26 neighbours in list order, then every
 * completed Border cell.  A target is won by the first attempt in that order
 * that finds it deformed.  The same outcome is computed here by three kernels:
 *
 *   k_sweep    chunk-ordered pass over the occupied slots of both lists.  A slot
 *              costs one 4-byte read (its predicted completion step); cells
 *              infected in the previous step get their increment and prediction,
 *              and after a change of the per-step scalars every cell is
 *              re-anchored.  Completed and INACTIVE slots are ballot-compacted
 *              per 64-slot tile and placed by a decoupled look-back over chunks,
 *              giving the FullRX lists (ascending slot) and the capped recycling
 *              lists in one pass.
 *   k_propose  every (parent p, direction rank r) attempt on a deformed cell does
 *              atomicMin(claim[target], p*32 + r): the key orders attempts exactly
 *              as the serial loop issues them.
 *   k_infect   chunk-ordered pass over parents: an attempt wins iff its key is
 *              the minimum; winners per parent and destination list are counted,
 *              a look-back over chunks ranks every winner among the infections
 *              into its list (rank k takes the k-th recycled slot, else appends),
 *              the new records are written and the parents retire.
 *
 * FP64 throughout, compiled with -ffp-contract=off so every product and sum is
 * rounded exactly as in the reference's build without FMA.
 */
#include "gm_internal.h"
#include "gm_radd.h"

#define GM_RH 2
#define GM_RHO_RX (1.0e11)
#define GM_DOUBLE_ACCURACY (5.0e-16)

/* velocity shape factors per infector code (SCORE_Kernel.cpp:4051-4053) */
__constant__ double c_shape[27] = {0.6494, 0.7957, 0.6494, 0.7957, 1.0000, 0.7957, 0.6494, 0.7957, 0.6494,
                                   0.7957, 1.0000, 0.7957, 1.0000, 1.0000, 0.7957, 1.0000, 0.7957,
                                   0.6494, 0.7957, 0.6494, 0.7957, 1.0000, 0.7957, 0.6494, 0.7957, 0.6494, 1.0};

/* the 26 attempts by rank = call order of SCORE_Kernel.cpp:4611-4638:
 * offsets, infector code and overshoot class (1 face 0.0682, 2 edge 0.0542, 3 corner 0.0484) */
__constant__ int8_t c_dx[26] = {-1, 0, 1, -1, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1, -1, 0, 1};
__constant__ int8_t c_dy[26] = {-1, -1, -1, 0, 0, 1, 1, 1, 1, 1, 1, -1, -1, -1, 0, 0, 0, -1, -1, -1, 0, 0, 0, 1, 1, 1};
__constant__ int8_t c_dz[26] = {0, 0, 0, 0, 0, 0, 0, 0, -1, -1, -1, -1, -1, -1, -1, -1, -1, 1, 1, 1, 1, 1, 1, 1, 1, 1};
__constant__ uint8_t c_code[26] = {16, 15, 14, 13, 12, 11, 10, 9, 19, 18, 17, 25, 24, 23, 22, 21, 20,
                                   8, 7, 6, 5, 4, 3, 2, 1, 0};
__constant__ uint8_t c_cls[26] = {2, 1, 2, 1, 1, 2, 1, 2, 3, 2, 3, 3, 2, 3, 2, 1, 2, 3, 2, 3, 2, 1, 2, 3, 2, 3};

static __device__ __forceinline__ uint32_t lanemask_lt_count(uint64_t m) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
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
    for (uint32_t probe = 0; probe < d.hcap; probe++) {
        unsigned long long k = __atomic_load_n(&d.hkeys[h], __ATOMIC_RELAXED);
        if (k == 0ull) {
            unsigned long long old = atomicCAS(&d.hkeys[h], 0ull, key);
            if (old == 0ull) {
                uint32_t i = atomicAdd(&d.st->npend, 1u);
                d.pend_slot[i] = h;
                d.pend_key[i] = key;
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
    atomicOr(&d.st->err, GM_ERRF_PCACHE_FULL);
    return pending_box(0);
}

/* ---------------------------------------------------------------- helpers */

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}

static unsigned fill_blocks(size_t n) {
    size_t b = (n + GM_BLOCK - 1) / GM_BLOCK;
    if (b > 8192) b = 8192;
    if (b < 1) b = 1;
    return (unsigned)b;
}

hipError_t gm_launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_fill_u32<<<fill_blocks(n), GM_BLOCK, 0, s>>>(p, v, n);
    return hipGetLastError();
}
hipError_t gm_launch_fill_u64(unsigned long long *p, unsigned long long v, size_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_fill_u64<<<fill_blocks(n), GM_BLOCK, 0, s>>>(p, v, n);
    return hipGetLastError();
}
hipError_t gm_launch_fill_f64_bits(double *p, unsigned long long bits, size_t n, hipStream_t s) {
    return gm_launch_fill_u64((unsigned long long *)p, bits, n, s);
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

/* ------------------------------------------------------------------- sweep */

/* rxFrac increment of one cell: SCORE's v = m*p; v *= shape; rxFrac += (v*dt)*inv_dx
 * (SCORE_Kernel.cpp:4077-4092, :5459-5470); returns false when p <= DOUBLE_ACCURACY
 * (the cell is not advanced) */
static __device__ __forceinline__ bool cell_increment(const GmDev &d, const GmSweepParams &sp, uint32_t s,
                                                      double &inc) {
    double P = d.P[s];
    if (is_pending(P)) {
        P = d.hvals[(uint32_t)__double_as_longlong(P)];
        d.P[s] = P;
    }
    double m;
    if (sp.model == GM_RH) m = P * sp.mRH;
    else if (P < 0.0) m = sp.mLAGB;
    else m = (P * sp.mGS) + ((1.0 - P) * sp.mHAGB);
    double p = (sp.Gbhalfsq * (d.rho[d.defg[s]] - GM_RHO_RX)) - sp.zener;
    if (!(p > GM_DOUBLE_ACCURACY)) return false;
    double v = m * p;
    v *= c_shape[d.inf[s]];
    inc = (v * sp.dt) * sp.inv_dx;
    return true;
}

/* Completion step of slot s from its value f before step k: the first step
 * >= k in which repeated rounded additions of the increment reach 1.0.  Stores
 * the increment (0.0 for a cell SCORE does not advance), the completion step
 * and the value after the completing addition, whose excess over 1.0 the
 * infection carries on. */
static __device__ __forceinline__ uint32_t predict_slot(const GmDev &d, const GmSweepArgs &A, uint32_t s, double f) {
    double inc;
    uint32_t cs = GM_CS_NEVER;
    if (cell_increment(d, A.sp, s, inc)) {
        if (A.k <= A.stepcap) {
            const uint64_t j = gm_radd_until_one(f, inc, (uint64_t)(A.stepcap - A.k) + 1u);
            if (j) {
                cs = A.k + (uint32_t)(j - 1u);
                d.fin[s] = f;
            }
        }
    } else {
        inc = 0.0;
    }
    d.inc[s] = inc;
    d.cstep[s] = cs;
    return cs;
}

/* calcGrowthStepInside/Border (SCORE_Kernel.cpp:4046-4249) for one slot whose
 * completion step is cs: returns 1 when the cell completes in step k, 2 when the
 * slot is INACTIVE, 0 otherwise */
static __device__ __forceinline__ int sweep_slot(const GmDev &d, const GmSweepArgs &A, uint32_t s, uint32_t cs) {
    if (cs == GM_CS_INACTIVE) return 2;
    const uint32_t k = A.k;
    if (A.reanchor) {
        double f = d.frac0[s];
        if (cs != GM_CS_UNKNOWN) f = gm_radd_n(f, d.inc[s], k - d.s0[s]);
        d.frac0[s] = f;
        d.s0[s] = k;
        cs = predict_slot(d, A, s, f);
    } else if (cs == GM_CS_UNKNOWN) {
        cs = predict_slot(d, A, s, d.frac0[s]);
    } else if (cs < k) {
        atomicOr(&d.st->err, GM_ERRF_MISSED);
    }
    return cs == k ? 1 : 0;
}

/* cells infected in the previous step, listed by the infection pass, predicted
 * with every lane of a wave busy */
__global__ void __launch_bounds__(GM_BLOCK) k_predict_new(GmDev d, GmSweepArgs A, uint32_t n) {
    for (uint32_t i = blockIdx.x * GM_BLOCK + threadIdx.x; i < n; i += gridDim.x * GM_BLOCK) {
        const uint32_t s = d.newl[i];
        if (d.cstep[s] == GM_CS_UNKNOWN) predict_slot(d, A, s, d.frac0[s]);
    }
}

/* per-block list counts: [0] full Inside, [1] recycled Inside, [2] full Border,
 * [3] recycled Border, [4] increments computed; each GM_MAX_BLOCKS long */
#define BC_FI 0
#define BC_RI 1
#define BC_FB 2
#define BC_RB 3
#define BC_UN 4

/* static partition of `n` items over the grid in multiples of `q` */
static __device__ __forceinline__ void part_of(uint32_t n, uint32_t q, uint32_t &lo, uint32_t &hi) {
    uint32_t per = (n + gridDim.x - 1) / gridDim.x;
    per = (per + q - 1) / q * q;
    lo = blockIdx.x * per;
    hi = lo + per;
    if (lo > n) lo = n;
    if (hi > n) hi = n;
}

static __device__ __forceinline__ uint32_t tile_slot0(const GmDev &d, uint32_t tile, uint32_t tI) {
    return tile < tI ? tile * 64u : d.capI + (tile - tI) * 64u;
}

/* pass 1: advance/predict every occupied slot, ballot the completed and the
 * INACTIVE ones into per-tile masks, count them per block and list.  Each wave
 * loads the completion steps of GM_SW_UNROLL tiles before working on them. */
#define GM_SW_UNROLL 4
__global__ void __launch_bounds__(GM_BLOCK) k_sweep(GmDev d, GmSweepArgs A, uint32_t tI, uint32_t T) {
    __shared__ uint32_t red[GM_BLOCK / 64][5];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t lo, hi;
    part_of(T, GM_BLOCK / 64, lo, hi);
    uint32_t cFI = 0, cRI = 0, cFB = 0, cRB = 0, unknown = 0;
    for (uint32_t t0 = lo + w; t0 < hi; t0 += (GM_BLOCK / 64) * GM_SW_UNROLL) {
        uint32_t cs[GM_SW_UNROLL];
        bool in[GM_SW_UNROLL];
#pragma unroll
        for (int u = 0; u < GM_SW_UNROLL; u++) {
            const uint32_t tile = t0 + u * (GM_BLOCK / 64);
            const uint32_t s = tile_slot0(d, tile, tI) + lane;
            in[u] = tile < hi && s < (tile < tI ? A.nI : d.capI + A.nB);
            cs[u] = in[u] ? d.cstep[s] : GM_CS_INACTIVE;
        }
#pragma unroll
        for (int u = 0; u < GM_SW_UNROLL; u++) {
            const uint32_t tile = t0 + u * (GM_BLOCK / 64);
            if (tile >= hi) break;
            const uint32_t s = tile_slot0(d, tile, tI) + lane;
            int r = 0;
            if (in[u]) {
                if (!A.reanchor && cs[u] == GM_CS_UNKNOWN) unknown++;
                r = sweep_slot(d, A, s, cs[u]);
            }
            const uint64_t bf = __ballot(r == 1), br = __ballot(r == 2);
            if (lane == 0) {
                d.mF[tile] = bf;
                d.mR[tile] = br;
                if (tile < tI) { cFI += (uint32_t)__popcll(bf); cRI += (uint32_t)__popcll(br); }
                else { cFB += (uint32_t)__popcll(bf); cRB += (uint32_t)__popcll(br); }
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) unknown += __shfl_down(unknown, o, 64);
    if (lane == 0) {
        red[w][0] = cFI; red[w][1] = cRI; red[w][2] = cFB; red[w][3] = cRB; red[w][4] = unknown;
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        uint32_t t = 0;
        for (int i = 0; i < GM_BLOCK / 64; i++) t += red[i][threadIdx.x];
        d.bcnt[threadIdx.x * GM_MAX_BLOCKS + blockIdx.x] = t;
    }
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

/* pass 2: per-block offsets and the list lengths; also resets the counters the
 * infection phase accumulates */
__global__ void __launch_bounds__(1024) k_sweep_scan(GmDev d, uint32_t nb) {
    uint32_t tot[5];
    scan_block_counts<5>(d.bcnt, nb, tot);
    if (threadIdx.x == 0) {
        GmStats *st = d.st;
        st->nunknown += tot[BC_UN];
        st->nfullI = tot[BC_FI];
        st->totRecI = tot[BC_RI];
        st->nfullB = tot[BC_FB];
        st->totRecB = tot[BC_RB];
        st->recLenI = tot[BC_RI] < d.recCapI ? tot[BC_RI] : d.recCapI;
        st->recLenB = tot[BC_RB] < d.recCapB ? tot[BC_RB] : d.recCapB;
        st->npend = 0;
        st->pone = 0;
        st->nzeroed = 0;
    }
}

/* block-wide exclusive scan of one value per thread (GM_BLOCK threads) */
static __device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *total, uint32_t (*ws)[GM_BLOCK / 64],
                                                           int c) {
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t incl = wave_incl_scan(v);
    if (lane == 63) ws[c][w] = incl;
    __syncthreads();
    uint32_t pre = 0, t = 0;
#pragma unroll
    for (int i = 0; i < GM_BLOCK / 64; i++) {
        if (i < (int)w) pre += ws[c][i];
        t += ws[c][i];
    }
    *total = t;
    return pre + incl - v;
}

/* pass 3: append_to_fullrx_*_list / append_to_recycling_*_list
 * (SCORE_Kernel.cpp:3994-4042, :4133-4180) as ordered compaction: FullRX lists
 * hold global slots in ascending order, recycling lists are capped */
__global__ void __launch_bounds__(GM_BLOCK) k_compact(GmDev d, uint32_t tI, uint32_t T) {
    __shared__ uint64_t sF[GM_BLOCK], sR[GM_BLOCK];
    __shared__ uint32_t off[4][GM_BLOCK];
    __shared__ uint32_t ws[4][GM_BLOCK / 64];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t lo, hi;
    part_of(T, GM_BLOCK / 64, lo, hi);
    uint32_t run[4];
    for (int c = 0; c < 4; c++) run[c] = d.bcnt[c * GM_MAX_BLOCKS + blockIdx.x];
    for (uint32_t batch = lo; batch < hi; batch += GM_BLOCK) {
        const uint32_t t = batch + threadIdx.x;
        const bool valid = t < hi;
        const uint64_t bf = valid ? d.mF[t] : 0ull, br = valid ? d.mR[t] : 0ull;
        const bool isB = t >= tI;
        const uint32_t cf = (uint32_t)__popcll(bf), cr = (uint32_t)__popcll(br);
        const uint32_t v[4] = {isB ? 0u : cf, isB ? 0u : cr, isB ? cf : 0u, isB ? cr : 0u};
        uint32_t tot[4];
#pragma unroll
        for (int c = 0; c < 4; c++) off[c][threadIdx.x] = run[c] + block_excl_scan(v[c], &tot[c], ws, c);
        sF[threadIdx.x] = bf;
        sR[threadIdx.x] = br;
        __syncthreads();
        const uint32_t nt = hi - batch < GM_BLOCK ? hi - batch : GM_BLOCK;
        for (uint32_t j = w; j < nt; j += GM_BLOCK / 64) {
            const uint32_t tile = batch + j;
            const uint32_t s = tile_slot0(d, tile, tI) + lane;
            const uint64_t mf = sF[j], mr = sR[j];
            const bool tb = tile >= tI;
            if ((mf >> lane) & 1ull) {
                const uint32_t pos = lanemask_lt_count(mf);
                if (tb) d.fullB[off[BC_FB][j] + pos] = s;
                else d.fullI[off[BC_FI][j] + pos] = s;
            }
            if ((mr >> lane) & 1ull) {
                const uint32_t pos = lanemask_lt_count(mr);
                if (tb) {
                    const uint32_t r = off[BC_RB][j] + pos;
                    if (r < d.recCapB) d.recB[r] = s - d.capI;
                } else {
                    const uint32_t r = off[BC_RI][j] + pos;
                    if (r < d.recCapI) d.recI[r] = s;
                }
            }
        }
        for (int c = 0; c < 4; c++) run[c] += tot[c];
        __syncthreads();
    }
}

static uint32_t grid_for(uint32_t items, uint32_t per_block) {
    uint32_t b = (items + per_block - 1) / per_block;
    if (b > GM_MAX_BLOCKS) b = GM_MAX_BLOCKS;
    if (b < 1) b = 1;
    return b;
}

hipError_t gm_launch_sweep(const GmDev &d, const GmSweepArgs &a, hipStream_t s) {
    if (!a.reanchor && a.nnew) {
        const uint32_t n = a.nnew < d.newCap ? a.nnew : d.newCap;
        k_predict_new<<<grid_for(n, GM_BLOCK), GM_BLOCK, 0, s>>>(d, a, n);
    }
    const uint32_t tI = (a.nI + 63) / 64, T = tI + (a.nB + 63) / 64;
    const uint32_t nb = grid_for(T, 4 * (GM_BLOCK / 64));
    k_sweep<<<nb, GM_BLOCK, 0, s>>>(d, a, tI, T);
    k_sweep_scan<<<1, 1024, 0, s>>>(d, nb);
    k_compact<<<nb, GM_BLOCK, 0, s>>>(d, tI, T);
    return hipGetLastError();
}

/* ----------------------------------------------------------------- infection */

static __device__ __forceinline__ uint32_t parent_slot(const GmDev &d, uint32_t p, uint32_t nfullI) {
    return p < nfullI ? d.fullI[p] : d.fullB[p - nfullI];
}

/* Target of attempt rank r from the parent in global slot s at cell `pos`.
 * Inside parents (omp_infect_OutOfRXFrontInside, SCORE_Kernel.cpp:4252-4344) never
 * wrap; they send every target to the Border list unless they lie more than one
 * cell away from every face (:4606-4609).  Border parents (:4347-4493) wrap x and
 * y, drop targets outside z, and send strictly interior targets to the Inside list. */
static __device__ __forceinline__ bool attempt_target(const GmDev &d, bool fromB, uint32_t pos, int r, uint32_t &t,
                                                      bool &toB) {
    uint32_t z = pos / d.nxy;
    uint32_t rem = pos - z * d.nxy;
    uint32_t y = rem / d.nx;
    uint32_t x = rem - y * d.nx;
    int tx = (int)x + c_dx[r], ty = (int)y + c_dy[r], tz = (int)z + c_dz[r];
    const int nx = (int)d.nx, ny = (int)d.ny, nz = (int)d.nz;
    if (fromB) {
        if (tx < 0) tx += nx;
        if (tx >= nx) tx -= nx;
        if (ty < 0) ty += ny;
        if (ty >= ny) ty -= ny;
        if (tz < 0 || tz >= nz) return false;
        toB = !(tx > 0 && tx < nx - 1 && ty > 0 && ty < ny - 1 && tz > 0 && tz < nz - 1);
    } else {
        /* Inside cells are strictly interior, so this never fires; it keeps a
         * corrupted record from addressing outside the grid */
        if (tx < 0 || tx >= nx || ty < 0 || ty >= ny || tz < 0 || tz >= nz) return false;
        toB = !((int)x > 1 && (int)x < nx - 2 && (int)y > 1 && (int)y < ny - 2 && (int)z > 1 && (int)z < nz - 2);
    }
    t = (uint32_t)tx + (uint32_t)ty * d.nx + (uint32_t)tz * d.nxy;
    return true;
}

/* all attempts claim their target: the smallest key p*32+r is the serial winner */
__global__ void __launch_bounds__(GM_BLOCK) k_propose(GmDev d) {
    const uint32_t nfullI = d.st->nfullI, nfull = nfullI + d.st->nfullB;
    const uint32_t lane = threadIdx.x & 31;
    const uint32_t ngroups = gridDim.x * (GM_BLOCK / 32);
    for (uint32_t p = blockIdx.x * (GM_BLOCK / 32) + (threadIdx.x >> 5); p < nfull; p += ngroups) {
        if (lane >= 26) continue;
        const uint32_t s = parent_slot(d, p, nfullI);
        uint32_t t;
        bool toB;
        if (!attempt_target(d, s >= d.capI, d.cell[s], (int)lane, t, toB)) continue;
        if (d.grid[t] < d.ndefg) atomicMin(&d.claim[t], p * 32u + lane);
    }
}

/* winners per parent and destination list, as 26-bit masks over the ranks;
 * per-block totals for the ranking scan */
__global__ void __launch_bounds__(GM_BLOCK) k_decide(GmDev d) {
    __shared__ uint32_t red[GM_BLOCK / 32][2];
    const uint32_t nfullI = d.st->nfullI, nfull = nfullI + d.st->nfullB;
    const uint32_t lane = threadIdx.x & 31, grp = threadIdx.x >> 5, half = threadIdx.x & 32;
    uint32_t lo, hi;
    part_of(nfull, GM_BLOCK / 32, lo, hi);
    uint32_t cI = 0, cB = 0;
    for (uint32_t p = lo + grp; p < hi; p += GM_BLOCK / 32) {
        bool wI = false, wB = false;
        if (lane < 26) {
            const uint32_t s = parent_slot(d, p, nfullI);
            uint32_t t;
            bool toB;
            if (attempt_target(d, s >= d.capI, d.cell[s], (int)lane, t, toB) && d.grid[t] < d.ndefg &&
                d.claim[t] == p * 32u + lane) {
                wB = toB;
                wI = !toB;
            }
        }
        const uint32_t bI = (uint32_t)(__ballot(wI) >> half), bB = (uint32_t)(__ballot(wB) >> half);
        if (lane == 0) {
            d.pmI[p] = bI;
            d.pmB[p] = bB;
            cI += (uint32_t)__popc(bI);
            cB += (uint32_t)__popc(bB);
        }
    }
    if (lane == 0) { red[grp][0] = cI; red[grp][1] = cB; }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint32_t t = 0;
        for (int i = 0; i < GM_BLOCK / 32; i++) t += red[i][threadIdx.x];
        d.bcnt[threadIdx.x * GM_MAX_BLOCKS + blockIdx.x] = t;
    }
}

/* one block: offsets of every block's winners in each list, the new list lengths
 * (recycled slots first, then appended) */
__global__ void __launch_bounds__(1024) k_decide_scan(GmDev d, uint32_t nb) {
    uint32_t tot[2];
    scan_block_counts<2>(d.bcnt, nb, tot);
    if (threadIdx.x == 0) {
        GmStats *st = d.st;
        st->winI = tot[0];
        st->winB = tot[1];
        /* the commit pass reads the lengths before this step's appends */
        st->nIold = st->nI;
        st->nBold = st->nB;
        if (tot[0] > st->recLenI) st->nI += tot[0] - st->recLenI;
        if (tot[1] > st->recLenB) st->nB += tot[1] - st->recLenB;
    }
}

/* new records as omp_infect_OutOfRXFront* write them (SCORE_Kernel.cpp:4281-4310,
 * :4402-4426), slots as omp_getNextFreeSlotIn* hand them out (:3757-3874); then
 * the block's parents retire as updateFullRX* does it (:4640-4653) */
__global__ void __launch_bounds__(GM_BLOCK) k_commit(GmDev d, GmPParams pp, uint32_t k) {
    __shared__ uint32_t sI[GM_BLOCK], sB[GM_BLOCK], oI[GM_BLOCK], oB[GM_BLOCK];
    __shared__ uint32_t ws[2][GM_BLOCK / 64];
    GmStats *st = d.st;
    const uint32_t nfullI = st->nfullI, nfull = nfullI + st->nfullB;
    const uint32_t recLenI = st->recLenI, recLenB = st->recLenB;
    const uint32_t nI = st->nIold, nB = st->nBold, winI = st->winI;
    const uint32_t lane = threadIdx.x & 31, grp = threadIdx.x >> 5;
    uint32_t lo, hi;
    part_of(nfull, GM_BLOCK / 32, lo, hi);
    uint32_t runI = d.bcnt[0 * GM_MAX_BLOCKS + blockIdx.x], runB = d.bcnt[1 * GM_MAX_BLOCKS + blockIdx.x];
    for (uint32_t batch = lo; batch < hi; batch += GM_BLOCK) {
        const uint32_t pt = batch + threadIdx.x;
        const uint32_t mI = pt < hi ? d.pmI[pt] : 0u, mB = pt < hi ? d.pmB[pt] : 0u;
        uint32_t totI, totB;
        oI[threadIdx.x] = runI + block_excl_scan((uint32_t)__popc(mI), &totI, ws, 0);
        oB[threadIdx.x] = runB + block_excl_scan((uint32_t)__popc(mB), &totB, ws, 1);
        sI[threadIdx.x] = mI;
        sB[threadIdx.x] = mB;
        __syncthreads();
        const uint32_t np = hi - batch < GM_BLOCK ? hi - batch : GM_BLOCK;
        const uint32_t bit = 1u << lane;
        for (uint32_t j = grp; j < np; j += GM_BLOCK / 32) {
            const uint32_t wI = sI[j], wB = sB[j];
            if (lane >= 26 || !((wI | wB) & bit)) continue;
            const uint32_t p = batch + j;
            const uint32_t s = parent_slot(d, p, nfullI);
            uint32_t t;
            bool toB;
            attempt_target(d, s >= d.capI, d.cell[s], (int)lane, t, toB);
            uint32_t gslot = GM_CS_INACTIVE, nidx;
            if (wI & bit) {
                const uint32_t rank = oI[j] + (uint32_t)__popc(wI & (bit - 1u));
                const uint32_t rel = rank < recLenI ? d.recI[rank] : nI + (rank - recLenI);
                if (rel >= d.capI) atomicOr(&st->err, GM_ERRF_INSIDE_FULL);
                else gslot = rel;
                nidx = rank;
            } else {
                const uint32_t rank = oB[j] + (uint32_t)__popc(wB & (bit - 1u));
                const uint32_t rel = rank < recLenB ? d.recB[rank] : nB + (rank - recLenB);
                if (rel >= d.capB) atomicOr(&st->err, GM_ERRF_BORDER_FULL);
                else gslot = d.capI + rel;
                nidx = winI + rank;
            }
            if (gslot == GM_CS_INACTIVE) continue;
            if (nidx < d.newCap) d.newl[nidx] = gslot;
            const double overshoot = d.fin[s] - 1.0;
            const int cls = c_cls[lane];
            const double carry = cls == 1 ? overshoot * 0.0682 : (cls == 2 ? overshoot * 0.0542 : overshoot * 0.0484);
            const uint32_t g = d.grid[t];
            const uint32_t prxg = d.rxg[s];
            double Pn = d.P[s];
            if (d.defg[s] != g) Pn = pair_weight(d, pp, prxg, g);
            d.cstep[gslot] = GM_CS_UNKNOWN;
            d.frac0[gslot] = carry;
            d.s0[gslot] = k + 1u;
            d.inf[gslot] = c_code[lane];
            d.cell[gslot] = t;
            d.P[gslot] = Pn;
            d.defg[gslot] = g;
            d.rxg[gslot] = prxg;
            d.grid[t] = GM_INFECTED;
            d.claim[t] = GM_FREE_CLAIM;
        }
        runI += totI;
        runB += totB;
        __syncthreads();
    }
    /* retire this block's parents; a deformed grain losing its last cell is
     * reported so the host can update rhomax (update_myrhomax, SCORE_RecoveryModels.cpp:146-165) */
    for (uint32_t p = lo + threadIdx.x; p < hi; p += GM_BLOCK) {
        const uint32_t s = parent_slot(d, p, nfullI);
        const uint32_t r = d.rxg[s], g = d.defg[s];
        d.grid[d.cell[s]] = d.ndefg + r;
        atomicAdd(&d.rxcnt[r], 1u);
        if (atomicSub(&d.defcnt[g], 1u) == 1u) {
            const uint32_t i = atomicAdd(&st->nzeroed, 1u);
            if (i < GM_ZEROED_INLINE) st->zeroed_ids[i] = g;
            d.zlist[i] = g;
        }
        d.cstep[s] = GM_CS_INACTIVE;
    }
}

static unsigned parent_blocks(uint32_t nupper) {
    uint32_t b = (nupper + (GM_BLOCK / 32) - 1) / (GM_BLOCK / 32);
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    return b;
}

hipError_t gm_launch_infect(const GmDev &d, const GmPParams &pp, uint32_t k, uint32_t parents_upper, hipStream_t s) {
    k_propose<<<parent_blocks(parents_upper), GM_BLOCK, 0, s>>>(d);
    const uint32_t nb = grid_for(parents_upper, 8 * (GM_BLOCK / 32));
    k_decide<<<nb, GM_BLOCK, 0, s>>>(d);
    k_decide_scan<<<1, 1024, 0, s>>>(d, nb);
    k_commit<<<nb, GM_BLOCK, 0, s>>>(d, pp, k);
    return hipGetLastError();
}

/* ------------------------------------------------- device-wide exclusive scan
 * (nucleation only) Two u32 arrays scanned in place over n entries.  Fixed grid:
 * block b owns a contiguous chunk; phase 2 scans the block sums. */

static __device__ __forceinline__ uint32_t block_incl_scan(uint32_t v, uint32_t *total) {
    __shared__ uint32_t ws[GM_BLOCK / 64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = wave_incl_scan(v);
    if (lane == 63) ws[w] = inc;
    __syncthreads();
    uint32_t pre = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < GM_BLOCK / 64; i++) {
        uint32_t x = ws[i];
        if (i < w) pre += x;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return inc + pre;
}

static __device__ __forceinline__ void chunk_of(uint32_t n, uint32_t &lo, uint32_t &hi) {
    uint32_t per = (n + gridDim.x - 1) / gridDim.x;
    per = (per + GM_BLOCK - 1) / GM_BLOCK * GM_BLOCK;
    lo = blockIdx.x * per;
    hi = lo + per;
    if (lo > n) lo = n;
    if (hi > n) hi = n;
}

__global__ void k_scan_p1(const uint32_t *a, const uint32_t *b, uint32_t n, uint32_t *bsum) {
    uint32_t lo, hi;
    chunk_of(n, lo, hi);
    uint32_t sa = 0, sb = 0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += GM_BLOCK) {
        sa += a[i];
        sb += b[i];
    }
    uint32_t ta, tb;
    block_incl_scan(sa, &ta);
    block_incl_scan(sb, &tb);
    if (threadIdx.x == 0) {
        bsum[blockIdx.x] = ta;
        bsum[gridDim.x + blockIdx.x] = tb;
    }
}

__global__ void k_scan_p2(uint32_t *bsum, uint32_t nb, uint32_t *totals) {
    __shared__ uint32_t wsa[16], wsb[16];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t va = threadIdx.x < nb ? bsum[threadIdx.x] : 0u;
    uint32_t vb = threadIdx.x < nb ? bsum[nb + threadIdx.x] : 0u;
    uint32_t ia = wave_incl_scan(va), ib = wave_incl_scan(vb);
    if (lane == 63) { wsa[w] = ia; wsb[w] = ib; }
    __syncthreads();
    uint32_t pa = 0, pb = 0, ta = 0, tb = 0;
    for (int i = 0; i < 16; i++) {
        if (i < w) { pa += wsa[i]; pb += wsb[i]; }
        ta += wsa[i];
        tb += wsb[i];
    }
    if (threadIdx.x < nb) {
        bsum[threadIdx.x] = ia - va + pa;
        bsum[nb + threadIdx.x] = ib - vb + pb;
    }
    if (threadIdx.x == 0) { totals[0] = ta; totals[1] = tb; }
}

__global__ void k_scan_p3(uint32_t *a, uint32_t *b, uint32_t n, const uint32_t *bsum) {
    uint32_t lo, hi;
    chunk_of(n, lo, hi);
    uint32_t ra = bsum[blockIdx.x], rb = bsum[gridDim.x + blockIdx.x];
    for (uint32_t base = lo; base < hi; base += GM_BLOCK) {
        uint32_t i = base + threadIdx.x;
        uint32_t va = i < hi ? a[i] : 0u, vb = i < hi ? b[i] : 0u;
        uint32_t ta, tb;
        uint32_t ia = block_incl_scan(va, &ta);
        uint32_t ib = block_incl_scan(vb, &tb);
        if (i < hi) {
            a[i] = ra + ia - va;
            b[i] = rb + ib - vb;
        }
        ra += ta;
        rb += tb;
    }
}

static hipError_t scan2(const GmDev &d, uint32_t *a, uint32_t *b, uint32_t n, hipStream_t s) {
    uint32_t nb = (n + GM_BLOCK - 1) / GM_BLOCK;
    if (nb > GM_SCAN_BLOCKS) nb = GM_SCAN_BLOCKS;
    if (nb < 1) nb = 1;
    k_scan_p1<<<nb, GM_BLOCK, 0, s>>>(a, b, n, d.scan_bsum);
    k_scan_p2<<<1, 1024, 0, s>>>(d.scan_bsum, nb, d.st->scan_tot);
    k_scan_p3<<<nb, GM_BLOCK, 0, s>>>(a, b, n, d.scan_bsum);
    return hipGetLastError();
}

/* ------------------------------------------------------------ nucleation
 * ompshar_sim_myCA_sitesaturatedNucleation + infect_JuvenileNucleation
 * (SCORE_GrowthMachine.cpp:50-189): nuclei in pool order, a site already taken
 * rejects later nuclei; interior sites append to Inside, shell sites to Border. */

__global__ void k_nuc_claim(GmDev d, const uint32_t *site, uint32_t n) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        uint32_t w = site[k];
        if (d.grid[w] < d.ndefg) atomicMin(&d.claim[w], k);
    }
}

static __device__ __forceinline__ bool site_interior(const GmDev &d, uint32_t w) {
    uint32_t z = w / d.nxy, rem = w - z * d.nxy, y = rem / d.nx, x = rem - y * d.nx;
    return x > 0 && x < d.nx - 1 && y > 0 && y < d.ny - 1 && z > 0 && z < d.nz - 1;
}

__global__ void k_nuc_flag(GmDev d, const uint32_t *site, uint32_t n) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        uint32_t w = site[k];
        bool win = d.grid[w] < d.ndefg && d.claim[w] == k;
        bool in = site_interior(d, w);
        d.nI_[k] = (win && in) ? 1u : 0u;
        d.nB_[k] = (win && !in) ? 1u : 0u;
        d.nflag[k] = win ? 1u : 0u;
    }
}

__global__ void k_nuc_commit(GmDev d, const uint32_t *site, const double *nucP, uint32_t n) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        if (!d.nflag[k]) continue;
        uint32_t w = site[k];
        bool in = site_interior(d, w);
        uint32_t gslot;
        if (in) {
            gslot = d.nI_[k];
            if (gslot >= d.capI) { atomicOr(&d.st->err, GM_ERRF_INSIDE_FULL); continue; }
        } else {
            uint32_t rel = d.nB_[k];
            if (rel >= d.capB) { atomicOr(&d.st->err, GM_ERRF_BORDER_FULL); continue; }
            gslot = d.capI + rel;
        }
        uint32_t g = d.grid[w];
        d.cstep[gslot] = GM_CS_UNKNOWN;
        d.frac0[gslot] = 1.0;
        d.s0[gslot] = 0u;
        d.inf[gslot] = 26;
        d.cell[gslot] = w;
        d.P[gslot] = nucP[k];
        d.defg[gslot] = g;
        d.rxg[gslot] = k;
        d.grid[w] = GM_INFECTED;
        d.claim[w] = GM_FREE_CLAIM;
    }
}

__global__ void k_nuc_tail(GmDev d) {
    d.st->nI = d.st->scan_tot[0];
    d.st->nB = d.st->scan_tot[1];
    if (d.st->nB > d.capB) atomicOr(&d.st->err, GM_ERRF_BORDER_FULL);
}

hipError_t gm_launch_nucleate(const GmDev &d, const uint32_t *site, const double *nucP, uint32_t n, hipStream_t s) {
    if (n == 0) {
        k_fill_u32<<<1, 32, 0, s>>>(d.st->scan_tot, 0u, 2);
        k_nuc_tail<<<1, 1, 0, s>>>(d);
        return hipGetLastError();
    }
    unsigned nb = fill_blocks(n);
    k_nuc_claim<<<nb, GM_BLOCK, 0, s>>>(d, site, n);
    k_nuc_flag<<<nb, GM_BLOCK, 0, s>>>(d, site, n);
    hipError_t e = scan2(d, d.nI_, d.nB_, n, s);
    if (e != hipSuccess) return e;
    k_nuc_commit<<<nb, GM_BLOCK, 0, s>>>(d, site, nucP, n);
    k_nuc_tail<<<1, 1, 0, s>>>(d);
    return hipGetLastError();
}

/* ------------------------------------------------------ mobility-weight cache */

__global__ void k_resolve_pending(GmDev d, const double *vals, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        d.hvals[d.pend_slot[i]] = vals[i];
}

hipError_t gm_launch_resolve_pending(const GmDev &d, const double *vals, uint32_t n, hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_resolve_pending<<<fill_blocks(n), GM_BLOCK, 0, s>>>(d, vals, n);
    return hipGetLastError();
}

/* SCORE's per-slot view before step k: activity, rxFrac (replayed from the
 * anchor; INACTIVE slots hold 0.0 as the reference leaves them), and P with
 * pending markers replaced by their cached values */
__global__ void k_materialize(GmDev d, uint32_t k, uint32_t base, uint32_t n, uint8_t *act, double *frac) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t s = base + i;
        const uint32_t cs = d.cstep[s];
        double P = d.P[s];
        if (is_pending(P)) {
            P = d.hvals[(uint32_t)__double_as_longlong(P)];
            d.P[s] = P;
        }
        if (!act) continue;
        if (cs == GM_CS_INACTIVE) {
            act[i] = 0;
            frac[i] = 0.0;
            continue;
        }
        act[i] = 1;
        double f = d.frac0[s];
        if (cs != GM_CS_UNKNOWN) f = gm_radd_n(f, d.inc[s], k - d.s0[s]);
        frac[i] = f;
    }
}

hipError_t gm_launch_materialize(const GmDev &d, uint32_t k, uint32_t base, uint32_t n, uint8_t *act, double *frac,
                                 hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_materialize<<<fill_blocks(n), GM_BLOCK, 0, s>>>(d, k, base, n, act, frac);
    return hipGetLastError();
}

__global__ void k_hash_insert(GmDev d, const unsigned long long *keys, const double *vals, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        unsigned long long key = keys[i];
        uint32_t h = hash_slot(key, d.hcap);
        for (uint32_t probe = 0; probe < d.hcap; probe++) {
            unsigned long long old = atomicCAS(&d.hkeys[h], 0ull, key);
            if (old == 0ull || old == key) {
                d.hvals[h] = vals[i];
                break;
            }
            h = (h + 1u) & (d.hcap - 1u);
        }
    }
}

hipError_t gm_launch_hash_insert(const GmDev &d, const unsigned long long *keys, const double *vals, uint32_t n,
                                 hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_hash_insert<<<fill_blocks(n), GM_BLOCK, 0, s>>>(d, keys, vals, n);
    return hipGetLastError();
}
