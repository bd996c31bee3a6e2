/*
 * gm_kernels.hip -- gfx950 kernels of the growth machine.
 *
 * One growth step of SCORE (SCORE_Kernel.cpp:6635-6654 with OMP_NUM_THREADS=1)
 * is, in serial order: sweep the Inside list, sweep the Border list, let every
 * completed Inside cell infect its 26 neighbours in list order, then every
 * completed Border cell.  A target is won by the first attempt in that order
 * that finds it deformed.  Here the same outcome is computed in parallel:
 *
 *   sweep    one wave per 64 slots: advance rxFrac, ballot the completed and the
 *            INACTIVE slots into per-tile masks and counts
 *   compact  exclusive scans of the tile counts give the FullRX list (Inside then
 *            Border, ascending slot) and the capped recycling lists
 *   propose  every (parent p, direction rank r) attempt on a deformed cell does
 *            atomicMin(claim[target], p*32 + r); the key orders attempts exactly
 *            as the serial loop issues them
 *   decide   an attempt wins iff its key is the minimum; winners per parent and
 *            per destination list are counted
 *   commit   scans of those counts rank every winner among the infections into
 *            its list; rank k takes the k-th recycled slot, else appends
 *   retire   completed cells become ndefg+rxg, grain counts move
 *
 * FP64 throughout, compiled with -ffp-contract=off so every product and sum is
 * rounded exactly as in the reference's build without FMA.
 */
#include "gm_internal.h"

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

/* ------------------------------------------------- device-wide exclusive scan
 * Two u32 arrays scanned in place over n entries (n from *ndev when given).
 * Fixed grid: block b owns a contiguous chunk; phase 2 scans the block sums. */

static __device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        uint32_t t = __shfl_up(v, o, 64);
        if (lane >= o) v += t;
    }
    return v;
}

/* inclusive block scan of one value per thread (GM_BLOCK threads); returns the
 * inclusive prefix and writes the block total to *total */
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

__global__ void k_scan_p1(const uint32_t *a, const uint32_t *b, const uint32_t *ndev, uint32_t nhost, uint32_t *bsum) {
    uint32_t n = ndev ? *ndev : nhost;
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
    /* one block of 1024 threads, nb <= 1024 */
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

__global__ void k_scan_p3(uint32_t *a, uint32_t *b, const uint32_t *ndev, uint32_t nhost, const uint32_t *bsum) {
    uint32_t n = ndev ? *ndev : nhost;
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

static unsigned scan_blocks(uint32_t nupper) {
    uint32_t b = (nupper + GM_BLOCK - 1) / GM_BLOCK;
    if (b > GM_SCAN_BLOCKS) b = GM_SCAN_BLOCKS;
    if (b < 1) b = 1;
    return b;
}

/* exclusive scan in place; totals land in st->scan_tot[0..1] */
static hipError_t scan2(const GmDev &d, uint32_t *a, uint32_t *b, const uint32_t *ndev, uint32_t nhost, uint32_t nupper,
                        hipStream_t s) {
    unsigned nb = scan_blocks(nupper);
    k_scan_p1<<<nb, GM_BLOCK, 0, s>>>(a, b, ndev, nhost, d.scan_bsum);
    k_scan_p2<<<1, 1024, 0, s>>>(d.scan_bsum, nb, d.st->scan_tot);
    k_scan_p3<<<nb, GM_BLOCK, 0, s>>>(a, b, ndev, nhost, d.scan_bsum);
    return hipGetLastError();
}

/* ------------------------------------------------------------------ step */

__global__ void k_begin_step(GmStats *st) {
    st->sv = 0;
    st->npend = 0;
    st->rhomax_bits = 0ull;
    st->winI = 0;
    st->winB = 0;
    st->pone = 0;
}

hipError_t gm_launch_begin_step(const GmDev &d, hipStream_t s) {
    k_begin_step<<<1, 1, 0, s>>>(d.st);
    return hipGetLastError();
}

/* calcGrowthStepInside/Border (SCORE_Kernel.cpp:4046-4130, :4183-4249), one wave per 64 slots */
__global__ void __launch_bounds__(GM_BLOCK) k_sweep(GmDev d, GmSweepParams sp, uint32_t nI, uint32_t nB, uint32_t tI,
                                                    uint32_t T) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t t = blockIdx.x * (GM_BLOCK / 64) + (threadIdx.x >> 6);
    const bool inTile = t < T;
    uint32_t s, lim;
    if (t < tI) { s = t * GM_TILE + lane; lim = nI; }
    else { s = d.capI + (t - tI) * GM_TILE + lane; lim = d.capI + nB; }
    const bool valid = inTile && s < lim;
    bool a = false, full = false;
    if (valid) {
        a = d.act[s] != 0;
        if (a) {
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
            if (p > GM_DOUBLE_ACCURACY) {
                double v = m * p;
                v *= c_shape[d.inf[s]];
                double f = d.frac[s] + ((v * sp.dt) * sp.inv_dx);
                d.frac[s] = f;
                full = f >= 1.0;
            }
        }
    }
    /* Sv needs no counter of its own: it is the scanned slots minus the INACTIVE ones */
    uint64_t bf = __ballot(full);
    uint64_t br = __ballot(valid && !a);
    if (inTile && lane == 0) {
        d.mfull[t] = bf;
        d.mrec[t] = br;
        d.tcntF[t] = (uint32_t)__popcll(bf);
        d.tcntR[t] = (uint32_t)__popcll(br);
    }
}

hipError_t gm_launch_sweep(const GmDev &d, const GmSweepParams &sp, uint32_t nI, uint32_t nB, hipStream_t s) {
    uint32_t tI = (nI + GM_TILE - 1) / GM_TILE, tB = (nB + GM_TILE - 1) / GM_TILE, T = tI + tB;
    if (T == 0) return hipSuccess;
    unsigned nblk = (T + (GM_BLOCK / 64) - 1) / (GM_BLOCK / 64);
    k_sweep<<<nblk, GM_BLOCK, 0, s>>>(d, sp, nI, nB, tI, T);
    return hipGetLastError();
}

/* list lengths after the tile scans */
__global__ void k_compact_stats(GmDev d, uint32_t tI, uint32_t T) {
    GmStats *st = d.st;
    uint32_t totF = st->scan_tot[0], totR = st->scan_tot[1];
    uint32_t fI = tI < T ? d.tcntF[tI] : totF;
    uint32_t rI = tI < T ? d.tcntR[tI] : totR;
    st->nfull = totF;
    st->nfullI = fI;
    st->sv = (st->nI + st->nB) - totR;
    st->totRecI = rI;
    st->totRecB = totR - rI;
    st->recLenI = rI < d.recCapI ? rI : d.recCapI;
    st->recLenB = (totR - rI) < d.recCapB ? (totR - rI) : d.recCapB;
}

/* append_to_fullrx_*_list / append_to_recycling_*_list (SCORE_Kernel.cpp:3994-4042, :4133-4180)
 * as ordered compaction: FullRX holds global slots, Inside first; recycling lists are capped */
__global__ void __launch_bounds__(GM_BLOCK) k_compact(GmDev d, uint32_t tI, uint32_t T) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t t = blockIdx.x * (GM_BLOCK / 64) + (threadIdx.x >> 6);
    if (t >= T) return;
    const uint64_t bf = d.mfull[t], br = d.mrec[t];
    const uint32_t s = t < tI ? t * GM_TILE + lane : d.capI + (t - tI) * GM_TILE + lane;
    if ((bf >> lane) & 1ull) d.full[d.tcntF[t] + lanemask_lt_count(bf)] = s;
    if ((br >> lane) & 1ull) {
        uint32_t r = d.tcntR[t] + lanemask_lt_count(br);
        if (t < tI) {
            if (r < d.recCapI) d.recI[r] = s;
        } else {
            r -= d.st->totRecI;
            if (r < d.recCapB) d.recB[r] = s - d.capI;
        }
    }
}

hipError_t gm_launch_compact(const GmDev &d, uint32_t nI, uint32_t nB, hipStream_t s) {
    uint32_t tI = (nI + GM_TILE - 1) / GM_TILE, tB = (nB + GM_TILE - 1) / GM_TILE, T = tI + tB;
    if (T == 0) {
        /* nothing in either list: the stats still need zero lengths */
        k_fill_u32<<<1, 32, 0, s>>>(d.st->scan_tot, 0u, 2);
        k_compact_stats<<<1, 1, 0, s>>>(d, 0, 0);
        return hipGetLastError();
    }
    hipError_t e = scan2(d, d.tcntF, d.tcntR, nullptr, T, T, s);
    if (e != hipSuccess) return e;
    k_compact_stats<<<1, 1, 0, s>>>(d, tI, T);
    unsigned nblk = (T + (GM_BLOCK / 64) - 1) / (GM_BLOCK / 64);
    k_compact<<<nblk, GM_BLOCK, 0, s>>>(d, tI, T);
    return hipGetLastError();
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
    const uint32_t nfull = d.st->nfull;
    const uint32_t lane = threadIdx.x & 31;
    const uint32_t ngroups = gridDim.x * (GM_BLOCK / 32);
    for (uint32_t p = blockIdx.x * (GM_BLOCK / 32) + (threadIdx.x >> 5); p < nfull; p += ngroups) {
        if (lane >= 26) continue;
        const uint32_t s = d.full[p];
        uint32_t t;
        bool toB;
        if (!attempt_target(d, s >= d.capI, d.cell[s], (int)lane, t, toB)) continue;
        if (d.grid[t] < d.ndefg) atomicMin(&d.claim[t], p * 32u + lane);
    }
}

/* winners per parent and destination list, as 26-bit masks over the ranks */
__global__ void __launch_bounds__(GM_BLOCK) k_decide(GmDev d) {
    const uint32_t nfull = d.st->nfull;
    const uint32_t lane = threadIdx.x & 31;
    const uint32_t half = threadIdx.x & 32;
    const uint32_t ngroups = gridDim.x * (GM_BLOCK / 32);
    for (uint32_t p = blockIdx.x * (GM_BLOCK / 32) + (threadIdx.x >> 5); p < nfull; p += ngroups) {
        bool winI = false, winB = false;
        if (lane < 26) {
            const uint32_t s = d.full[p];
            uint32_t t;
            bool toB;
            if (attempt_target(d, s >= d.capI, d.cell[s], (int)lane, t, toB)) {
                if (d.grid[t] < d.ndefg && d.claim[t] == p * 32u + lane) {
                    winB = toB;
                    winI = !toB;
                }
            }
        }
        uint32_t mI = (uint32_t)(__ballot(winI) >> half);
        uint32_t mB = (uint32_t)(__ballot(winB) >> half);
        if (lane == 0) {
            d.pwinI[p] = mI;
            d.pwinB[p] = mB;
            d.pcntI[p] = (uint32_t)__popc(mI);
            d.pcntB[p] = (uint32_t)__popc(mB);
        }
    }
}

/* new front records: slot from recycling (ascending) else append
 * (omp_getNextFreeSlotIn*, SCORE_Kernel.cpp:3757-3874); record contents as
 * omp_infect_OutOfRXFront* write them (SCORE_Kernel.cpp:4281-4310, :4402-4426) */
__global__ void __launch_bounds__(GM_BLOCK) k_commit(GmDev d, GmPParams pp) {
    const uint32_t nfull = d.st->nfull;
    const uint32_t lane = threadIdx.x & 31;
    const uint32_t ngroups = gridDim.x * (GM_BLOCK / 32);
    const uint32_t recLenI = d.st->recLenI, recLenB = d.st->recLenB;
    const uint32_t nI = d.st->nI, nB = d.st->nB;
    for (uint32_t p = blockIdx.x * (GM_BLOCK / 32) + (threadIdx.x >> 5); p < nfull; p += ngroups) {
        if (lane >= 26) continue;
        const uint32_t mI = d.pwinI[p], mB = d.pwinB[p];
        const uint32_t bit = 1u << lane;
        if (!((mI | mB) & bit)) continue;
        const uint32_t s = d.full[p];
        const bool fromB = s >= d.capI;
        const uint32_t pos = d.cell[s];
        uint32_t t;
        bool toB;
        attempt_target(d, fromB, pos, (int)lane, t, toB);
        const bool intoI = (mI & bit) != 0;
        uint32_t gslot;
        if (intoI) {
            uint32_t rank = d.pcntI[p] + (uint32_t)__popc(mI & (bit - 1u));
            uint32_t rel = rank < recLenI ? d.recI[rank] : nI + (rank - recLenI);
            if (rel >= d.capI) { atomicOr(&d.st->err, GM_ERRF_INSIDE_FULL); continue; }
            gslot = rel;
        } else {
            uint32_t rank = d.pcntB[p] + (uint32_t)__popc(mB & (bit - 1u));
            uint32_t rel = rank < recLenB ? d.recB[rank] : nB + (rank - recLenB);
            if (rel >= d.capB) { atomicOr(&d.st->err, GM_ERRF_BORDER_FULL); continue; }
            gslot = d.capI + rel;
        }
        const double overshoot = d.frac[s] - 1.0;
        const int cls = c_cls[lane];
        const double carry = cls == 1 ? overshoot * 0.0682 : (cls == 2 ? overshoot * 0.0542 : overshoot * 0.0484);
        const uint32_t g = d.grid[t];
        const uint32_t prxg = d.rxg[s];
        double Pn = d.P[s];
        if (d.defg[s] != g) Pn = pair_weight(d, pp, prxg, g);
        d.act[gslot] = 1;
        d.inf[gslot] = c_code[lane];
        d.cell[gslot] = t;
        d.frac[gslot] = carry;
        d.P[gslot] = Pn;
        d.defg[gslot] = g;
        d.rxg[gslot] = prxg;
        d.grid[t] = GM_INFECTED;
        d.claim[t] = GM_FREE_CLAIM;
    }
}

/* completed cells become transformed (SCORE_Kernel.cpp:4640-4653) */
__global__ void k_retire(GmDev d) {
    const uint32_t nfull = d.st->nfull;
    for (uint32_t p = blockIdx.x * blockDim.x + threadIdx.x; p < nfull; p += gridDim.x * blockDim.x) {
        const uint32_t s = d.full[p];
        const uint32_t r = d.rxg[s];
        d.grid[d.cell[s]] = d.ndefg + r;
        atomicAdd(&d.rxcnt[r], 1u);
        atomicSub(&d.defcnt[d.defg[s]], 1u);
        d.act[s] = 0;
        d.frac[s] = 0.0;
    }
}

__global__ void k_step_tail(GmDev d) {
    GmStats *st = d.st;
    uint32_t wI = st->scan_tot[0], wB = st->scan_tot[1];
    st->winI = wI;
    st->winB = wB;
    if (wI > st->recLenI) st->nI += wI - st->recLenI;
    if (wB > st->recLenB) st->nB += wB - st->recLenB;
}

static unsigned parent_blocks(uint32_t nupper) {
    uint32_t b = (nupper + (GM_BLOCK / 32) - 1) / (GM_BLOCK / 32);
    if (b > 4096) b = 4096;
    if (b < 1) b = 1;
    return b;
}

/* propose, decide, scan, commit, retire; the number of parents is read on the device */
hipError_t gm_launch_infect(const GmDev &d, const GmPParams &pp, hipStream_t s) {
    /* parents are at most the occupied slots of both lists */
    unsigned nb = parent_blocks(d.parentCap);
    k_propose<<<nb, GM_BLOCK, 0, s>>>(d);
    k_decide<<<nb, GM_BLOCK, 0, s>>>(d);
    hipError_t e = scan2(d, d.pcntI, d.pcntB, &d.st->nfull, 0, d.parentCap, s);
    if (e != hipSuccess) return e;
    k_commit<<<nb, GM_BLOCK, 0, s>>>(d, pp);
    unsigned nr = fill_blocks(d.parentCap);
    k_retire<<<nr, GM_BLOCK, 0, s>>>(d);
    k_step_tail<<<1, 1, 0, s>>>(d);
    return hipGetLastError();
}

/* update_myrhomax (SCORE_RecoveryModels.cpp:146-165): largest rho with cells left */
__global__ void k_rhomax(GmDev d) {
    for (uint32_t g = blockIdx.x * blockDim.x + threadIdx.x; g < d.ndefg; g += gridDim.x * blockDim.x)
        if (d.defcnt[g] > 0) atomicMax(&d.st->rhomax_bits, (unsigned long long)__double_as_longlong(d.rho[g]));
}

hipError_t gm_launch_rhomax(const GmDev &d, hipStream_t s) {
    k_rhomax<<<fill_blocks(d.ndefg), GM_BLOCK, 0, s>>>(d);
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
        d.pcntI[k] = (win && in) ? 1u : 0u;
        d.pcntB[k] = (win && !in) ? 1u : 0u;
        d.pwinI[k] = win ? 1u : 0u;
    }
}

__global__ void k_nuc_commit(GmDev d, const uint32_t *site, const double *nucP, uint32_t n) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        if (!d.pwinI[k]) continue;
        uint32_t w = site[k];
        bool in = site_interior(d, w);
        uint32_t gslot;
        if (in) {
            gslot = d.pcntI[k];
            if (gslot >= d.capI) { atomicOr(&d.st->err, GM_ERRF_INSIDE_FULL); continue; }
        } else {
            uint32_t rel = d.pcntB[k];
            if (rel >= d.capB) { atomicOr(&d.st->err, GM_ERRF_BORDER_FULL); continue; }
            gslot = d.capI + rel;
        }
        uint32_t g = d.grid[w];
        d.act[gslot] = 1;
        d.inf[gslot] = 26;
        d.cell[gslot] = w;
        d.frac[gslot] = 1.0;
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
    hipError_t e = scan2(d, d.pcntI, d.pcntB, nullptr, n, n, s);
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

/* replace pending markers in the front by their cached values */
__global__ void k_resolve_front(GmDev d, uint32_t nI, uint32_t nB) {
    size_t total = (size_t)nI + nB;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x) {
        uint32_t s = i < nI ? (uint32_t)i : d.capI + (uint32_t)(i - nI);
        double P = d.P[s];
        if (is_pending(P)) d.P[s] = d.hvals[(uint32_t)__double_as_longlong(P)];
    }
}

hipError_t gm_launch_resolve_front(const GmDev &d, uint32_t nI, uint32_t nB, hipStream_t s) {
    if ((size_t)nI + nB == 0) return hipSuccess;
    k_resolve_front<<<fill_blocks((size_t)nI + nB), GM_BLOCK, 0, s>>>(d, nI, nB);
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
