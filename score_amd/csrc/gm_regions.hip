/*
 * gm_regions.hip -- the growth step with several memory regions.
 *
 * SCORE run with OMP_NUM_THREADS = T > 1 splits the domain into T regions
 * (ompshar_ca2regions, SCORE_GridPartitioning.cpp:69-166), each with its own
 * Inside and Border lists, and runs their growth steps side by side
 * (SCORE_Kernel.cpp:6634-6654): each region sweeps its lists, lets its
 * completed Inside cells infect, then its completed Border cells.  An infection
 * aimed at a cell of another region is parked in a halo of the infecting
 * region, first come first served (omp_infect_OutOfRXFrontBorder, :4494-4555);
 * after the barrier each region takes up what its neighbours parked for it,
 * neighbour by neighbour in its mynborregions order and each halo in parking
 * order, into its Border list, with the weight looked up again
 * (ompshar_synchronize_haloregions_rx, :4777-4872).
 *
 * That outcome is one claim per target cell with a 32-bit key:
 *   attempt of the cell's own region    p*32 + r
 *   attempt parked in a halo            2^31 + n*2^28 + p*32 + r
 * p is the parent's position in the step's FullRX order (below 2^23), which
 * runs region by region, each Inside then Border, so that p orders one region's
 * parents as its serial loop does; r is the attempt's rank and n the rank of
 * the receiving region's neighbour index of the halo among the eight a halo can
 * have (3, 4, 5, 6, 15, 16, 17, 18).  Only a cell's own region attempts it directly,
 * so the smallest key is the serial winner: the own region's first attempt,
 * else the first parked one in the receiving region's order.  A halo cell,
 * once parked, is never parked again (its status is never reset, :4527-4531);
 * that can only matter for a cell that is no longer deformed, which no attempt
 * here reaches (it left the deformed bitmap).
 *
 * The 2T lists keep SCORE's slot numbering: a new record takes its list's k-th
 * recycled slot if there is one, else appends, k being its rank among the
 * list's new records in serial order (for a Border list: the region's own
 * infections, then the halo winners by neighbour).  Ranks come from prefix
 * counts over the parents, less the count at the region's first parent; the
 * infection passes give each block the parents of one region only, so that
 * count is the offset of the region's first block.
 *
 * Per step, in the profiling slots of the one-region kernels:
 *   k_reanchor (shared)   k_rg_sweep   k_rg_sweep_scan (+ per-list starts)
 *   k_rg_compact   k_rg_propose   k_rg_decide
 *   k_rg_decide_scan (+ per-region starts, list lengths)   k_rg_recycle   k_rg_commit
 */
#include <algorithm>

#include "gm_device.h"

/* opposite neighbour direction (mynborregions order, ompshar_init_neighbortopology,
 * SCORE_Kernel.cpp:2028-2080) */
static __constant__ uint8_t c_opp[27] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15,
                                         26, 25, 24, 23, 22, 21, 20, 19};
/* the halos an infection can be parked in, and each direction's index among them */
static __constant__ uint8_t c_hdir[GM_NHALO] = {3, 4, 5, 6, 15, 16, 17, 18};
/* neighbour index of the direction (0, sy, sz), at [(sy+1)*3 + sz+1] */
static __constant__ uint8_t c_ydir_nb[9] = {15, 3, 16, 5, 0, 6, 17, 4, 18};
static __constant__ uint8_t c_hidx[27] = {0xFF, 0xFF, 0xFF, 0, 1, 2, 3, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF,
                                          0xFF, 4, 5, 6, 7, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF, 0xFF};

/* rows of the sweep's per-block counts, and where the infection pass's rows
 * start (after them: the recycling pass reads the sweep's offsets last) */
#define RS_F 0
#define RS_R 1
#define RS_UN 2
#define RW_ROW0 3

/* the part_of block size for a grid of nb blocks, from any thread */
static __device__ __forceinline__ uint32_t part_per(uint32_t n, uint32_t nb, uint32_t q) {
    uint32_t per = (n + nb - 1u) / nb;
    return (per + q - 1u) / q * q;
}

/* Exclusive scan, in place, of `rows` rows of nb <= GM_MAX_BLOCKS per-block
 * counts by one block of 1024 threads, as scan_block_counts does but one row
 * at a time, reading each thread's eight entries again (and scanning them
 * across the wave again) for the offsets, so that few registers stay live and
 * no private array is indexed at run time: a kernel that spills to scratch pays
 * for its scratch set-up at every launch.  tot (shared) receives the totals. */
static __device__ uint32_t sum8(const uint32_t *row, uint32_t i0, uint32_t nb) {
    if (i0 >= nb) return 0u;
    const uint4 *p = reinterpret_cast<const uint4 *>(row + i0);
    const uint4 a = p[0], b = p[1];
    const uint32_t v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
    uint32_t s = 0;
#pragma unroll
    for (uint32_t j = 0; j < 8; j++) s += i0 + j < nb ? v[j] : 0u;
    return s;
}

static __device__ void scan_rows_lean(uint32_t *bcnt, uint32_t rows, uint32_t nb, uint32_t *tot) {
    __shared__ uint32_t ws[16][16];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const uint32_t i0 = threadIdx.x * 8u;
    for (uint32_t c = 0; c < rows; c++) {
        const uint32_t incl = wave_incl_scan(sum8(bcnt + c * GM_MAX_BLOCKS, i0, nb));
        if (lane == 63) ws[w][c] = incl;
    }
    __syncthreads();
    for (uint32_t c = 0; c < rows; c++) {
        uint32_t pre = 0, t = 0;
        for (uint32_t i = 0; i < 16; i++) {
            if (i < w) pre += ws[i][c];
            t += ws[i][c];
        }
        if (threadIdx.x == 0) tot[c] = t;
        uint32_t *row = bcnt + c * GM_MAX_BLOCKS;
        const uint32_t sum = sum8(row, i0, nb);
        const uint32_t incl = wave_incl_scan(sum);
        if (i0 >= nb) continue;
        uint4 *q = reinterpret_cast<uint4 *>(row + i0);
        const uint4 a = q[0], b = q[1];
        const uint32_t v[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
        uint32_t run = pre + incl - sum, o[8];
#pragma unroll
        for (uint32_t j = 0; j < 8; j++) {
            o[j] = run;
            run += i0 + j < nb ? v[j] : 0u;
        }
        q[0] = make_uint4(o[0], o[1], o[2], o[3]);
        q[1] = make_uint4(o[4], o[5], o[6], o[7]);
    }
    __syncthreads();
}

/* grid of the infection passes: at least one block more than there are regions,
 * so that every region's parents fit in whole blocks of GM_PCHUNK multiples */
static uint32_t rg_infect_blocks(const GmDev &d, uint32_t parents_upper) {
    const uint32_t nb = grid_for(parents_upper, GM_PCHUNK);
    return nb > d.nreg ? nb : d.nreg + 1u;
}

/* The parents of infection-pass block b: blocks are handed out region by region
 * (L.bstart[r] is region r's first block, st->rg_per parents each), so a block
 * holds one region's parents only; r is that region. */
static __device__ __forceinline__ void rg_part(const GmDev &d, uint32_t &lo, uint32_t &hi, uint32_t &r) {
    const uint32_t b = blockIdx.x;
    lo = hi = r = 0;
    r = d.L.breg[b];
    if (r >= d.nreg) return; /* a block past the ones in use */
    const uint32_t per = d.st->rg_per, end = d.L.par[2u * r + 2u];
    lo = d.L.par[2u * r] + (b - d.L.bstart[r]) * per;
    hi = lo + per < end ? lo + per : end;
}

/* sweep tiles of a list of n occupied slots: whole 8-tile chunks */
static __device__ __forceinline__ uint32_t list_tiles(uint32_t n) {
    return (n + 64u * GM_SW_TPC - 1u) / (64u * GM_SW_TPC) * GM_SW_TPC;
}

struct RgTarget {
    uint32_t x, y, z;
    uint32_t dl;                /* list the new record goes to */
    uint32_t hidx;              /* halo index among the GM_NHALO, for a parked attempt */
    uint32_t khi;               /* upper part of the claim key: 0 for the own region */
};

/* Target of attempt rank rk of a parent at (x,y,z) in list l, routed as the
 * reference does.  Inside parents (omp_infect_OutOfRXFrontInside,
 * SCORE_Kernel.cpp:4252-4344) stay in their region and send every target to its
 * Border list unless they lie more than one cell inside each face of the
 * region (:4606-4609).  Border parents (:4347-4556) wrap x and y, drop targets
 * outside z, send targets strictly inside the region to its Inside list and the
 * region's other targets to its Border list; a target in another region goes to
 * the first halo towards that region whose box, taken one cell beyond its
 * extent as the reference's test does (:4500-4512), holds it.  False: no
 * target. */
/* a region's box, loaded once by the caller for all attempts of its parents */
struct RgBox {
    int lo0, lo1, lo2, hi0, hi1, hi2;
};

/* region r's box from the region grid alone (as build_regions lays it out), so
 * that no load waits for it */
static __device__ __forceinline__ RgBox rg_box(const GmDev &d, uint32_t r) {
    const uint32_t iy = r % d.npy, iz = r / d.npy;
    const int lo1 = (int)(iy * d.nry), lo2 = (int)(iz * d.nrz);
    const int hi1 = iy + 1u == d.npy ? (int)d.ny - 1 : lo1 + (int)d.nry - 1;
    const int hi2 = iz + 1u == d.npz ? (int)d.nz - 1 : lo2 + (int)d.nrz - 1;
    return RgBox{0, lo1, lo2, (int)d.nx - 1, hi1, hi2};
}

static __device__ __forceinline__ bool rg_target(const GmDev &d, const RgBox &B, uint32_t l, uint32_t x, uint32_t y,
                                                  uint32_t z, int rk, RgTarget &o) {
    const uint32_t r = l >> 1;
    int tx = (int)x + c_dx[rk], ty = (int)y + c_dy[rk], tz = (int)z + c_dz[rk];
    const int nx = (int)d.nx, ny = (int)d.ny, nz = (int)d.nz;
    const int lo0 = B.lo0, lo1 = B.lo1, lo2 = B.lo2, hi0 = B.hi0, hi1 = B.hi1, hi2 = B.hi2;
    o.khi = 0u;
    o.hidx = GM_NHALO;
    if (!(l & 1u)) {
        /* an Inside cell is strictly inside its region: this never fires, and
         * keeps a corrupted record from addressing outside the grid */
        if (tx < 0 || tx >= nx || ty < 0 || ty >= ny || tz < 0 || tz >= nz) return false;
        const bool deep = (int)x > lo0 + 1 && (int)x < hi0 - 1 && (int)y > lo1 + 1 && (int)y < hi1 - 1 &&
                          (int)z > lo2 + 1 && (int)z < hi2 - 1;
        o.dl = deep ? 2u * r : 2u * r + 1u;
    } else {
        if (tx < 0) tx += nx;
        if (tx >= nx) tx -= nx;
        if (ty < 0) ty += ny;
        if (ty >= ny) ty -= ny;
        if (tz < 0 || tz >= nz) return false;
        if (tx > lo0 && tx < hi0 && ty > lo1 && ty < hi1 && tz > lo2 && tz < hi2) {
            o.dl = 2u * r;
        } else {
            /* the target lies in the region or next to it: the neighbour in the
             * direction it left along y (none with one slab along y, whose
             * region spans every y) and z */
            const int uy = (int)y + c_dy[rk], uz = (int)z + c_dz[rk];
            const int sy = d.npy == 1u ? 0 : (uy < lo1 ? -1 : (uy > hi1 ? 1 : 0));
            const int sz = uz < lo2 ? -1 : (uz > hi2 ? 1 : 0);
            const GmRegion &R = d.reg[r];
            const uint32_t q = R.nbor[c_ydir_nb[(sy + 1) * 3 + (sz + 1)]];
            if (q == r) {
                o.dl = 2u * r + 1u;
            } else {
                uint32_t h = 0;
                for (uint32_t c = 1; c < 27; c++) {
                    const GmHalo &H = R.halo[c];
                    if (H.inregion != q || H.ext[0] == 0) continue;
                    if (tx >= (int)H.org[0] && tx <= (int)H.org[0] + (int)H.ext[0] && ty >= (int)H.org[1] &&
                        ty <= (int)H.org[1] + (int)H.ext[1] && tz >= (int)H.org[2] && tz <= (int)H.org[2] + (int)H.ext[2]) {
                        h = c;
                        break;
                    }
                }
                const uint32_t hx = c_hidx[h];
                if (hx >= GM_NHALO) {
                    atomicOr(&d.st->err, GM_ERRF_HALO);
                    return false;
                }
                o.hidx = hx;
                o.dl = 2u * q + 1u;
                o.khi = (1u << 31) | ((uint32_t)c_hidx[c_opp[h]] << 28);
            }
        }
    }
    o.x = (uint32_t)tx;
    o.y = (uint32_t)ty;
    o.z = (uint32_t)tz;
    return true;
}

/* ------------------------------------------------------------- list tables */

/* One block of 1024 threads: the lists' occupied and tile prefixes from their
 * lengths, and the lengths summed over Inside and Border lists; thread t takes
 * region t's two lists (at most 1024 regions). */
static __device__ void rg_prefix_block(const GmDev &d) {
    __shared__ uint32_t ws[16][4];
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6;
    const bool v = t < d.nreg;
    const uint32_t nI = v ? d.L.n[2u * t] : 0u, nB = v ? d.L.n[2u * t + 1u] : 0u;
    const uint32_t tI = list_tiles(nI), tB = list_tiles(nB);
    const uint32_t val[4] = {nI + nB, tI + tB, nI, nB};
    uint32_t incl[4];
#pragma unroll
    for (int c = 0; c < 4; c++) {
        incl[c] = wave_incl_scan(val[c]);
        if (lane == 63u) ws[w][c] = incl[c];
    }
    __syncthreads();
    uint32_t pre[4] = {0, 0, 0, 0}, tot[4] = {0, 0, 0, 0};
    for (uint32_t i = 0; i < 16u; i++)
#pragma unroll
        for (int c = 0; c < 4; c++) {
            if (i < w) pre[c] += ws[i][c];
            tot[c] += ws[i][c];
        }
    if (v) {
        const uint32_t o = pre[0] + incl[0] - val[0], ti = pre[1] + incl[1] - val[1];
        d.L.occ[2u * t] = o;
        d.L.occ[2u * t + 1u] = o + nI;
        d.L.tile[2u * t] = ti;
        d.L.tile[2u * t + 1u] = ti + tI;
    }
    if (t == 0) {
        d.L.occ[d.nlists] = tot[0];
        d.L.tile[d.nlists] = tot[1];
        d.st->nI = tot[2];
        d.st->nB = tot[3];
    }
    __syncthreads();
}

__global__ void __launch_bounds__(1024) k_rg_prefix(GmDev d) { rg_prefix_block(d); }

hipError_t gm_launch_rg_prefix(const GmDev &d, hipStream_t s) {
    k_rg_prefix<<<1, 1024, 0, s>>>(d);
    return hipGetLastError();
}

/* nuclei placed by the host: records and codes scattered to their slots, and
 * their cells taken out of the deformed bitmap */
__global__ void k_rg_place(GmDev d, const uint32_t *slots, const GmRec *recs, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t s = slots[i];
        const GmRec r = recs[i];
        d.rec[s] = r;
        d.cstep[s] = (gm_cs_t)GM_CS_UNKNOWN;
        const uint32_t w = r.cell;
        const uint32_t z = w / d.nxy, rem = w - z * d.nxy, y = rem / d.nx, x = rem - y * d.nx;
        clear_deformed(d, x, y, z);
    }
}

hipError_t gm_launch_rg_place(const GmDev &d, const uint32_t *slots, const GmRec *recs, uint32_t n, hipStream_t s) {
    if (n) k_rg_place<<<fill_blocks(n), GM_BLOCK, 0, s>>>(d, slots, recs, n);
    return hipGetLastError();
}

/* ------------------------------------------------------------------- sweep */

/* Pass over the occupied slots of every list, as the one-region sweep does:
 * the lists' tiles follow one another, each list in whole 8-tile chunks, so a
 * chunk belongs to one list.  Completed and INACTIVE slots go into per-tile
 * masks, counted per block. */
__global__ void __launch_bounds__(GM_BLOCK) k_rg_sweep(GmDev d, GmSweepArgs A0) {
    __shared__ uint32_t red[GM_BLOCK / 64][3];
    __shared__ uint64_t sF[GM_BLOCK / 64][8], sR[GM_BLOCK / 64][8];
    if (d.st->halt) return;
    const GmSweepArgs A = step_args(d, A0);
    const uint32_t T = d.L.tile[d.nlists];
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t lo, hi;
    part_of(T, GM_SW_TPC, lo, hi);
    const uint32_t c_lo = lo / GM_SW_TPC, c_hi = hi / GM_SW_TPC; /* T is a whole number of chunks */
    const uint32_t kc = A.k & (GM_CS_MOD - 1u), horizon = d.horizon;
    uint32_t cF = 0, cR = 0, unknown = 0;
    for (uint32_t c = c_lo + w; c < c_hi; c += GM_BLOCK / 64) {
        const uint32_t tile0 = c * GM_SW_TPC;
        const uint32_t l = upper_index(d.L.tile, d.nlists, tile0);
        const uint32_t s0 = d.L.base[l] + (tile0 - d.L.tile[l]) * 64u;
        const uint32_t base = s0 + 8u * lane;
        const uint32_t lim = d.L.base[l] + d.L.n[l];
        const uint32_t nocc = lim > base ? (lim - base < 8u ? lim - base : 8u) : 0u;
        const uint4 cs = nocc ? *reinterpret_cast<const uint4 *>(d.cstep + base)
                              : make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
        const uint32_t pack = sweep_classify(cs, nocc, kc, horizon, unknown);
        cF += (uint32_t)__popc(pack & 0xFFu);
        cR += (uint32_t)__popc(pack >> 8);
        uint64_t myF, myR;
        sweep_masks(pack, sF[w], sR[w], myF, myR);
        if (lane < GM_SW_TPC) {
            d.mF[tile0 + lane] = myF;
            d.mR[tile0 + lane] = myR;
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        unknown += __shfl_down(unknown, o, 64);
        cF += __shfl_down(cF, o, 64);
        cR += __shfl_down(cR, o, 64);
    }
    if (lane == 0) {
        red[w][RS_F] = cF;
        red[w][RS_R] = cR;
        red[w][RS_UN] = unknown;
    }
    __syncthreads();
    if (threadIdx.x < 3) {
        uint32_t t = 0;
        for (int i = 0; i < GM_BLOCK / 64; i++) t += red[i][threadIdx.x];
        d.bcnt[threadIdx.x * GM_MAX_BLOCKS + blockIdx.x] = t;
    }
}

/* One block: per-block offsets; then each list's first completed and first
 * INACTIVE position (the block offset of the tile where the list starts plus
 * the block's tiles before it), the per-list counts, and their sums over the
 * Inside and the Border lists. */
__global__ void __launch_bounds__(1024) k_rg_sweep_scan(GmDev d, uint32_t nb, uint32_t nbd) {
    __shared__ uint32_t acc[6];
    __shared__ uint32_t ws[16];
    GmStats *st = d.st;
    if (st->halt) return;
    uint32_t tot[3];
    scan_block_counts<3>(d.bcnt, nb, tot);
    if (threadIdx.x < 6) acc[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        if (tot[RS_UN]) st->err |= GM_ERRF_MISSED;
        /* infection keys are (halo, neighbour, p, rank) in 32 bits, p < 2^23 */
        if ((uint64_t)tot[RS_F] > (1ull << 23)) st->err |= GM_ERRF_KEYS;
    }
    __syncthreads();
    const uint32_t T = d.L.tile[d.nlists];
    const uint32_t per = T ? part_per(T, nb, GM_SW_TPC) : 1u;
    /* one wave per list, the lanes sharing the block's tiles before the list's first */
    const uint32_t lane = threadIdx.x & 63u, wv = threadIdx.x >> 6;
    for (uint32_t l = wv; l <= d.nlists; l += 1024u / 64u) {
        const uint32_t t0 = d.L.tile[l];
        uint32_t ef = 0, er = 0, b = 0;
        if (t0 < T) {
            b = t0 / per;
            for (uint32_t t = b * per + lane; t < t0; t += 64) {
                ef += (uint32_t)__popcll(d.mF[t]);
                er += (uint32_t)__popcll(d.mR[t]);
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            ef += __shfl_down(ef, o, 64);
            er += __shfl_down(er, o, 64);
        }
        if (lane == 0) {
            d.L.tile_s[l] = t0; /* this step's tiles, for the recycling pass after the lengths change */
            d.L.par[l] = t0 < T ? d.bcnt[RS_F * GM_MAX_BLOCKS + b] + ef : tot[RS_F];
            d.L.rst[l] = t0 < T ? d.bcnt[RS_R * GM_MAX_BLOCKS + b] + er : tot[RS_R];
        }
    }
    __syncthreads();
    uint32_t a[6] = {0, 0, 0, 0, 0, 0};
    for (uint32_t l = threadIdx.x; l < d.nlists; l += 1024) {
        const uint32_t nf = d.L.par[l + 1] - d.L.par[l], nr = d.L.rst[l + 1] - d.L.rst[l];
        const uint32_t rl = nr < d.L.rcap[l] ? nr : d.L.rcap[l];
        d.L.rlen[l] = rl;
        const uint32_t k = (l & 1u) * 3u;
        a[k] += nf;
        a[k + 1] += nr;
        a[k + 2] += rl;
    }
#pragma unroll
    for (int c = 0; c < 6; c++)
        if (a[c]) atomicAdd(&acc[c], a[c]);
    /* the infection passes' blocks, region by region: rg_per parents each (a
     * multiple of GM_PCHUNK), few enough that all regions fit in nbd blocks */
    const uint32_t nfull = tot[RS_F];
    uint32_t pper = (nfull + (nbd - d.nreg) - 1u) / (nbd - d.nreg);
    pper = pper ? (pper + GM_PCHUNK - 1u) / GM_PCHUNK * GM_PCHUNK : GM_PCHUNK;
    const uint32_t r = threadIdx.x;
    const uint32_t nr = r < d.nreg ? d.L.par[2u * r + 2u] - d.L.par[2u * r] : 0u;
    const uint32_t blk = (nr + pper - 1u) / pper;
    const uint32_t incl = wave_incl_scan(blk);
    if (lane == 63u) ws[wv] = incl;
    __syncthreads();
    uint32_t pre = 0, totb = 0;
    for (uint32_t i = 0; i < 16u; i++) {
        if (i < wv) pre += ws[i];
        totb += ws[i];
    }
    if (r < d.nreg) d.L.bstart[r] = pre + incl - blk;
    if (threadIdx.x == 0) d.L.bstart[d.nreg] = totb;
    __syncthreads();
    /* each block's region; blocks past the ones in use find none */
    for (uint32_t b = threadIdx.x; b < nbd; b += 1024)
        d.L.breg[b] = b < totb ? upper_index(d.L.bstart, d.nreg, b) : 0xFFFFFFFFu;
    if (threadIdx.x == 0) {
        st->rg_per = pper;
        st->nfullI = acc[0];
        st->totRecI = acc[1];
        st->recLenI = acc[2];
        st->nfullB = acc[3];
        st->totRecB = acc[4];
        st->recLenB = acc[5];
    }
}

/* One wave per sweep block, one lane per tile: the completed slots in FullRX
 * order (append_to_fullrx_*_list, SCORE_Kernel.cpp:3994-4042, :4133-4180), the
 * lists one after another. */
__global__ void __launch_bounds__(64) k_rg_compact(GmDev d) {
    if (d.st->halt) return;
    const uint32_t lane = threadIdx.x;
    const uint32_t T = d.L.tile[d.nlists];
    uint32_t lo, hi;
    part_of(T, GM_SW_TPC, lo, hi);
    uint32_t runF = d.bcnt[RS_F * GM_MAX_BLOCKS + blockIdx.x];
    for (uint32_t batch = lo; batch < hi; batch += 64) {
        const uint32_t t = batch + lane;
        const uint64_t bf = t < hi ? d.mF[t] : 0ull;
        const uint32_t cf = (uint32_t)__popcll(bf);
        const uint32_t incf = wave_incl_scan(cf);
        uint32_t of = runF + incf - cf;
        runF += __shfl(incf, 63, 64);
        if (bf) {
            const uint32_t l = upper_index(d.L.tile, d.nlists, t);
            const uint32_t s0 = d.L.base[l] + (t - d.L.tile[l]) * 64u;
            for (uint64_t m = bf; m; m &= m - 1ull) {
                d.plist[of] = l;
                d.pslot[of++] = s0 + (uint32_t)__ffsll((long long)m) - 1u;
            }
        }
    }
}

/* Each list's INACTIVE slots in ascending order (append_to_recycling_*_list,
 * SCORE_Kernel.cpp:4037-4041, :4175-4179), written once the step's new records
 * per list are known and only as far as they are handed out: the first
 * min(new records, recycled slots) of each list.  The sweep's partition and
 * offsets; one wave per sweep block. */
__global__ void __launch_bounds__(64) k_rg_recycle(GmDev d) {
    if (d.st->halt) return;
    const uint32_t lane = threadIdx.x;
    uint32_t lo, hi;
    part_of(d.L.tile_s[d.nlists], GM_SW_TPC, lo, hi);
    uint32_t runR = d.bcnt[RS_R * GM_MAX_BLOCKS + blockIdx.x];
    for (uint32_t batch = lo; batch < hi; batch += 64) {
        const uint32_t t = batch + lane;
        const uint64_t br = t < hi ? d.mR[t] : 0ull;
        const uint32_t cr = (uint32_t)__popcll(br);
        const uint32_t incr = wave_incl_scan(cr);
        const uint32_t orr = runR + incr - cr;
        runR += __shfl(incr, 63, 64);
        if (br) {
            const uint32_t l = upper_index(d.L.tile_s, d.nlists, t);
            const uint32_t w = d.L.win[l], rl = d.L.rlen[l];
            const uint32_t lim = w < rl ? w : rl;
            uint32_t k = orr - d.L.rst[l];
            if (k < lim) {
                const uint32_t rel0 = (t - d.L.tile_s[l]) * 64u;
                uint32_t *rb = d.L.rbuf + d.L.rbase[l];
                for (uint64_t m = br; m && k < lim; m &= m - 1ull) rb[k++] = rel0 + (uint32_t)__ffsll((long long)m) - 1u;
            }
        }
    }
}

/* grid of the sweep, compaction and recycling passes for about nI + nB occupied slots */
static uint32_t rg_sweep_blocks(const GmDev &d, const GmSweepArgs &a) {
    const uint64_t tiles = (uint64_t)(a.nI + a.nB) / 64u + (uint64_t)GM_SW_TPC * (d.nlists + 1u);
    return grid_for((uint32_t)std::min<uint64_t>(tiles, 0xFFFFFFFFull), GM_SW_TPC * (GM_BLOCK / 64));
}

hipError_t gm_launch_rg_sweep(const GmDev &d, const GmSweepArgs &a, uint32_t parents_upper, hipStream_t s,
                              hipEvent_t *ev) {
    hipError_t e;
    GmSweepArgs A = a;
    A.reanchor = 0;
    if ((e = mark(ev, 0, s)) != hipSuccess) return e;
    if (a.reanchor && (e = gm_launch_reanchor(d, a, s)) != hipSuccess) return e;
    const uint32_t nb = rg_sweep_blocks(d, a);
    if ((e = mark(ev, 1, s)) != hipSuccess) return e;
    k_rg_sweep<<<nb, GM_BLOCK, 0, s>>>(d, A);
    if ((e = mark(ev, 2, s)) != hipSuccess) return e;
    k_rg_sweep_scan<<<1, 1024, 0, s>>>(d, nb, rg_infect_blocks(d, parents_upper));
    if ((e = mark(ev, 3, s)) != hipSuccess) return e;
    k_rg_compact<<<nb, 64, 0, s>>>(d);
    return hipGetLastError();
}

/* --------------------------------------------------------------- infection */

/* GM_PG lanes per parent, as the one-region propose pass: the parent's data
 * into the dense array (with its list in place of the Border flag), its
 * retirement into the log, and every attempt on a deformed cell claims it with
 * its 64-bit key. */
__global__ void __launch_bounds__(GM_BLOCK) k_rg_propose(GmDev d) {
    if (d.st->halt) return;
    const uint32_t nfull = d.st->nfullI + d.st->nfullB;
    const unsigned long long nret = d.st->nret;
    const bool logged = nret + nfull <= d.rlogCap;
    if (!logged && blockIdx.x == 0 && threadIdx.x == 0) atomicOr(&d.st->err, GM_ERRF_RLOG_FULL);
    const uint32_t lg = threadIdx.x & (GM_PG - 1);
    const uint32_t ngroups = gridDim.x * (GM_BLOCK / GM_PG);
    for (uint32_t p = blockIdx.x * (GM_BLOCK / GM_PG) + threadIdx.x / GM_PG; p < nfull; p += ngroups) {
        uint4 q2 = make_uint4(0u, 0u, 0u, 0u);
        if (lg == 0) {
            const uint32_t s = d.pslot[p];
            const GmRec *r = d.rec + s;
            const RecC c = rec_c(r);
            const uint32_t rxg = r->rxg;
            const uint32_t cell = c.cell;
            const uint32_t z = cell / d.nxy, rem = cell - z * d.nxy, y = rem / d.nx, x = rem - y * d.nx;
            q2 = make_uint4(x, y, z, d.plist[p]);
            par_plane(d, GM_PAR_IDS)[p] = make_uint4(s, cell, c.defg, rxg);
            /* the value after the completing addition follows in k_rg_decide */
            reinterpret_cast<double2 *>(par_plane(d, GM_PAR_FP))[p] = make_double2(0.0, c.P);
            par_plane(d, GM_PAR_XYZ)[p] = q2;
            if (logged) d.rlog[nret + p] = make_uint2(cell, d.ndefg + rxg);
        }
        q2.x = (uint32_t)__shfl((int)q2.x, 0, GM_PG);
        q2.y = (uint32_t)__shfl((int)q2.y, 0, GM_PG);
        q2.z = (uint32_t)__shfl((int)q2.z, 0, GM_PG);
        q2.w = (uint32_t)__shfl((int)q2.w, 0, GM_PG);
        const uint32_t gsh = (threadIdx.x & 63) & ~(GM_PG - 1);
        const RgBox box = rg_box(d, q2.w >> 1);
        uint32_t am = 0;
#pragma unroll
        for (uint32_t i = 0; i < GM_PG_ITERS; i++) {
            const uint32_t rk = lg + GM_PG * i;
            bool att = false;
            if (rk < 26) {
                RgTarget o;
                if (rg_target(d, box, q2.w, q2.x, q2.y, q2.z, (int)rk, o) && is_deformed(d, o.x, o.y, o.z)) {
                    att = true;
                    atomicMin(&d.claim[cell_of(d, o.x, o.y, o.z)], o.khi | (p * 32u + rk));
                }
            }
            am |= (uint32_t)((__ballot(att) >> gsh) & ((1ull << GM_PG) - 1u)) << (GM_PG * i);
        }
        if (lg == 0) d.pmI[p] = am;
    }
}

/* One lane per parent: the attempts that hold their target's smallest key, as
 * masks by destination (own Inside, own Border, a halo), counted per block by
 * row (GM_RW_*); a parent that won a target gets its value after the completing
 * addition in the parent array.  The partition into blocks (one region per
 * block) is the commit pass's. */
__global__ void __launch_bounds__(64) k_rg_decide(GmDev d) {
    if (d.st->halt) return;
    const uint32_t k = d.st->kcur;
    uint32_t lo, hi, reg;
    rg_part(d, lo, hi, reg);
    const RgBox box = rg_box(d, reg);
    uint32_t cnt[GM_RROWS];
#pragma unroll
    for (int c = 0; c < GM_RROWS; c++) cnt[c] = 0;
    for (uint32_t p = lo + threadIdx.x; p < hi; p += 64) {
        const uint4 q2 = par_plane(d, GM_PAR_XYZ)[p];
        const uint32_t am = d.pmI[p];
        const uint32_t ps = d.pslot[p]; /* loaded with the rest, used if the parent wins */
        uint32_t mI = 0, mB = 0, mH = 0;
        for (uint32_t m = am; m; m &= m - 1u) {
            const uint32_t rk = (uint32_t)__ffs(m) - 1u;
            RgTarget o;
            rg_target(d, box, q2.w, q2.x, q2.y, q2.z, (int)rk, o);
            if (d.claim[cell_of(d, o.x, o.y, o.z)] != (o.khi | (p * 32u + rk))) continue;
            if (o.khi) {
                mH |= 1u << rk;
#pragma unroll
                for (uint32_t j = 0; j < GM_NHALO; j++) cnt[GM_RW_H0 + j] += o.hidx == j ? 1u : 0u;
            } else if (o.dl & 1u) {
                mB |= 1u << rk;
            } else {
                mI |= 1u << rk;
            }
        }
        d.pmI[p] = mI;
        d.pmB[p] = mB;
        d.pmH[p] = mH;
        if (mI | mB | mH) reinterpret_cast<double *>(par_plane(d, GM_PAR_FP))[2 * (size_t)p] = parent_fin(d, ps, k);
        cnt[GM_RW_I] += (uint32_t)__popc(mI);
        cnt[GM_RW_B] += (uint32_t)__popc(mB);
        cnt[GM_RW_A] += (uint32_t)__popc(am);
    }
#pragma unroll
    for (int c = 0; c < GM_RROWS; c++) {
        uint32_t v = cnt[c];
        for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
        if (threadIdx.x == 0) d.bcnt[(RW_ROW0 + c) * GM_MAX_BLOCKS + blockIdx.x] = v;
    }
}

/* the halo index of every parked winner of parent p (whose data is q2), as
 * GM_NHALO masks */
static __device__ __forceinline__ void halo_masks(const GmDev &d, const RgBox &box, const uint4 &q2, uint32_t mH,
                                                  uint32_t *mh) {
#pragma unroll
    for (uint32_t j = 0; j < GM_NHALO; j++) mh[j] = 0;
    for (uint32_t m = mH; m; m &= m - 1u) {
        const uint32_t rk = (uint32_t)__ffs(m) - 1u;
        RgTarget o;
        rg_target(d, box, q2.w, q2.x, q2.y, q2.z, (int)rk, o);
#pragma unroll
        for (uint32_t j = 0; j < GM_NHALO; j++) mh[j] |= o.hidx == j ? 1u << rk : 0u;
    }
}

/* One block: per-block offsets of every row; then, per region, the counts of
 * each row before its first parent (block offset plus the block's parents
 * before it); per region, where each neighbour's halo winners start in its
 * Border list (after its own winners, in its neighbour order: the infecting
 * region S = nbor[nb] parked them in its halo opp(nb)); the new records per
 * list, the lists' new lengths (recycled slots first, then appended) and their
 * prefixes for the next step. */
__global__ void __launch_bounds__(1024) k_rg_decide_scan(GmDev d, uint32_t nb) {
    __shared__ uint32_t acc[2];
    GmStats *st = d.st;
    if (st->halt) return;
    /* only the blocks in use hold counts */
    const uint32_t nused = d.L.bstart[d.nreg];
    __shared__ uint32_t tot[GM_RROWS];
    scan_rows_lean(d.bcnt + RW_ROW0 * GM_MAX_BLOCKS, GM_RROWS, nused < nb ? nused : nb, tot);
    if (threadIdx.x < 2) acc[threadIdx.x] = 0;
    if (threadIdx.x == 0) st->nattempt = tot[GM_RW_A];
    /* each region's counts at its first parent: the offset of its first block
     * (the next region's, or the totals, for a region without parents) */
    for (uint32_t r = threadIdx.x; r <= d.nreg; r += 1024) {
        const uint32_t b = d.L.bstart[r];
#pragma unroll
        for (int c = 0; c < GM_RROWS; c++)
            d.L.rs[r * GM_RROWS + c] = b < nused ? d.bcnt[(RW_ROW0 + c) * GM_MAX_BLOCKS + b] : tot[c];
    }
    __syncthreads();
    for (uint32_t q = threadIdx.x; q < d.nreg; q += 1024) {
        /* every load first (the stores below could alias them), then the prefix */
        const uint32_t *a = d.L.rs + q * GM_RROWS, *b = a + GM_RROWS;
        const uint32_t dI = b[GM_RW_I] - a[GM_RW_I], dB = b[GM_RW_B] - a[GM_RW_B];
        uint32_t S[GM_NHALO], cnt[GM_NHALO];
#pragma unroll
        for (uint32_t j = 0; j < GM_NHALO; j++) S[j] = d.reg[q].nbor[c_hdir[j]];
#pragma unroll
        for (uint32_t j = 0; j < GM_NHALO; j++) {
            const uint32_t hs = GM_RW_H0 + c_hidx[c_opp[c_hdir[j]]];
            cnt[j] = S[j] != q ? d.L.rs[(S[j] + 1u) * GM_RROWS + hs] - d.L.rs[S[j] * GM_RROWS + hs] : 0u;
        }
        uint32_t run = dB;
#pragma unroll
        for (uint32_t j = 0; j < GM_NHALO; j++) {
            d.L.hoff[q * GM_NHALO + j] = run;
            run += cnt[j];
        }
        d.L.win[2u * q] = dI;
        d.L.win[2u * q + 1u] = run;
    }
    __syncthreads();
    uint32_t wI = 0, wB = 0;
    for (uint32_t l = threadIdx.x; l < d.nlists; l += 1024) {
        const uint32_t n = d.L.n[l], w = d.L.win[l], rl = d.L.rlen[l];
        d.L.nold[l] = n;
        d.L.n[l] = w > rl ? n + (w - rl) : n;
        if (l & 1u) wB += w;
        else wI += w;
    }
    if (wI) atomicAdd(&acc[0], wI);
    if (wB) atomicAdd(&acc[1], wB);
    __syncthreads();
    if (threadIdx.x == 0) {
        st->winI = acc[0];
        st->winB = acc[1];
    }
    rg_prefix_block(d);
}

/* New records and retiring parents as the one-region commit pass writes them;
 * a winner's slot is its rank in its list: own Inside and own Border winners
 * by the parents' prefix counts less the count at the region's first parent,
 * halo winners after the receiving region's own, by neighbour, then by the
 * infecting region's prefix count of that halo.  A halo winner's weight is
 * looked up again whatever its deformed grain (SCORE_Kernel.cpp:4853). */
__global__ void __launch_bounds__(64) k_rg_commit(GmDev d, GmPParams pp, GmSweepArgs A) {
    __shared__ uint32_t sI[64], sB[64], oI[64], oB[64];
    __shared__ uint32_t sH[GM_NHALO][64], oH[GM_NHALO][64];
    __shared__ uint4 sq0[64], sq2[64];
    __shared__ double sfin[64], sP[64];
    __shared__ uint16_t wl[26 * 64];
    GmStats *st = d.st;
    if (st->halt) return;
    const uint32_t k = st->kcur;
    const uint32_t lane = threadIdx.x;
    __shared__ uint32_t srs[GM_RROWS], sL[2][5];
    uint32_t lo, hi, reg;
    rg_part(d, lo, hi, reg);
    if (lo >= hi) return; /* no parents: nothing to commit */
    const RgBox box = rg_box(d, reg);
    /* the block's region: its counts at its first parent, and its two lists' tables */
    if (lane < GM_RROWS) srs[lane] = d.L.rs[reg * GM_RROWS + lane];
    if (lane < 2u) {
        const uint32_t l = 2u * reg + lane;
        sL[lane][0] = d.L.rlen[l];
        sL[lane][1] = d.L.rbase[l];
        sL[lane][2] = d.L.nold[l];
        sL[lane][3] = d.L.cap[l];
        sL[lane][4] = d.L.base[l];
    }
    uint32_t run[GM_RROWS];
#pragma unroll
    for (int c = 0; c < GM_RROWS; c++) run[c] = d.bcnt[(RW_ROW0 + c) * GM_MAX_BLOCKS + blockIdx.x];
    for (uint32_t batch = lo; batch < hi; batch += 64) {
        const uint32_t pt = batch + lane;
        const bool valid = pt < hi;
        const uint32_t mI = valid ? d.pmI[pt] : 0u, mB = valid ? d.pmB[pt] : 0u, mH = valid ? d.pmH[pt] : 0u;
        uint4 q0 = make_uint4(0u, 0u, 0xFFFFFFFFu, 0xFFFFFFFFu), q2 = make_uint4(0u, 0u, 0u, 0u);
        double2 fp = make_double2(0.0, 0.0);
        if (valid) {
            q0 = par_plane(d, GM_PAR_IDS)[pt];
            fp = reinterpret_cast<const double2 *>(par_plane(d, GM_PAR_FP))[pt];
            q2 = par_plane(d, GM_PAR_XYZ)[pt];
        }
        uint32_t mh[GM_NHALO];
        halo_masks(d, box, q2, mH, mh);
        /* exclusive prefixes within the batch, two 16-bit counts per scan
         * (a batch holds at most 64*26 winners of a row) */
        uint32_t cnt[GM_RROWS], ex[GM_RROWS], bt[GM_RROWS];
        cnt[GM_RW_I] = (uint32_t)__popc(mI);
        cnt[GM_RW_B] = (uint32_t)__popc(mB);
        cnt[GM_RW_A] = 0u;
#pragma unroll
        for (uint32_t j = 0; j < GM_NHALO; j++) cnt[GM_RW_H0 + j] = (uint32_t)__popc(mh[j]);
#pragma unroll
        for (int c = 0; c < GM_RROWS; c += 2) {
            const uint32_t hi16 = c + 1 < GM_RROWS ? cnt[c + 1] : 0u;
            const uint32_t packed = (cnt[c] << 16) | hi16;
            const uint32_t incl = wave_incl_scan(packed);
            const uint32_t e = incl - packed, t = __shfl(incl, 63, 64);
            ex[c] = e >> 16;
            bt[c] = t >> 16;
            if (c + 1 < GM_RROWS) {
                ex[c + 1] = e & 0xFFFFu;
                bt[c + 1] = t & 0xFFFFu;
            }
        }
        oI[lane] = run[GM_RW_I] + ex[GM_RW_I];
        oB[lane] = run[GM_RW_B] + ex[GM_RW_B];
#pragma unroll
        for (uint32_t j = 0; j < GM_NHALO; j++) {
            oH[j][lane] = run[GM_RW_H0 + j] + ex[GM_RW_H0 + j];
            sH[j][lane] = mh[j];
        }
        sI[lane] = mI;
        sB[lane] = mB;
        sq0[lane] = q0;
        sq2[lane] = q2;
        sfin[lane] = fp.x;
        sP[lane] = fp.y;
        /* this parent's winners at their place in the flat list, by rank */
        const uint32_t mall = mI | mB | mH, call = (uint32_t)__popc(mall);
        const uint32_t inc = wave_incl_scan(call);
        const uint32_t nw = __shfl(inc, 63, 64);
        uint32_t pos = inc - call;
        for (uint32_t m = mall; m; m &= m - 1u) wl[pos++] = (uint16_t)(lane | ((uint32_t)__ffs(m) - 1u) << 8);
        __syncthreads();
        for (uint32_t w = lane; w < nw; w += 64) {
            const uint32_t e = wl[w], j = e & 0xFFu, r = e >> 8, bit = 1u << r;
            const uint4 pq0 = sq0[j], pq2 = sq2[j];
            RgTarget o;
            rg_target(d, box, pq2.w, pq2.x, pq2.y, pq2.z, (int)r, o);
            const uint32_t t = cell_of(d, o.x, o.y, o.z);
            const uint32_t g = d.grid[t];
            /* the rank in the destination list, and that list's tables: the block's
             * own two lists' from LDS, a halo's receiving list's from memory */
            uint32_t rank, ix;
            if (o.khi) {
                const uint32_t q = o.dl >> 1, nbi = (o.khi >> 28) & 7u;
                rank = d.L.hoff[q * GM_NHALO + nbi] + (oH[o.hidx][j] - srs[GM_RW_H0 + o.hidx]) +
                       (uint32_t)__popc(sH[o.hidx][j] & (bit - 1u));
                ix = 2u;
            } else if (o.dl & 1u) {
                rank = oB[j] - srs[GM_RW_B] + (uint32_t)__popc(sB[j] & (bit - 1u));
                ix = 1u;
            } else {
                rank = oI[j] - srs[GM_RW_I] + (uint32_t)__popc(sI[j] & (bit - 1u));
                ix = 0u;
            }
            const uint32_t l = o.dl;
            uint32_t rl, rb, no, cp, bs;
            if (ix < 2u) {
                rl = sL[ix][0]; rb = sL[ix][1]; no = sL[ix][2]; cp = sL[ix][3]; bs = sL[ix][4];
            } else {
                rl = d.L.rlen[l]; rb = d.L.rbase[l]; no = d.L.nold[l]; cp = d.L.cap[l]; bs = d.L.base[l];
            }
            const uint32_t rel = rank < rl ? d.L.rbuf[rb + rank] : no + (rank - rl);
            if (rel >= cp) {
                atomicOr(&st->err, (l & 1u) ? GM_ERRF_BORDER_FULL : GM_ERRF_INSIDE_FULL);
                continue;
            }
            infect_one(d, pp, A, k, sfin[j], sP[j], pq0, r, t, g, o.x, o.y, o.z, bs + rel, o.khi != 0u);
        }
        retire_batch(d, q0, valid);
#pragma unroll
        for (int c = 0; c < GM_RROWS; c++) run[c] += bt[c];
        __syncthreads();
    }
}

hipError_t gm_launch_rg_infect(const GmDev &d, const GmPParams &pp, const GmSweepArgs &a, uint32_t parents_upper,
                               hipStream_t s, hipEvent_t *ev) {
    hipError_t e;
    if ((e = mark(ev, 4, s)) != hipSuccess) return e;
    k_rg_propose<<<grid_for(parents_upper, GM_BLOCK / GM_PG), GM_BLOCK, 0, s>>>(d);
    const uint32_t nb = rg_infect_blocks(d, parents_upper);
    if ((e = mark(ev, 5, s)) != hipSuccess) return e;
    k_rg_decide<<<nb, 64, 0, s>>>(d);
    if ((e = mark(ev, 6, s)) != hipSuccess) return e;
    k_rg_decide_scan<<<1, 1024, 0, s>>>(d, nb);
    if ((e = mark(ev, 7, s)) != hipSuccess) return e;
    k_rg_recycle<<<rg_sweep_blocks(d, a), 64, 0, s>>>(d);
    if ((e = mark(ev, 8, s)) != hipSuccess) return e;
    GmSweepArgs A = a;
    A.reanchor = 0;
    k_rg_commit<<<nb, 64, 0, s>>>(d, pp, A);
    if ((e = mark(ev, 9, s)) != hipSuccess) return e;
    return hipGetLastError();
}
