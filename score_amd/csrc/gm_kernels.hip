/*
 * gm_kernels.hip -- gfx950 kernels of the growth machine.
 *
 * One growth step of SCORE (SCORE_Kernel.cpp:6635-6654 with OMP_NUM_THREADS=1)
 * is, in serial order: sweep the Inside list, sweep the Border list, let every
 * completed Inside cell infect its 26 neighbours in list order, then every
 * completed Border cell.  A target is won by the first attempt in that order
 * that finds it deformed.  The same outcome is computed by nine kernels:
 *
 *   k_reanchor     only in steps whose scalars changed: every active cell is
 *                  advanced to the step with its old increment and predicted
 *                  again with the new one, one thread per slot.
 *   k_sweep        pass over the occupied slots of both lists, statically
 *                  partitioned over the grid.  A slot costs one 4-byte read (its
 *                  predicted completion step).  Completed and INACTIVE slots are
 *                  balloted into per-tile masks and counted per block.
 *   k_sweep_scan   one block scans the per-block counts into offsets.
 *   k_compact      one wave per sweep block writes the completed slots in FullRX
 *                  order (Inside list, then Border) at those offsets.
 *   k_propose      copies what the infection needs of each completed cell (ids,
 *                  rxFrac, P, coordinates) into a dense parent array; then every
 *                  (parent p, direction rank r) attempt on a deformed cell
 *                  (one bit of a brick-blocked bitmap) does
 *                  atomicMin(claim[target], p*32 + r): the key orders attempts
 *                  exactly as the serial loop issues them.
 *   k_decide       an attempt wins iff its key is the minimum; winners per parent
 *                  and destination list as bit masks, counted per block; a
 *                  parent that won gets its rxFrac after the completing
 *                  addition, replayed from its anchor.
 *   k_decide_scan  one block scans those counts: rank k into a list takes the
 *                  k-th recycled slot, else appends.
 *   k_recycle      writes the recycled slots that will be handed out: the first
 *                  min(winners, cap) INACTIVE slots of each list, ascending.
 *   k_commit       one thread per winner writes the new record with its
 *                  increment and completion step (certified from the real
 *                  sums, gm_radd_count_to_one), and clears the cell's bitmap
 *                  bit; then the parents retire.
 *
 * No kernel waits on another block: every cross-block dependency is a kernel
 * boundary, so the passes use only relaxed device atomics.  FP64 throughout,
 * compiled with -ffp-contract=off so every product and sum is rounded exactly as
 * in the reference's build without FMA.
 */
#include "gm_device.h"

/* ---------------------------------------------------------------- helpers */

__global__ void k_fill_u32(uint32_t *p, uint32_t v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
}
__global__ void k_fill_u64(unsigned long long *p, unsigned long long v, size_t n) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = v;
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

/* The per-step scalars changed: every active cell is advanced to the start of
 * step k with its old increment (its anchor moves to k) and predicted again
 * with the new one.  One thread per occupied slot.  Codes of cells that
 * completed before step k are turned into INACTIVE. */
__global__ void __launch_bounds__(GM_BLOCK) k_reanchor(GmDev d, GmSweepArgs A0) {
    if (d.st->halt) return;
    const GmSweepArgs A = step_args(d, A0);
    const uint32_t nI = d.st->nI, n = nI + d.st->nB;
    const uint32_t kc = A.k & (GM_CS_MOD - 1u);
    for (uint32_t i = blockIdx.x * GM_BLOCK + threadIdx.x; i < n; i += gridDim.x * GM_BLOCK) {
        const uint32_t s = occ_slot(d, nI, i);
        const uint32_t cs = d.cstep[s];
        if (cs_inactive(cs, kc, d.horizon)) {
            /* a cell that completed in an earlier step: its code becomes the sentinel */
            if (cs != GM_CS_INACTIVE) d.cstep[s] = (gm_cs_t)GM_CS_INACTIVE;
            continue;
        }
        GmRec *r = d.rec + s;
        const RecA a = rec_a(r);
        const RecC c = rec_c(r);
        double f = a.frac0;
        if (cs != GM_CS_UNKNOWN) {
            const double inc = r->inc;
            /* one step since the anchor: the one addition itself */
            f = A.k - a.s0 == 1u ? f + inc : gm_radd_n(f, inc, A.k - a.s0);
        }
        const double P = resolved(d, c.P);
        if (is_pending(c.P)) r->P = P;
        predict_slot(d, A, s, f, A.k, P, c.defg, a.inf);
    }
}

/* list lengths of the current step as tiles: the Inside list rounded up to
 * whole sweep chunks (tI), then the Border list (T in all) */
static __device__ __forceinline__ void step_tiles(uint32_t nI, uint32_t nB, uint32_t &tI, uint32_t &T) {
    tI = gm_tiles_inside(nI);
    T = tI + (nB + 63) / 64;
}

/* per-block list counts: [0] full Inside, [1] recycled Inside, [2] full Border,
 * [3] recycled Border, [4] cells without a valid prediction; each GM_MAX_BLOCKS long */
#define BC_FI 0
#define BC_RI 1
#define BC_FB 2
#define BC_RB 3
#define BC_UN 4
/* rows of the infection pass: winners into Inside / Border, attempts */
#define BC_WI 5
#define BC_WB 6
#define BC_WA 7

static __device__ __forceinline__ uint32_t tile_slot0(const GmDev &d, uint32_t tile, uint32_t tI) {
    return tile < tI ? tile * 64u : d.capI + (tile - tI) * 64u;
}

/* pass 1: every occupied slot, eight per lane with one 16-byte load of 2-byte
 * codes: a wave reads a 512-slot chunk of GM_SW_TPC = 8 tiles per load, and
 * GM_SW_UNROLL chunks at once.  Each lane classifies its eight slots without
 * branches (completed bits 0-7, INACTIVE bits 8-15); tile t's masks are balloted
 * from lanes 8t..8t+7, lane t keeps them and one store writes the chunk's.  The
 * Inside list spans tI tiles (whole chunks), the Border list follows from tile tI. */
#ifndef GM_SW_UNROLL
#define GM_SW_UNROLL 2
#endif
static_assert(GM_SW_TPC == 8, "a 16-byte load of 2-byte codes covers 8 slots per lane");
__global__ void __launch_bounds__(GM_BLOCK) k_sweep(GmDev d, GmSweepArgs A0) {
    __shared__ uint32_t red[GM_BLOCK / 64][5];
    __shared__ uint64_t sF[GM_BLOCK / 64][8], sR[GM_BLOCK / 64][8];
    if (d.st->halt) return;
    const GmSweepArgs A = step_args(d, A0);
    const uint32_t nI = d.st->nI, nB = d.st->nB;
    uint32_t tI, T;
    step_tiles(nI, nB, tI, T);
    const uint32_t lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t lo, hi;
    part_of(T, GM_SW_TPC, lo, hi); /* whole chunks, except an end at T */
    /* chunks [c_lo, c_hi); a block left empty (lo == hi == T) has none */
    const uint32_t c_lo = lo / GM_SW_TPC, c_hi = lo < hi ? (hi + GM_SW_TPC - 1u) / GM_SW_TPC : c_lo;
    const uint32_t kc = A.k & (GM_CS_MOD - 1u), horizon = d.horizon;
    uint32_t cFI = 0, cRI = 0, cFB = 0, cRB = 0, unknown = 0;
    for (uint32_t c0 = c_lo + w; c0 < c_hi; c0 += (GM_BLOCK / 64) * GM_SW_UNROLL) {
        uint4 cs[GM_SW_UNROLL];
        uint32_t nocc[GM_SW_UNROLL];
#pragma unroll
        for (int u = 0; u < GM_SW_UNROLL; u++) {
            const uint32_t c = c0 + u * (GM_BLOCK / 64);
            const uint32_t tile0 = c * GM_SW_TPC;
            const uint32_t base = tile_slot0(d, tile0, tI) + 8u * lane;
            /* slots below lim are occupied and inside this block's tiles */
            const uint32_t lend = tile0 < tI ? nI : d.capI + nB;
            const uint32_t tend = tile_slot0(d, tile0, tI) + 64u * (hi - tile0 < GM_SW_TPC ? hi - tile0 : GM_SW_TPC);
            const uint32_t lim = c < c_hi ? (lend < tend ? lend : tend) : 0u;
            nocc[u] = lim > base ? (lim - base < 8u ? lim - base : 8u) : 0u; /* occupied slots of this lane */
            cs[u] = nocc[u] ? *reinterpret_cast<const uint4 *>(d.cstep + base)
                            : make_uint4(0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu, 0xFFFFFFFFu);
        }
#pragma unroll
        for (int u = 0; u < GM_SW_UNROLL; u++) {
            const uint32_t c = c0 + u * (GM_BLOCK / 64);
            if (c >= c_hi) break;
            const uint32_t pack = sweep_classify(cs[u], nocc[u], kc, horizon, unknown);
            const uint32_t tile0 = c * GM_SW_TPC;
            /* per-lane counts, summed over the wave once at the end */
            const uint32_t nf = (uint32_t)__popc(pack & 0xFFu), nr = (uint32_t)__popc(pack >> 8);
            if (tile0 >= tI) {
                cFB += nf;
                cRB += nr;
            } else {
                cFI += nf;
                cRI += nr;
            }
            uint64_t myF, myR;
            sweep_masks(pack, sF[w], sR[w], myF, myR);
            if (lane < GM_SW_TPC && tile0 + lane < hi) {
                d.mF[tile0 + lane] = myF;
                d.mR[tile0 + lane] = myR;
            }
        }
    }
    for (int o = 32; o > 0; o >>= 1) {
        unknown += __shfl_down(unknown, o, 64);
        cFI += __shfl_down(cFI, o, 64);
        cRI += __shfl_down(cRI, o, 64);
        cFB += __shfl_down(cFB, o, 64);
        cRB += __shfl_down(cRB, o, 64);
    }
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

/* pass 2: per-block offsets and the list lengths.  Cells the sweep met without
 * a prediction (UNKNOWN) raise GM_ERRF_MISSED. */
__global__ void __launch_bounds__(1024) k_sweep_scan(GmDev d, uint32_t nb) {
    if (d.st->halt) return;
    uint32_t tot[5];
    scan_block_counts<5>(d.bcnt, nb, tot);
    if (threadIdx.x == 0) {
        GmStats *st = d.st;
        if (tot[BC_UN]) st->err |= GM_ERRF_MISSED;
        /* infection keys are p*32 + rank in 32 bits, p < nfull */
        if ((uint64_t)tot[BC_FI] + tot[BC_FB] > (1ull << 27)) st->err |= GM_ERRF_KEYS;
        st->nfullI = tot[BC_FI];
        st->totRecI = tot[BC_RI];
        st->nfullB = tot[BC_FB];
        st->totRecB = tot[BC_RB];
        st->recLenI = tot[BC_RI] < d.recCapI ? tot[BC_RI] : d.recCapI;
        st->recLenB = tot[BC_RB] < d.recCapB ? tot[BC_RB] : d.recCapB;
    }
}

/* Running offsets of the tiles [batch, batch+64) of a sweep block: lane t
 * holds tile batch+t and gets, per list count c (BC_FI..BC_RB), the offset of
 * its tile's first entry; run[] advances past the batch. */
static __device__ __forceinline__ void tile_offsets(const GmDev &d, uint32_t tI, uint32_t t, bool valid, uint64_t &bf,
                                                    uint64_t &br, uint32_t *run, uint32_t *off) {
    bf = valid ? d.mF[t] : 0ull;
    br = valid ? d.mR[t] : 0ull;
    const bool isB = t >= tI;
    const uint32_t cf = (uint32_t)__popcll(bf), cr = (uint32_t)__popcll(br);
    const uint32_t v[4] = {isB ? 0u : cf, isB ? 0u : cr, isB ? cf : 0u, isB ? cr : 0u};
#pragma unroll
    for (int c = 0; c < 4; c++) {
        const uint32_t incl = wave_incl_scan(v[c]);
        off[c] = run[c] + incl - v[c];
        run[c] += __shfl(incl, 63, 64);
    }
}

/* pass 3: append_to_fullrx_*_list (SCORE_Kernel.cpp:3994-4042, :4133-4180) as
 * ordered compaction: the completed slots in ascending order, Inside then
 * Border, in one array.  One wave per sweep block, one lane per tile: the lane
 * writes its tile's completed slots at its offset. */
__global__ void __launch_bounds__(64) k_compact(GmDev d) {
    if (d.st->halt) return;
    const uint32_t lane = threadIdx.x;
    uint32_t tI, T;
    step_tiles(d.st->nI, d.st->nB, tI, T);
    uint32_t lo, hi;
    part_of(T, GM_SW_TPC, lo, hi);
    uint32_t run[4];
#pragma unroll
    for (int c = 0; c < 4; c++) run[c] = d.bcnt[c * GM_MAX_BLOCKS + blockIdx.x];
    run[BC_FB] += d.st->nfullI; /* Border parents follow the Inside ones */
    for (uint32_t batch = lo; batch < hi; batch += 64) {
        const uint32_t t = batch + lane;
        uint64_t bf, br;
        uint32_t off[4];
        tile_offsets(d, tI, t, t < hi, bf, br, run, off);
        const uint32_t s0 = tile_slot0(d, t, tI);
        uint32_t o = off[t >= tI ? BC_FB : BC_FI];
        for (uint64_t m = bf; m; m &= m - 1ull) d.pslot[o++] = s0 + (uint32_t)__ffsll((long long)m) - 1u;
    }
}

/* append_to_recycling_*_list (SCORE_Kernel.cpp:4037-4041, :4175-4179): the
 * recycled slots in ascending order, of which only the first min(winners, cap)
 * are ever handed out; written after the winners are counted, so each sweep
 * block writes only the part of its INACTIVE slots that falls below that. */
__global__ void __launch_bounds__(64) k_recycle(GmDev d) {
    const GmStats *st = d.st;
    if (st->halt) return;
    const uint32_t lane = threadIdx.x;
    /* the sweep's tiles and partition: the lengths before this step's appends */
    uint32_t tI, T;
    step_tiles(st->nIold, st->nBold, tI, T);
    const uint32_t limI = st->winI < st->recLenI ? st->winI : st->recLenI;
    const uint32_t limB = st->winB < st->recLenB ? st->winB : st->recLenB;
    uint32_t run[4];
#pragma unroll
    for (int c = 0; c < 4; c++) run[c] = d.bcnt[c * GM_MAX_BLOCKS + blockIdx.x];
    uint32_t lo, hi;
    part_of(T, GM_SW_TPC, lo, hi);
    for (uint32_t batch = lo; batch < hi; batch += 64) {
        if (run[BC_RI] >= limI && run[BC_RB] >= limB) return; /* nothing further is handed out */
        const uint32_t t = batch + lane;
        uint64_t bf, br;
        uint32_t off[4];
        tile_offsets(d, tI, t, t < hi, bf, br, run, off);
        const uint32_t s0 = tile_slot0(d, t, tI);
        if (t >= tI) {
            uint32_t o = off[BC_RB];
            for (uint64_t m = br; m && o < limB; m &= m - 1ull) d.recB[o++] = s0 - d.capI + (uint32_t)__ffsll((long long)m) - 1u;
        } else {
            uint32_t o = off[BC_RI];
            for (uint64_t m = br; m && o < limI; m &= m - 1ull) d.recI[o++] = s0 + (uint32_t)__ffsll((long long)m) - 1u;
        }
    }
}

/* grid of the sweep, compaction and recycling passes for lists of about nI, nB */
#ifndef GM_SW_CHUNKS
#define GM_SW_CHUNKS 2 /* 512-slot chunks per sweep wave */
#endif
static uint32_t sweep_blocks(uint32_t nI, uint32_t nB) {
    const uint32_t T = gm_tiles_inside(nI) + (nB + 63) / 64;
    return grid_for(T, GM_SW_TPC * (GM_BLOCK / 64) * GM_SW_CHUNKS);
}

hipError_t gm_launch_reanchor(const GmDev &d, const GmSweepArgs &a, hipStream_t s) {
    k_reanchor<<<grid_for(a.nI + a.nB, GM_BLOCK), GM_BLOCK, 0, s>>>(d, a);
    return hipGetLastError();
}

hipError_t gm_launch_sweep(const GmDev &d, const GmSweepArgs &a, hipStream_t s, hipEvent_t *ev) {
    hipError_t e;
    GmSweepArgs A = a;
    A.reanchor = 0;
    if ((e = mark(ev, 0, s)) != hipSuccess) return e;
    if (a.reanchor) k_reanchor<<<grid_for(a.nI + a.nB, GM_BLOCK), GM_BLOCK, 0, s>>>(d, a);
    if ((e = gm_launch_arrivals(d, A, s)) != hipSuccess) return e;
    const uint32_t nb = sweep_blocks(a.nI, a.nB);
    if ((e = mark(ev, 1, s)) != hipSuccess) return e;
    k_sweep<<<nb, GM_BLOCK, 0, s>>>(d, A);
    if ((e = mark(ev, 2, s)) != hipSuccess) return e;
    k_sweep_scan<<<1, 1024, 0, s>>>(d, nb);
    if ((e = mark(ev, 3, s)) != hipSuccess) return e;
    k_compact<<<nb, 64, 0, s>>>(d);
    return hipGetLastError();
}

/* ----------------------------------------------------------------- infection */

/* every word of the bitmap from the cell grid (cells outside the domain stay 0) */
__global__ void k_dmap_init(GmDev d, uint32_t nwords) {
    for (uint32_t w = blockIdx.x * blockDim.x + threadIdx.x; w < nwords; w += gridDim.x * blockDim.x) {
        const uint32_t line = w >> 5, sub = w & 31u;
        const uint32_t bx = line % d.dnbx, rest = line / d.dnbx, by = rest % d.dnby, bz = rest / d.dnby;
        uint32_t v = 0;
        for (uint32_t i = 0; i < 32; i++) {
            const uint32_t b = sub * 32u + i;
            const uint32_t x = bx * 16u + (b & 15u), y = by * 8u + ((b >> 4) & 7u), z = bz * 8u + (b >> 7);
            if (x < d.nx && y < d.ny && z < d.nz && d.grid[x + y * d.nx + z * d.nxy] < d.ndefg) v |= 1u << i;
        }
        d.dmap[w] = v;
    }
}

hipError_t gm_launch_dmap_init(const GmDev &d, hipStream_t s) {
    const uint32_t nwords = d.dnbx * d.dnby * d.dnbz * 32u;
    k_dmap_init<<<fill_blocks(nwords), GM_BLOCK, 0, s>>>(d, nwords);
    return hipGetLastError();
}

/* Target of attempt rank r from a parent at (x,y,z).  Inside parents
 * (omp_infect_OutOfRXFrontInside, SCORE_Kernel.cpp:4252-4344) never wrap; they send
 * every target to the Border list unless they lie more than one cell away from
 * every face (:4606-4609).  Border parents (:4347-4493) wrap x and y, drop
 * targets outside z, and send strictly interior targets to the Inside list. */
static __device__ __forceinline__ bool attempt_target(const GmDev &d, bool fromB, uint32_t x, uint32_t y, uint32_t z,
                                                      int r, uint32_t &tx_, uint32_t &ty_, uint32_t &tz_, bool &toB) {
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
    tx_ = (uint32_t)tx;
    ty_ = (uint32_t)ty;
    tz_ = (uint32_t)tz;
    return true;
}


/* GM_PG lanes per parent: lane 0 copies what the infection needs of the
 * completed cell from its record into the dense parent array (FullRX order);
 * lane l makes the attempts of rank l, l+GM_PG, ... (< 26), so the bitmap reads
 * of a lane are independent and a wave keeps 64/GM_PG parents in flight.  Every
 * attempt on a deformed cell claims its target (the smallest key p*32+r is the
 * serial winner); their mask over the ranks is kept in pmI for the decide pass. */
__global__ void __launch_bounds__(GM_BLOCK) k_propose(GmDev d) {
    if (d.st->halt) return;
    const uint32_t nfull = d.st->nfullI + d.st->nfullB;
    /* the step's parents retire into the log at [nret, nret + nfull) */
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
            q2 = make_uint4(x, y, z, s >= d.capI ? 1u : 0u);
            par_plane(d, GM_PAR_IDS)[p] = make_uint4(s, cell, c.defg, rxg);
            /* the value after the completing addition follows in k_decide */
            reinterpret_cast<double2 *>(par_plane(d, GM_PAR_FP))[p] = make_double2(0.0, c.P);
            par_plane(d, GM_PAR_XYZ)[p] = q2;
            if (logged) d.rlog[nret + p] = make_uint2(cell, d.ndefg + rxg);
        }
        q2.x = (uint32_t)__shfl((int)q2.x, 0, GM_PG);
        q2.y = (uint32_t)__shfl((int)q2.y, 0, GM_PG);
        q2.z = (uint32_t)__shfl((int)q2.z, 0, GM_PG);
        q2.w = (uint32_t)__shfl((int)q2.w, 0, GM_PG);
        const bool fromB = q2.w != 0u;
        const uint32_t gsh = (threadIdx.x & 63) & ~(GM_PG - 1); /* first lane of the group in the wave */
        uint32_t am = 0;
#pragma unroll
        for (uint32_t i = 0; i < GM_PG_ITERS; i++) {
            const uint32_t rk = lg + GM_PG * i;
            bool att = false;
            if (rk < 26) {
                uint32_t tx, ty, tz;
                bool toB;
                if (attempt_target(d, fromB, q2.x, q2.y, q2.z, (int)rk, tx, ty, tz, toB) && is_deformed(d, tx, ty, tz)) {
                    att = true;
                    atomicMin(&d.claim[cell_of(d, tx, ty, tz)], p * 32u + rk);
                }
            }
            am |= (uint32_t)((__ballot(att) >> gsh) & ((1ull << GM_PG) - 1u)) << (GM_PG * i);
        }
        if (lg == 0) d.pmI[p] = am;
    }
}

/* One lane per parent: which of its attempts won, as 26-bit masks per
 * destination list; per-block totals for the ranking scan and the attempts.
 * A parent that won a target gets its value after the completing addition, the
 * carry's source, in the parent array.  The partition into blocks is the commit
 * pass's. */
__global__ void __launch_bounds__(64) k_decide(GmDev d) {
    if (d.st->halt) return;
    const uint32_t nfull = d.st->nfullI + d.st->nfullB;
    const uint32_t k = d.st->kcur;
    uint32_t lo, hi;
    part_of(nfull, GM_PCHUNK, lo, hi);
    uint32_t cI = 0, cB = 0, cA = 0;
    for (uint32_t p = lo + threadIdx.x; p < hi; p += 64) {
        const uint4 q2 = par_plane(d, GM_PAR_XYZ)[p];
        const uint32_t am = d.pmI[p];
        const uint32_t ps = d.pslot[p]; /* loaded with the rest, used if the parent wins */
        uint32_t mI = 0, mB = 0;
        for (uint32_t m = am; m; m &= m - 1u) {
            const uint32_t rk = (uint32_t)__ffs(m) - 1u;
            uint32_t tx, ty, tz;
            bool toB;
            attempt_target(d, q2.w != 0u, q2.x, q2.y, q2.z, (int)rk, tx, ty, tz, toB);
            if (d.claim[cell_of(d, tx, ty, tz)] == p * 32u + rk) {
                if (toB) mB |= 1u << rk;
                else mI |= 1u << rk;
            }
        }
        d.pmI[p] = mI;
        d.pmB[p] = mB;
        if (mI | mB) reinterpret_cast<double *>(par_plane(d, GM_PAR_FP))[2 * (size_t)p] = parent_fin(d, ps, k);
        cI += (uint32_t)__popc(mI);
        cB += (uint32_t)__popc(mB);
        cA += (uint32_t)__popc(am);
    }
    for (int o = 32; o > 0; o >>= 1) {
        cI += __shfl_down(cI, o, 64);
        cB += __shfl_down(cB, o, 64);
        cA += __shfl_down(cA, o, 64);
    }
    if (threadIdx.x == 0) {
        d.bcnt[BC_WI * GM_MAX_BLOCKS + blockIdx.x] = cI;
        d.bcnt[BC_WB * GM_MAX_BLOCKS + blockIdx.x] = cB;
        d.bcnt[BC_WA * GM_MAX_BLOCKS + blockIdx.x] = cA;
    }
}

/* one block: offsets of every block's winners in each list, the new list lengths
 * (recycled slots first, then appended) */
__global__ void __launch_bounds__(1024) k_decide_scan(GmDev d, uint32_t nb) {
    if (d.st->halt) return;
    uint32_t tot[3];
    scan_block_counts<3>(d.bcnt + BC_WI * GM_MAX_BLOCKS, nb, tot);
    if (threadIdx.x == 0) {
        GmStats *st = d.st;
        st->winI = tot[0];
        st->winB = tot[1];
        st->nattempt = tot[2];
        /* the commit pass reads the lengths before this step's appends */
        st->nIold = st->nI;
        st->nBold = st->nB;
        if (tot[0] > st->recLenI) st->nI += tot[0] - st->recLenI;
        if (tot[1] > st->recLenB) st->nB += tot[1] - st->recLenB;
    }
}

/* New records as omp_infect_OutOfRXFront* write them (SCORE_Kernel.cpp:4281-4310,
 * :4402-4426), slots as omp_getNextFreeSlotIn* hand them out (:3757-3874); then
 * the parents retire as updateFullRX* does it (:4640-4653).  One wave per block
 * takes the block's parents (the partition of k_decide) 64 at a time; their
 * winners are flattened into a list so that every lane commits one infection.
 * Each lane loads its own parent's data with its winner masks and shares it
 * through LDS, and a winner issues its target's grid load and its slot lookup
 * together, so a wave waits on few dependent memory round trips. */
__global__ void __launch_bounds__(64) k_commit(GmDev d, GmPParams pp, GmSweepArgs A) {
    __shared__ uint32_t sI[64], sB[64], oI[64], oB[64];
    __shared__ uint4 sq0[64], sq2[64];
    __shared__ double sfin[64], sP[64];
    __shared__ uint16_t wl[26 * 64];
    GmStats *st = d.st;
    if (st->halt) return;
    const uint32_t k = st->kcur;
    const uint32_t nfull = st->nfullI + st->nfullB;
    const uint32_t recLenI = st->recLenI, recLenB = st->recLenB;
    const uint32_t nI = st->nIold, nB = st->nBold;
    const uint32_t lane = threadIdx.x;
    uint32_t lo, hi;
    part_of(nfull, GM_PCHUNK, lo, hi);
    uint32_t runI = d.bcnt[BC_WI * GM_MAX_BLOCKS + blockIdx.x], runB = d.bcnt[BC_WB * GM_MAX_BLOCKS + blockIdx.x];
    for (uint32_t batch = lo; batch < hi; batch += 64) {
        const uint32_t pt = batch + lane;
        const bool valid = pt < hi;
        const uint32_t mI = valid ? d.pmI[pt] : 0u, mB = valid ? d.pmB[pt] : 0u;
        /* this lane's parent: slot, cell, defg, rxg | fin, P | x, y, z, fromB */
        uint4 q0 = make_uint4(0u, 0u, 0xFFFFFFFFu, 0xFFFFFFFFu), q2 = make_uint4(0u, 0u, 0u, 0u);
        double2 fp = make_double2(0.0, 0.0);
        if (valid) {
            q0 = par_plane(d, GM_PAR_IDS)[pt];
            fp = reinterpret_cast<const double2 *>(par_plane(d, GM_PAR_FP))[pt];
            q2 = par_plane(d, GM_PAR_XYZ)[pt];
        }
        const uint32_t cI = (uint32_t)__popc(mI), cB = (uint32_t)__popc(mB);
        /* one scan of both counts (each batch total < 2^16) */
        const uint32_t packed = (cI << 16) | cB;
        const uint32_t incl = wave_incl_scan(packed);
        const uint32_t ex = incl - packed, tot = __shfl(incl, 63, 64);
        const uint32_t totI = tot >> 16, totB = tot & 0xFFFFu;
        oI[lane] = runI + (ex >> 16);
        oB[lane] = runB + (ex & 0xFFFFu);
        sI[lane] = mI;
        sB[lane] = mB;
        sq0[lane] = q0;
        sq2[lane] = q2;
        sfin[lane] = fp.x;
        sP[lane] = fp.y;
        /* winners of this parent at [exI + exB, ...) in the flat list, by rank */
        uint32_t pos = (ex >> 16) + (ex & 0xFFFFu);
        for (uint32_t m = mI | mB; m; m &= m - 1u) wl[pos++] = (uint16_t)(lane | ((uint32_t)__ffs(m) - 1u) << 8);
        __syncthreads();
        const uint32_t nw = totI + totB;
        for (uint32_t w = lane; w < nw; w += 64) {
            const uint32_t e = wl[w], j = e & 0xFFu, r = e >> 8, bit = 1u << r;
            const uint4 pq0 = sq0[j], pq2 = sq2[j];
            uint32_t tx, ty, tz;
            bool toB;
            attempt_target(d, pq2.w != 0u, pq2.x, pq2.y, pq2.z, (int)r, tx, ty, tz, toB);
            const uint32_t t = cell_of(d, tx, ty, tz);
            const uint32_t g = d.grid[t];
            /* rank k into a list takes its k-th recycled slot, else appends */
            const uint32_t wI = sI[j];
            const bool intoI = (wI & bit) != 0u;
            const uint32_t wm = intoI ? wI : sB[j];
            const uint32_t rank = (intoI ? oI[j] : oB[j]) + (uint32_t)__popc(wm & (bit - 1u));
            const uint32_t recLen = intoI ? recLenI : recLenB;
            const uint32_t *recl = intoI ? d.recI : d.recB;
            const uint32_t rel = rank < recLen ? recl[rank] : (intoI ? nI : nB) + (rank - recLen);
            if (rel >= (intoI ? d.capI : d.capB)) {
                atomicOr(&st->err, intoI ? GM_ERRF_INSIDE_FULL : GM_ERRF_BORDER_FULL);
                continue;
            }
            const uint32_t gslot = intoI ? rel : d.capI + rel;
            infect_one(d, pp, A, k, sfin[j], sP[j], pq0, r, t, g, tx, ty, tz, gslot);
        }
        retire_batch(d, q0, valid);
        runI += totI;
        runB += totB;
        __syncthreads();
    }
}

hipError_t gm_launch_infect(const GmDev &d, const GmPParams &pp, const GmSweepArgs &a, uint32_t parents_upper,
                            hipStream_t s, hipEvent_t *ev) {
    hipError_t e;
    if ((e = mark(ev, 4, s)) != hipSuccess) return e;
    k_propose<<<grid_for(parents_upper, GM_BLOCK / GM_PG), GM_BLOCK, 0, s>>>(d);
    const uint32_t nb = grid_for(parents_upper, GM_PCHUNK);
    if ((e = mark(ev, 5, s)) != hipSuccess) return e;
    k_decide<<<nb, 64, 0, s>>>(d);
    if ((e = mark(ev, 6, s)) != hipSuccess) return e;
    k_decide_scan<<<1, 1024, 0, s>>>(d, nb);
    /* the sweep's grid, for the recycling lists */
    if ((e = mark(ev, 7, s)) != hipSuccess) return e;
    k_recycle<<<sweep_blocks(a.nI, a.nB), 64, 0, s>>>(d);
    if ((e = mark(ev, 8, s)) != hipSuccess) return e;
    GmSweepArgs A = a;
    A.reanchor = 0;
    k_commit<<<nb, 64, 0, s>>>(d, pp, A);
    if ((e = mark(ev, 9, s)) != hipSuccess) return e;
    return hipGetLastError();
}

/* exported map: retired cells from the log, in retirement order (a cell retired
 * twice keeps the later entry, as SCORE's grid keeps its last write) */
__global__ void k_apply_rlog(GmDev d, uint32_t *out) {
    const unsigned long long n = d.st->nret < d.rlogCap ? d.st->nret : d.rlogCap;
    for (unsigned long long i = (unsigned long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (unsigned long long)gridDim.x * blockDim.x) {
        const uint2 e = d.rlog[i];
        out[e.x] = e.y;
    }
}

/* exported map: cells held by active front records before step k are INFECTED */
__global__ void k_overlay_infected(GmDev d, uint32_t nI, uint32_t nB, uint32_t k, uint32_t *out) {
    const uint32_t n = nI + nB, kc = k & (GM_CS_MOD - 1u);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t s = occ_slot(d, nI, i);
        if (!cs_inactive(d.cstep[s], kc, d.horizon)) out[d.rec[s].cell] = GM_INFECTED;
    }
}

hipError_t gm_launch_overlay_infected(const GmDev &d, uint32_t nI, uint32_t nB, uint32_t k, uint32_t *out,
                                      hipStream_t s) {
    k_apply_rlog<<<8192, GM_BLOCK, 0, s>>>(d, out);
    if (nI + nB == 0) return hipGetLastError();
    k_overlay_infected<<<fill_blocks(nI + nB), GM_BLOCK, 0, s>>>(d, nI, nB, k, out);
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

static __device__ __forceinline__ bool site_deformed(const GmDev &d, uint32_t w) {
    const uint32_t z = w / d.nxy, rem = w - z * d.nxy, y = rem / d.nx, x = rem - y * d.nx;
    return is_deformed(d, x, y, z);
}

__global__ void k_nuc_claim(GmDev d, const uint32_t *site, uint32_t n) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        uint32_t w = site[k];
        if (w != GM_NO_SITE && site_deformed(d, w)) atomicMin(&d.claim[w], k);
    }
}

static __device__ __forceinline__ bool site_interior(const GmDev &d, uint32_t w) {
    uint32_t z = w / d.nxy, rem = w - z * d.nxy, y = rem / d.nx, x = rem - y * d.nx;
    return x > 0 && x < d.nx - 1 && y > 0 && y < d.ny - 1 && z > 0 && z < d.nz - 1;
}

__global__ void k_nuc_flag(GmDev d, const uint32_t *site, uint32_t n) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
        uint32_t w = site[k];
        bool win = w != GM_NO_SITE && site_deformed(d, w) && d.claim[w] == k;
        bool in = win && site_interior(d, w);
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
        GmRec *r = d.rec + gslot;
        rec_set_a(r, 1.0, 0u, 26u);
        rec_set_b(r, 0.0);
        rec_set_c(r, nucP[k], w, g);
        r->rxg = k;
        d.cstep[gslot] = (gm_cs_t)GM_CS_UNKNOWN;
        const uint32_t z = w / d.nxy, rem = w - z * d.nxy, y = rem / d.nx, x = rem - y * d.nx;
        clear_deformed(d, x, y, z);
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

/* the stats into the host's pinned mirror (mapped, coherent), word by word:
 * cheaper than a copy command at the end of every window.  Only the filled part
 * of the inline id and key arrays crosses the bus; the host reads no more. */
__global__ void __launch_bounds__(GM_BLOCK) k_stats_out(GmDev d, uint32_t *out) {
    const uint32_t *in = reinterpret_cast<const uint32_t *>(d.st);
    const GmStats *st = d.st;
    const uint32_t nz = st->nzeroed < GM_ZEROED_INLINE ? st->nzeroed : GM_ZEROED_INLINE;
    const uint32_t np = st->npend < GM_PEND_INLINE ? st->npend : GM_PEND_INLINE;
    const uint32_t z0 = offsetof(GmStats, zeroed_ids) / 4, p0 = offsetof(GmStats, pend_keys) / 4;
    const uint32_t t0 = offsetof(GmStats, kcur) / 4, end = sizeof(GmStats) / 4;
    for (uint32_t i = threadIdx.x; i < end; i += GM_BLOCK) {
        const bool skip = (i >= z0 + nz && i < p0) || (i >= p0 + 2u * np && i < t0);
        if (!skip) out[i] = in[i];
    }
}

hipError_t gm_launch_stats_out(const GmDev &d, GmStats *out, hipStream_t s) {
    k_stats_out<<<1, GM_BLOCK, 0, s>>>(d, reinterpret_cast<uint32_t *>(out));
    return hipGetLastError();
}

/* SCORE's per-slot view before step k: activity, rxFrac (replayed from the
 * anchor; INACTIVE slots hold 0.0 as the reference leaves them), and the record
 * fields, with pending weights replaced by their cached values in the record
 * itself.  With o.act null only the weights are resolved. */
__global__ void k_materialize(GmDev d, uint32_t k, uint32_t base, uint32_t n, GmFrontOut o) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t s = base + i;
        const uint32_t cs = d.cstep[s];
        GmRec *r = d.rec + s;
        const RecA a = rec_a(r);
        const RecC c = rec_c(r);
        const double P = resolved(d, c.P);
        if (is_pending(c.P)) r->P = P;
        if (!o.act) continue;
        o.P[i] = P;
        o.inf[i] = (uint8_t)a.inf;
        o.cell[i] = c.cell;
        o.defg[i] = c.defg;
        o.rxg[i] = r->rxg;
        if (cs_inactive(cs, k & (GM_CS_MOD - 1u), d.horizon)) {
            o.act[i] = 0;
            o.frac[i] = 0.0;
            continue;
        }
        o.act[i] = 1;
        double f = a.frac0;
        if (cs != GM_CS_UNKNOWN) f = gm_radd_n(f, r->inc, k - a.s0);
        o.frac[i] = f;
    }
}

hipError_t gm_launch_materialize(const GmDev &d, uint32_t k, uint32_t base, uint32_t n, const GmFrontOut &o,
                                 hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_materialize<<<fill_blocks(n), GM_BLOCK, 0, s>>>(d, k, base, n, o);
    return hipGetLastError();
}

__global__ void k_hash_insert(GmDev d, const unsigned long long *keys, const double *vals, uint32_t n) {
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        unsigned long long key = keys[i];
        uint32_t h = hash_slot(key, d.hcap);
        const uint32_t nprobe = d.hcap < GM_PROBE_MAX ? d.hcap : GM_PROBE_MAX;
        bool stored = false;
        for (uint32_t probe = 0; probe < nprobe; probe++) {
            unsigned long long old = atomicCAS(&d.hkeys[h], 0ull, key);
            if (old == 0ull || old == key) {
                d.hvals[h] = vals[i];
                stored = true;
                break;
            }
            h = (h + 1u) & (d.hcap - 1u);
        }
        /* lookups stop after nprobe slots, so a key stored further away would be
         * lost; the host grows the table and inserts again */
        if (!stored) atomicOr(&d.st->err, GM_ERRF_PCACHE_FULL);
    }
}

hipError_t gm_launch_hash_insert(const GmDev &d, const unsigned long long *keys, const double *vals, uint32_t n,
                                 hipStream_t s) {
    if (n == 0) return hipSuccess;
    k_hash_insert<<<fill_blocks(n), GM_BLOCK, 0, s>>>(d, keys, vals, n);
    return hipGetLastError();
}

/* ------------------------------------------------------- defragmentation
 * ompshar_defragmentation (SCORE_Kernel.cpp:3877-3934), one list at a time: the
 * first INACTIVE slot is found (past nextSlotNeverActive if every slot below it
 * is ACTIVE, as the reference's unbounded search does), then every ACTIVE slot
 * after it moves down, in order, to close the gaps.  The reference copies the
 * records and leaves the slots behind the compacted block as they were; an
 * ACTIVE record left there is frozen here at its value before step k (anchor k,
 * increment 0), which is what the reference holds in that slot from then on. */

__global__ void __launch_bounds__(GM_BLOCK) k_defrag_first(GmDev d, uint32_t base, uint32_t lo, uint32_t hi,
                                                          uint32_t kc) {
    const uint32_t lane = threadIdx.x & 63u, stride = gridDim.x * GM_BLOCK;
    for (uint32_t i0 = lo + blockIdx.x * GM_BLOCK + (threadIdx.x & ~63u); i0 < hi; i0 += stride) {
        const uint32_t i = i0 + lane;
        const uint64_t m = __ballot(i < hi && cs_inactive(d.cstep[base + i], kc, d.horizon));
        if (m) {
            /* later chunks of this wave only hold larger slots */
            if (lane == 0) atomicMin(&d.st->dfirst, i0 + (uint32_t)__ffsll((long long)m) - 1u);
            return;
        }
        if (hi - i0 <= stride) return;
    }
}

hipError_t gm_launch_defrag_first_inactive(const GmDev &d, uint32_t base, uint32_t lo, uint32_t hi, uint32_t k,
                                           hipStream_t s) {
    k_fill_u32<<<1, 64, 0, s>>>(&d.st->dfirst, hi, 1);
    if (hi > lo) k_defrag_first<<<fill_blocks(hi - lo), GM_BLOCK, 0, s>>>(d, base, lo, hi, k & (GM_CS_MOD - 1u));
    return hipGetLastError();
}

/* ACTIVE slots among list slots first+1+i, i in [0, n), per block chunk */
__global__ void __launch_bounds__(GM_BLOCK) k_defrag_count(GmDev d, uint32_t base, uint32_t first, uint32_t n,
                                                          uint32_t kc) {
    uint32_t lo, hi;
    chunk_of(n, lo, hi);
    uint32_t c = 0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += GM_BLOCK)
        c += cs_inactive(d.cstep[base + first + 1u + i], kc, d.horizon) ? 0u : 1u;
    uint32_t tot;
    block_incl_scan(c, &tot);
    if (threadIdx.x == 0) d.bcnt[blockIdx.x] = tot;
}

/* exclusive scan of one row of per-block counts; the total goes to st->dcount */
__global__ void __launch_bounds__(1024) k_scan_row(GmDev d, uint32_t nb) {
    uint32_t tot[1];
    scan_block_counts<1>(d.bcnt, nb, tot);
    if (threadIdx.x == 0) d.st->dcount = tot[0];
}

/* Each ACTIVE record after `first` is copied, in slot order, to the scratch
 * array; one that the compacted block will not overwrite is then frozen. */
__global__ void __launch_bounds__(GM_BLOCK) k_defrag_gather(GmDev d, uint32_t base, uint32_t first, uint32_t n,
                                                           uint32_t k, GmRec *trec, gm_cs_t *tcs) {
    uint32_t lo, hi;
    chunk_of(n, lo, hi);
    uint32_t run = d.bcnt[blockIdx.x];
    const uint32_t keep = first + d.st->dcount; /* list slots below this receive moved records */
    for (uint32_t b0 = lo; b0 < hi; b0 += GM_BLOCK) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t c = first + 1u + i;
        const uint32_t cs = i < hi ? d.cstep[base + c] : GM_CS_INACTIVE;
        const uint32_t act = cs_inactive(cs, k & (GM_CS_MOD - 1u), d.horizon) ? 0u : 1u;
        /* the code of a cell that completed in an earlier step becomes the
         * sentinel, so that no past code stays behind the compacted block */
        if (i < hi && !act && cs != GM_CS_INACTIVE) d.cstep[base + c] = (gm_cs_t)GM_CS_INACTIVE;
        uint32_t tot;
        const uint32_t incl = block_incl_scan(act, &tot);
        if (act) {
            const uint32_t o = run + incl - 1u;
            GmRec *r = d.rec + base + c;
            const uint4 *src = reinterpret_cast<const uint4 *>(r);
            const uint4 q0 = src[0], q1 = src[1], q2 = src[2], q3 = src[3];
            uint4 *dst = reinterpret_cast<uint4 *>(trec + o);
            dst[0] = q0;
            dst[1] = q1;
            dst[2] = q2;
            dst[3] = q3;
            tcs[o] = (gm_cs_t)cs;
            if (c >= keep) {
                double f = __hiloint2double((int)q0.y, (int)q0.x);
                const double inc = __hiloint2double((int)q1.y, (int)q1.x);
                if (cs != GM_CS_UNKNOWN) f = gm_radd_n(f, inc, k - q0.z);
                const double P = resolved(d, __hiloint2double((int)q2.y, (int)q2.x));
                rec_set_a(r, f, k, q0.w);
                rec_set_b(r, 0.0);
                rec_set_c(r, P, q2.z, q2.w);
                d.cstep[base + c] = (gm_cs_t)GM_CS_NEVER;
            }
        }
        run += tot;
    }
}

__global__ void __launch_bounds__(GM_BLOCK) k_defrag_scatter(GmDev d, uint32_t base, uint32_t first,
                                                            const GmRec *trec, const gm_cs_t *tcs) {
    const uint32_t n = d.st->dcount;
    for (uint32_t i = blockIdx.x * GM_BLOCK + threadIdx.x; i < n; i += gridDim.x * GM_BLOCK) {
        const uint4 *src = reinterpret_cast<const uint4 *>(trec + i);
        uint4 *dst = reinterpret_cast<uint4 *>(d.rec + base + first + i);
        const uint4 q0 = src[0], q1 = src[1], q2 = src[2], q3 = src[3];
        dst[0] = q0;
        dst[1] = q1;
        dst[2] = q2;
        dst[3] = q3;
        d.cstep[base + first + i] = tcs[i];
    }
}

hipError_t gm_launch_defrag_move(const GmDev &d, uint32_t base, uint32_t first, uint32_t nnever, uint32_t k,
                                 GmRec *trec, gm_cs_t *tcs, hipStream_t s) {
    if (first + 1u >= nnever) return hipSuccess;
    const uint32_t n = nnever - first - 1u;
    const uint32_t nb = grid_for(n, 4 * GM_BLOCK);
    k_defrag_count<<<nb, GM_BLOCK, 0, s>>>(d, base, first, n, k & (GM_CS_MOD - 1u));
    k_scan_row<<<1, 1024, 0, s>>>(d, nb);
    k_defrag_gather<<<nb, GM_BLOCK, 0, s>>>(d, base, first, n, k, trec, tcs);
    k_defrag_scatter<<<fill_blocks(n), GM_BLOCK, 0, s>>>(d, base, first, trec, tcs);
    return hipGetLastError();
}

/* ------------------------------------------------------------ status vector
 * ompcrit_clarify_status_infectedcells (SCORE_Io_Characterization.cpp:925-997):
 * every ACTIVE record, Inside list then Border list in slot order, as its
 * linear cell index and the grain the cell shows before step k: ndefg+rxg once
 * rxFrac has reached the threshold, else its deformed grain.  Item i of
 * [0, nI+nB) is Inside slot i or Border slot i-nI. */

static __device__ __forceinline__ uint32_t item_slot(const GmDev &d, uint32_t nI, uint32_t i) {
    return occ_slot(d, nI, i);
}

__global__ void __launch_bounds__(GM_BLOCK) k_status_count(GmDev d, uint32_t nI, uint32_t n, uint32_t kc) {
    uint32_t lo, hi;
    chunk_of(n, lo, hi);
    uint32_t c = 0;
    for (uint32_t i = lo + threadIdx.x; i < hi; i += GM_BLOCK)
        c += cs_inactive(d.cstep[item_slot(d, nI, i)], kc, d.horizon) ? 0u : 1u;
    uint32_t tot;
    block_incl_scan(c, &tot);
    if (threadIdx.x == 0) d.bcnt[blockIdx.x] = tot;
}

__global__ void __launch_bounds__(GM_BLOCK) k_status_write(GmDev d, uint32_t nI, uint32_t n, uint32_t k, double thr,
                                                          uint32_t *ixyz, uint32_t *gid) {
    uint32_t lo, hi;
    chunk_of(n, lo, hi);
    uint32_t run = d.bcnt[blockIdx.x];
    for (uint32_t b0 = lo; b0 < hi; b0 += GM_BLOCK) {
        const uint32_t i = b0 + threadIdx.x;
        const uint32_t s = i < hi ? item_slot(d, nI, i) : 0u;
        const uint32_t cs = i < hi ? d.cstep[s] : GM_CS_INACTIVE;
        const uint32_t act = cs_inactive(cs, k & (GM_CS_MOD - 1u), d.horizon) ? 0u : 1u;
        uint32_t tot;
        const uint32_t incl = block_incl_scan(act, &tot);
        if (act) {
            const GmRec *r = d.rec + s;
            const RecA a = rec_a(r);
            const RecC c = rec_c(r);
            double f = a.frac0;
            if (cs != GM_CS_UNKNOWN) f = gm_radd_n(f, r->inc, k - a.s0);
            const uint32_t o = run + incl - 1u;
            ixyz[o] = c.cell;
            gid[o] = f < thr ? c.defg : d.ndefg + r->rxg;
        }
        run += tot;
    }
}

hipError_t gm_launch_status(const GmDev &d, uint32_t nI, uint32_t nB, uint32_t k, double thr, uint32_t *ixyz,
                            uint32_t *gid, hipStream_t s) {
    const uint32_t n = nI + nB;
    if (n == 0) {
        k_fill_u32<<<1, 64, 0, s>>>(&d.st->dcount, 0u, 1);
        return hipGetLastError();
    }
    const uint32_t nb = grid_for(n, 4 * GM_BLOCK);
    k_status_count<<<nb, GM_BLOCK, 0, s>>>(d, nI, n, k & (GM_CS_MOD - 1u));
    k_scan_row<<<1, 1024, 0, s>>>(d, nb);
    k_status_write<<<nb, GM_BLOCK, 0, s>>>(d, nI, n, k, thr, ixyz, gid);
    return hipGetLastError();
}

/* ------------------------------------------------------------- arrivals
 * Nuclei that arrive after step 0 are placed at the start of their step, before
 * the sweep, as infect_JuvenileNucleation places the site-saturated ones
 * (SCORE_GrowthMachine.cpp:124-189): in pool order, a nucleus whose cell is no
 * longer deformed (taken by the front or by an earlier nucleus) is rejected;
 * the record (rxFrac 1.0, infector 26, the host's weight) is appended to the
 * Inside list for a strictly interior cell, else to the Border list.  The
 * reference marks this injection point but has no model behind it
 * (SCORE_Kernel.cpp:6632).  One block: every nucleus of the step claims its
 * cell (the lowest pool id wins), then the winners are ranked in pool order,
 * 1024 at a time, and each record is predicted under the step's scalars. */
__global__ void __launch_bounds__(1024) k_arrivals(GmDev d, GmSweepArgs A0) {
    __shared__ uint32_t ws[16][2];
    GmStats *st = d.st;
    if (st->halt) return;
    const uint32_t k = st->kcur;
    if (k == 0u || k > d.arr_last) return;
    const GmSweepArgs A = step_args(d, A0);
    const uint32_t lo = d.arr_off[k], hi = d.arr_off[k + 1u];
    if (lo == hi) return;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    for (uint32_t i = lo + tid; i < hi; i += 1024u) {
        const uint32_t id = d.arr_ids[i], c = d.nsite[id];
        if (site_deformed(d, c)) atomicMin(&d.claim[c], id);
    }
    __syncthreads();
    uint32_t runI = st->nI, runB = st->nB, placed = 0;
    unsigned long long pk = 0ull;
    for (uint32_t base = lo; base < hi; base += 1024u) {
        const uint32_t i = base + tid;
        uint32_t id = 0, c = 0;
        bool win = false, in = false;
        if (i < hi) {
            id = d.arr_ids[i];
            c = d.nsite[id];
            win = site_deformed(d, c) &&
                  __hip_atomic_load(&d.claim[c], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == id;
            in = win && site_interior(d, c);
        }
        const uint32_t vI = (win && in) ? 1u : 0u, vB = (win && !in) ? 1u : 0u;
        const uint32_t packed = (vI << 16) | vB;
        const uint32_t incl = wave_incl_scan(packed);
        if (lane == 63u) {
            ws[w][0] = incl >> 16;
            ws[w][1] = incl & 0xFFFFu;
        }
        __syncthreads();
        uint32_t preI = 0, preB = 0, totI = 0, totB = 0;
        for (uint32_t q = 0; q < 16u; q++) {
            if (q < w) { preI += ws[q][0]; preB += ws[q][1]; }
            totI += ws[q][0];
            totB += ws[q][1];
        }
        if (win) {
            const uint32_t ex = incl - packed;
            const uint32_t rel = in ? runI + preI + (ex >> 16) : runB + preB + (ex & 0xFFFFu);
            if (rel >= (in ? d.capI : d.capB)) {
                atomicOr(&st->err, in ? GM_ERRF_INSIDE_FULL : GM_ERRF_BORDER_FULL);
            } else {
                const uint32_t gslot = in ? rel : d.capI + rel;
                const uint32_t g = d.grid[c];
                const double P = d.nucP[id];
                double inc = 0.0;
                const gm_cs_t cs = cs_code(d, predict_values(d, A, k, 1.0, P, g, 26u, inc), k);
                GmRec *r = d.rec + gslot;
                rec_set_a(r, 1.0, k, 26u);
                rec_set_b(r, inc);
                rec_set_c(r, P, c, g);
                *reinterpret_cast<uint4 *>(&r->rxg) = make_uint4(id, 0u, 0u, 0u);
                d.cstep[gslot] = cs;
                const uint32_t z = c / d.nxy, rem = c - z * d.nxy, y = rem / d.nx, x = rem - y * d.nx;
                clear_deformed(d, x, y, z);
                const unsigned long long key = gm_pkey(P);
                pk = key > pk ? key : pk;
            }
        }
        runI += totI;
        runB += totB;
        placed += totI + totB;
        __syncthreads();
    }
    /* the block's largest weight */
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long other = (unsigned long long)__shfl_down((long long)pk, o, 64);
        pk = other > pk ? other : pk;
    }
    if (lane == 0u && pk) atomicMax(&st->arr_pkey, pk);
    if (tid == 0) {
        st->nI = runI < d.capI ? runI : d.capI;
        st->nB = runB < d.capB ? runB : d.capB;
        st->nplaced = placed;
    }
}

hipError_t gm_launch_arrivals(const GmDev &d, const GmSweepArgs &a, hipStream_t s) {
    if (d.arr_last == 0u) return hipSuccess;
    k_arrivals<<<1, 1024, 0, s>>>(d, a);
    return hipGetLastError();
}

/* ----------------------------------------------------------- launch windows
 * Up to GM_WMAX steps are launched back to back.  Every kernel of a step reads
 * the step index and list lengths from the stats and returns at once once a
 * step has set `halt`, so that the host takes over at the step after it. */

__global__ void k_window_begin(GmDev d, uint32_t k, double xcells) {
    GmStats *st = d.st;
    st->kcur = k;
    st->halt = 0;
    st->nexec = 0;
    st->ndeferred = 0;
    st->npend = 0;
    st->npend0 = 0;
    st->pone = 0;
    st->nzeroed = 0;
    st->cwrap = 0;
    st->novf = 0;
    st->nplaced = 0;
    st->arr_pkey = 0ull;
    st->xcells = xcells;
}

hipError_t gm_launch_window_begin(const GmDev &d, uint32_t k, double xcells, hipStream_t s) {
    k_window_begin<<<1, 1, 0, s>>>(d, k, xcells);
    return hipGetLastError();
}


/* One wave after the commit pass: the step's row and whether the next step of
 * the window may run.  (A separate kernel: the kernel boundary is the grid-wide
 * barrier, where a last-block handoff inside the commit pass would need a
 * device-scope release per block, which on this GPU writes back the L2 of the
 * block's XCD.) */
__global__ void __launch_bounds__(64) k_step_tail(GmDev d, GmWindow w) {
    if (d.st->halt) return;
    step_tail(d, w);
}

hipError_t gm_launch_step_tail(const GmDev &d, const GmWindow &w, hipStream_t s) {
    k_step_tail<<<1, 64, 0, s>>>(d, w);
    return hipGetLastError();
}

/* DEFERRED cells: anchored at the step after their infection with their carry;
 * every step since then added the increment of the window's scalars, so the
 * value before step a.k is replayed from the anchor, then predicted.  The cells
 * come from the deferred list, or from a scan of every occupied slot when the
 * list overflowed. */
__global__ void __launch_bounds__(GM_BLOCK) k_predict_deferred(GmDev d, GmSweepArgs A) {
    const uint32_t nd = d.st->ndeferred;
    const bool listed = nd <= d.dlistCap;
    const uint32_t nI = d.st->nI, n = listed ? nd : nI + d.st->nB;
    for (uint32_t i = blockIdx.x * GM_BLOCK + threadIdx.x; i < n; i += gridDim.x * GM_BLOCK) {
        const uint32_t s = listed ? d.dlist[i] : occ_slot(d, nI, i);
        if (d.cstep[s] != GM_CS_DEFERRED) continue;
        GmRec *r = d.rec + s;
        const RecA a = rec_a(r);
        const RecC c = rec_c(r);
        const double P = resolved(d, c.P);
        if (is_pending(P)) {
            atomicOr(&d.st->err, GM_ERRF_MISSED);
            continue;
        }
        r->P = P;
        double inc, f = a.frac0;
        if (cell_increment(d, A.sp, P, c.defg, a.inf, inc)) f = gm_radd_n(f, inc, A.k - a.s0);
        predict_slot(d, A, s, f, A.k, P, c.defg, a.inf);
    }
}

hipError_t gm_launch_predict_deferred(const GmDev &d, const GmSweepArgs &a, uint32_t ndeferred, hipStream_t s) {
    if (a.nI + a.nB == 0) return hipSuccess;
    const uint32_t n = ndeferred <= d.dlistCap ? ndeferred : a.nI + a.nB;
    k_predict_deferred<<<grid_for(n, GM_BLOCK), GM_BLOCK, 0, s>>>(d, a);
    return hipGetLastError();
}

/* Cells whose weight lookup spilled (the cache had no room near the key's home):
 * with the cache grown, each looks its (nucleus, deformed grain) pair up again
 * from its record, which inserts a pending entry the host resolves before the
 * cell is predicted.  The cells come from the deferred list, or from a scan of
 * every occupied slot when the list overflowed. */
__global__ void k_spill_begin(GmDev d) {
    d.st->npend = 0;
    d.st->npend0 = 0;
    d.st->novf = 0;
}

__global__ void __launch_bounds__(GM_BLOCK) k_reinsert_spilled(GmDev d, GmPParams pp, uint32_t ndeferred, uint32_t nI,
                                                              uint32_t nB) {
    const bool listed = ndeferred <= d.dlistCap;
    const uint32_t n = listed ? ndeferred : nI + nB;
    for (uint32_t i = blockIdx.x * GM_BLOCK + threadIdx.x; i < n; i += gridDim.x * GM_BLOCK) {
        const uint32_t s = listed ? d.dlist[i] : occ_slot(d, nI, i);
        if (d.cstep[s] != GM_CS_DEFERRED) continue;
        GmRec *r = d.rec + s;
        if (!is_spilled(r->P)) continue;
        r->P = pair_weight(d, pp, r->rxg, r->defg);
    }
}

hipError_t gm_launch_reinsert_spilled(const GmDev &d, const GmPParams &pp, uint32_t ndeferred, uint32_t nI, uint32_t nB,
                                      hipStream_t s) {
    k_spill_begin<<<1, 1, 0, s>>>(d);
    const uint32_t n = ndeferred <= d.dlistCap ? ndeferred : nI + nB;
    if (n) k_reinsert_spilled<<<grid_for(n, GM_BLOCK), GM_BLOCK, 0, s>>>(d, pp, ndeferred, nI, nB);
    return hipGetLastError();
}
