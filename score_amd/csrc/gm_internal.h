/*
 * gm_internal.h -- device layout and kernel launchers of the growth machine.
 *
 * One domain lives in HBM as:
 *   grid   uint32[N]   deformed-grain id (< ndefg); not written during the run: only
 *                      deformed cells' entries are ever read then, and a cell
 *                      that transforms is appended to the retirement log
 *   rlog   (cell, ndefg+rxg) per retired cell in retirement order, written
 *                      coalesced with the step's parent data; the exported map
 *                      is the grid with the log applied and the front's cells
 *                      marked INFECTED
 *   claim  uint32[N]   per-cell infection claim key, 0xFFFFFFFF until claimed; a
 *                      claimed cell is infected in the same step and never
 *                      claimed again, so claims are never reset
 *   dmap   1 bit per cell, set while the cell is deformed (infectable), blocked so
 *          that one 128-byte line holds a 16x8x8 brick of cells: a 3x3x3
 *          neighbourhood usually lies in one or two lines
 *   front  capI+capB slots: Inside list slots [0,capI), Border list slots [capI,capI+capB)
 *          cstep uint16[slots]   completion step (mod GM_CS_MOD) or a GM_CS_*
 *                                sentinel; the only array every step reads in full
 *          rec   GmRec[slots]    64-byte record, read and written only when a cell
 *                                is infected, predicted, re-anchored or completes
 * The two lists keep SCORE's slot numbering (SCORE_Kernel.cpp:3587-3874) so that
 * slot-level state can be compared with a serial run.
 *
 * rxFrac is kept as an anchor (frac0 at the start of step s0) plus the increment
 * inc = ((v*dt)*inv_dx) that SCORE adds every step (SCORE_Kernel.cpp:4092).  While
 * the per-step scalars stay bit-identical, rxFrac after step k is frac0 plus
 * (k-s0+1) rounded additions of inc, so the step in which it first reaches 1.0
 * (cstep) is computed in advance (gm_radd.h), and the value it then holds is
 * replayed from the anchor when the cell completes.  When any scalar changes,
 * every active cell is re-anchored with the old increment and predicted again
 * with the new one.
 */
#ifndef SCORE_GM_INTERNAL_H
#define SCORE_GM_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <mutex>

#define GM_INFECTED 4096002000u
#define GM_FREE_CLAIM 0xFFFFFFFFu
#define GM_PENDING_TAG 0x7FF8000000000000ull  /* quiet NaN; low 32 bits carry the P-cache slot */
#define GM_PENDING_MASK 0xFFF8000000000000ull

/* Completion steps are stored in 16 bits.  A step cs is kept as cs % GM_CS_MOD
 * when it lies less than the horizon ahead of the step that predicted it, else
 * the cell is FAR; every cell is predicted again at each multiple of the
 * horizon.  So at step k every live cell's stored step lies in [k, k + horizon),
 * and cs % GM_CS_MOD names it.  A cell that completed keeps its code, which from
 * the next step on names a past step and reads as INACTIVE until a
 * re-anchoring (at the latest at the next multiple of the horizon) or a
 * defragmentation writes the INACTIVE sentinel.  Codes at or above GM_CS_MOD
 * are sentinels. */
typedef uint16_t gm_cs_t;
#define GM_CS_MOD 0x8000u
#define GM_CS_HORIZON_MAX 0x4000u  /* horizon in steps, at most GM_CS_MOD / 2 */
static_assert(GM_CS_HORIZON_MAX <= GM_CS_MOD / 2, "a past code must not wrap into the horizon before it is cleared");
#define GM_CS_INACTIVE 0xFFFFu
#define GM_CS_NEVER 0xFFFEu       /* active, does not reach rxFrac >= 1 within the run */
#define GM_CS_UNKNOWN 0xFFFDu     /* placed by nucleation, increment not yet known (the first step re-anchors) */
#define GM_CS_DEFERRED 0xFFFCu    /* mobility weight still pending: predicted at the end of the window */
#define GM_CS_FAR 0xFFFBu         /* active, completes at least a horizon after its prediction */
#define GM_STEP_NONE 0xFFFFFFFFu  /* no completion step */
#define GM_STEP_MAX 0xFFFFFFF0u   /* steps of a run stay below this */
#define GM_STEP_NEXT 0xFFFFFFFEu  /* does not complete in the predicted step; the next step predicts it again */

/* error flags raised on the device */
#define GM_ERRF_BORDER_FULL 0x1u
#define GM_ERRF_INSIDE_FULL 0x2u
#define GM_ERRF_PCACHE_FULL 0x4u
#define GM_ERRF_MISSED 0x8u       /* a prediction lies in the past: internal error */
#define GM_ERRF_KEYS 0x10u        /* more completed cells in one step than infection keys hold (2^27; 2^23 with regions) */
#define GM_ERRF_RLOG_FULL 0x20u   /* more retirements than the log holds (only cells retired twice can do it) */
#define GM_ERRF_HALO 0x40u        /* an infection into another region found no halo to park in: internal error */

/* the mobility-weight cache probes at most this many slots from a key's home;
 * a key that finds no room within them spills: its cells hold the pending
 * marker of slot GM_OVF_SLOT and the host re-inserts them into a larger table */
#define GM_PROBE_MAX 1024u
#define GM_NO_SITE 0xFFFFFFFFu    /* a nucleus left out of the site-saturated placement */

/* order-preserving unsigned key of a double (larger value, larger key; never 0
 * for a finite value), for atomicMax over weights */
static inline __host__ __device__ unsigned long long gm_pkey(double v) {
    unsigned long long b;
    __builtin_memcpy(&b, &v, sizeof(b));
    return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
static inline __host__ __device__ double gm_pkey_value(unsigned long long k) {
    const unsigned long long b = (k >> 63) ? (k & 0x7FFFFFFFFFFFFFFFull) : ~k;
    double v;
    __builtin_memcpy(&v, &b, sizeof(v));
    return v;
}
#define GM_OVF_SLOT 0xFFFFFFFFu

#define GM_BLOCK 256
/* The sweep reads a list in groups of GM_SW_TPC 64-slot tiles per wave; the
 * Inside list is rounded up to whole groups and the Border list's tiles follow. */
#define GM_SW_TPC 8u
static inline __host__ __device__ uint32_t gm_tiles_inside(uint32_t nI) {
    return ((nI + 63u) / 64u + GM_SW_TPC - 1u) & ~(GM_SW_TPC - 1u);
}
#define GM_MAX_BLOCKS 8192        /* grid of the partitioned passes (scanned by one 1024-thread block) */
#define GM_SCAN_BLOCKS 1024       /* fixed grid of the device-wide scans (nucleation) */
#define GM_ZEROED_INLINE 1022     /* exhausted grain ids returned with the per-step stats */
#define GM_PEND_INLINE 2048       /* new P-cache keys returned with the per-step stats */
#define GM_NK 9                   /* kernels per step (SCORE_GM_K_* order) */
#define GM_WMAX 16                /* steps per launch window at most (also below imin - 1) */

/* One front record: SCORE's `cell` (SCORE_Kernel.h:124-137) with rxFrac held as
 * anchor + increment.  The first 32-byte sector holds
 * what a prediction or a re-anchoring rewrites, so those write one sector. */
struct __attribute__((aligned(64))) GmRec {
    double frac0;   /* rxFrac before step s0 */
    uint32_t s0;    /* anchor step */
    uint32_t inf;   /* infector code 0..26 */
    double inc;     /* per-step increment, 0.0 when SCORE does not advance the cell */
    double spare;   /* unused, written as 0.0 with inc */
    double P;       /* mobility weight, or a pending marker */
    uint32_t cell;  /* global cell id x + y*nx + z*nx*ny */
    uint32_t defg;  /* deformed grain being consumed (mydefgid) */
    uint32_t rxg;   /* growing nucleus (myrxgid) */
    uint32_t pad[3];
};
static_assert(sizeof(GmRec) == 64, "GmRec is one 64-byte record");
/* the kernels move records as four 16-byte quarters */
static_assert(__builtin_offsetof(GmRec, inc) == 16 && __builtin_offsetof(GmRec, spare) == 24 &&
                  __builtin_offsetof(GmRec, P) == 32 && __builtin_offsetof(GmRec, cell) == 40 &&
                  __builtin_offsetof(GmRec, defg) == 44 && __builtin_offsetof(GmRec, rxg) == 48,
              "GmRec quarters: (frac0, s0, inf) | (inc, spare) | (P, cell, defg) | (rxg, pad)");

/* The completed cells of the current step, in FullRX list order (Inside list
 * first): what the infection kernels need of each parent, as three planes of
 * parentCap 16-byte entries, so that a wave's loads of one plane are contiguous.
 *   GM_PAR_IDS  (slot, cell, defg, rxg)
 *   GM_PAR_FP   (fin, P) as two doubles: rxFrac after the completing addition
 *               (written by the decide pass for parents that won a target) and
 *               the mobility weight
 *   GM_PAR_XYZ  (x, y, z, Border flag or list) */
#define GM_PAR_IDS 0u
#define GM_PAR_FP 1u
#define GM_PAR_XYZ 2u
#define GM_PAR_PLANES 3u

/* what one step of a window leaves for the host's master section */
struct GmStepRow {
    uint32_t nI, nB;          /* list lengths after the step */
    uint32_t nfullI, nfullB;
    uint32_t totRecI, totRecB;
    uint32_t recLenI, recLenB;
    uint32_t winI, winB;
    uint32_t npend;           /* P-cache entries the step inserted */
    uint32_t nattempt;
    uint32_t nzeroed;         /* exhausted grains so far in the window */
    uint32_t err;
    uint32_t nplaced;         /* nuclei placed at the start of the step */
    uint32_t pone;            /* a new pair's weight was short-circuited to 1.0 in the step */
    unsigned long long pkey;  /* largest weight of the nuclei placed in the step (gm_pkey order), 0: none */
};

struct alignas(8) GmStats {
    uint32_t nI, nB;          /* nextSlotNeverActive of both lists */
    uint32_t sv;              /* ACTIVE slots seen by the sweep */
    uint32_t nfull;           /* parents (completed cells), Inside then Border */
    uint32_t nfullI, nfullB;
    uint32_t totRecI, totRecB;/* INACTIVE slots seen by the sweep */
    uint32_t recLenI, recLenB;/* recycled slots available (capped) */
    uint32_t winI, winB;      /* successful infections into each list */
    uint32_t nIold, nBold;    /* list lengths before this step's appends */
    uint32_t npend;           /* P-cache entries inserted this step, still unresolved */
    uint32_t err;             /* GM_ERRF_* */
    uint32_t pone;            /* set when a pair weight was short-circuited to 1.0 */
    uint32_t scan_tot[4];     /* totals written by the device-wide scans */
    uint32_t nattempt;        /* attempts that found a deformed target this step */
    uint32_t nzeroed;         /* deformed grains that lost their last cell this step */
    uint32_t cwrap;           /* a deformed-grain count was decremented past zero this step */
    uint32_t dfirst;          /* defragmentation: first INACTIVE slot of the list */
    uint32_t dcount;          /* defragmentation: ACTIVE slots moved; status vector: entries */
    uint32_t novf;            /* weight lookups that found the cache full (their cells wait for a larger table) */
    uint32_t rg_per;          /* with several regions: parents per block of the infection passes */
    uint32_t nplaced;         /* nuclei placed at the start of the current step */
    unsigned long long arr_pkey; /* largest weight of those nuclei (gm_pkey order) */
    unsigned long long nret;  /* entries of the retirement log (cells retired so far) */
    uint32_t zeroed_ids[GM_ZEROED_INLINE]; /* the first of them; all are in GmDev::zlist */
    unsigned long long pend_keys[GM_PEND_INLINE]; /* the first new P-cache keys; all are in GmDev::pend_key */
    /* launch windows: several steps run back to back; the host reads the rows at the end */
    uint32_t kcur;            /* index of the next step */
    uint32_t halt;            /* set after a step that needs the host: later steps of the window do nothing */
    uint32_t nexec;           /* steps executed in the window */
    uint32_t ndeferred;       /* cells whose prediction waits for a pending weight */
    uint32_t npend0;          /* npend before the current step */
    double xcells;            /* Xcells, advanced by the device within a window */
    GmStepRow rows[GM_WMAX];
};

/* stop conditions of a window, checked after every step */
struct GmWindow {
    double xstop;             /* Xcells at which the host must act (termination, defragmentation) */
    uint32_t top_grain;       /* the grain that sets rhomax (0xFFFFFFFF: none) */
    uint32_t pend_limit;      /* new P-cache entries that keep the cache at most half full */
    uint32_t defer;           /* pending weights may wait for the end of the window */
    uint32_t pad;
};

/* Several memory regions (SCORE with OMP_NUM_THREADS = T > 1, ompshar_ca2regions,
 * SCORE_GridPartitioning.cpp:69-166).  Region r = iz*npy + iy owns front lists
 * 2r (Inside) and 2r+1 (Border), laid out one after another in the slot arrays.
 * One halo of a region (ompshar_initialize_halo, SCORE_Kernel.cpp:2085-2127): the
 * region it lies in, its origin and extent in cells (extent 0: a dummy). */
struct GmHalo {
    uint32_t inregion;
    uint16_t org[3], ext[3];
};
static_assert(sizeof(GmHalo) == 16, "GmHalo is 16 bytes");

struct GmRegion {
    uint32_t lo[3], hi[3];      /* inclusive box */
    uint32_t nbor[27];          /* mynborregions ids: myself, faces, edges, corners */
    GmHalo halo[27];            /* halo towards each neighbour; [0] unused */
};

/* halos an infection can be parked in (the other 18 are dummies or always
 * shadowed by one of these, which come first and span the whole x range) */
#define GM_NHALO 8
/* rows of the per-block counts of the infection pass with several regions:
 * winners into the own Inside / Border list, attempts, halo winners by halo */
#define GM_RW_I 0
#define GM_RW_B 1
#define GM_RW_A 2
#define GM_RW_H0 3
#define GM_RROWS (GM_RW_H0 + GM_NHALO)
#define GM_BCNT_ROWS 16

/* per-list tables with several regions: nlists entries, the prefix arrays nlists+1 */
struct GmLists {
    uint32_t *base;             /* first global slot of each list (multiples of 64) */
    uint32_t *cap;              /* slots of each list */
    uint32_t *rcap;             /* recycling-list capacity of each list (fixed, as SCORE's) */
    uint32_t *rbase;            /* start of each list's recycled slots in rbuf */
    uint32_t *n;                /* nextSlotNeverActive */
    uint32_t *nold;             /* the same before the current step's appends */
    uint32_t *rlen;             /* recycled slots that may be handed out this step */
    uint32_t *win;              /* this step's new records into each list */
    uint32_t *occ;              /* exclusive prefix of n: occupied slot i lies in list l iff occ[l] <= i < occ[l+1] */
    uint32_t *tile;             /* exclusive prefix of the sweep's tiles per list (whole 8-tile chunks) */
    uint32_t *tile_s;           /* the same as the current step's sweep saw it (tile moves on with the lengths) */
    uint32_t *par;              /* completed cells before each list's first, in FullRX order */
    uint32_t *rst;              /* INACTIVE slots before each list's first */
    uint32_t *rbuf;             /* recycled slots (list-relative), ascending per list */
    uint32_t *rs;               /* per region r (and r = T: totals), GM_RROWS prefix counts at its first parent */
    uint32_t *bstart;           /* first block of each region in the infection passes (nreg+1: blocks in use) */
    uint32_t *breg;             /* region of each block of the infection passes (GM_MAX_BLOCKS; ~0u past those in use) */
    uint32_t *hoff;             /* per region, GM_NHALO offsets of the halo winners in its Border list */
};

struct GmDev {
    /* geometry */
    uint32_t nx, ny, nz, nxy;
    uint64_t N;
    uint32_t ndefg, nrxg;
    uint32_t capI, capB;
    uint32_t recCapI, recCapB;
    /* state */
    uint32_t *grid, *claim;
    uint2 *rlog;                /* retirement log, rlogCap entries */
    uint64_t rlogCap;
    uint32_t *dmap;
    uint32_t dnbx, dnby, dnbz;  /* bricks of the deformed-cell bitmap per axis */
    gm_cs_t *cstep;
    uint32_t horizon;           /* steps, in [16, GM_CS_HORIZON_MAX] */
    GmRec *rec;
    /* per-step lists */
    uint32_t *pslot;            /* completed cells (global slots), Inside then Border, ascending each */
    uint4 *par;                 /* the same cells' parent data: GM_PAR_PLANES planes */
    uint32_t *recI, *recB;      /* recycled slots (list-relative), ascending, capped */
    uint64_t *mF, *mR;          /* per 64-slot tile: completed / INACTIVE masks */
    uint32_t *pmI, *pmB;        /* per parent: winning ranks into Inside / Border */
    /* pslot, par, pmI and pmB hold parentCap entries */
    uint32_t *bcnt;             /* per-block counts and offsets, 8 x GM_MAX_BLOCKS */
    uint32_t *dlist;            /* slots of the cells left DEFERRED in a window (dlistCap entries) */
    uint32_t dlistCap;
    uint32_t tileCap, parentCap;
    /* grains */
    double *rho, *defq, *rxq;
    uint32_t *defcnt, *rxcnt;
    uint32_t *zlist;            /* ids of grains exhausted in the current step */
    /* mobility-weight cache: open addressing on key rxg*ndefg+defg+1 */
    unsigned long long *hkeys;
    double *hvals;
    uint32_t hcap;
    uint32_t *pend_slot;
    unsigned long long *pend_key;
    /* nuclei arriving after step 0: pool ids grouped by step (arr_off[k] ..
     * arr_off[k+1] for step k, 1 <= k <= arr_last; arr_last 0: none), and every
     * nucleus' site and host-computed weight */
    uint32_t *arr_ids, *arr_off;
    uint32_t arr_last;
    uint32_t *nsite;
    double *nucP;
    /* nucleation scratch */
    uint32_t *nflag, *nI_, *nB_;
    uint32_t *scan_bsum;        /* 2*GM_SCAN_BLOCKS */
    GmStats *st;
    uint32_t *prog;             /* mapped host word: (steps the window has finished << 1) | halt */
    /* several memory regions (nreg > 1); with one region none of this is used */
    uint32_t nreg, nlists;      /* T and 2T */
    uint32_t npy, npz, nry, nrz;/* region grid and slab thickness along y and z */
    GmRegion *reg;              /* nreg entries */
    GmLists L;
    uint32_t *pmH;              /* per parent: winning ranks into another region's halo */
    uint32_t *plist;            /* per parent: its list */
};

/* mobility-weight evaluation parameters the kernels need */
struct GmPParams {
    int model;                  /* 2 = Rollett-Holm: device short-circuits P == 1.0 */
    double omega_hi;            /* omega <= omega_hi  =>  P == 1.0 exactly (RH); < 0 disables */
    double fac;                 /* 1/sqrt(2) as computed by the host */
};

struct GmSweepParams {
    int model;
    double mRH, mLAGB, mHAGB, mGS;
    double Gbhalfsq, zener, dt, inv_dx;
};

/* The step index and list lengths live in GmStats (kcur, nI, nB) so that a
 * window of steps runs without the host; nI and nB here only size the grids
 * (upper estimates for the window). */
struct GmSweepArgs {
    GmSweepParams sp;
    uint32_t k;                 /* step index (kernels read st->kcur) */
    uint32_t reanchor;          /* scalars changed since the previous step */
    uint32_t nI, nB;            /* occupied slots, for the grid sizes */
    uint32_t stepcap;           /* last step a completion can fall in */
    uint32_t next_only;         /* every following step re-anchors: predict only whether a cell
                                 * completes in the step it is predicted for (else FAR) */
    uint32_t pad;
};

/* launchers (gm_kernels.hip); all asynchronous on `s`.  With `ev` non-null,
 * ev[i] is recorded before kernel i of the step and ev[GM_NK] after the last. */
hipError_t gm_launch_sweep(const GmDev &d, const GmSweepArgs &a, hipStream_t s, hipEvent_t *ev);
/* every active cell advanced to step a.k with its old increment and predicted again */
hipError_t gm_launch_reanchor(const GmDev &d, const GmSweepArgs &a, hipStream_t s);
/* the same two halves of a step with several regions (gm_regions.hip) */
hipError_t gm_launch_rg_sweep(const GmDev &d, const GmSweepArgs &a, uint32_t parents_upper, hipStream_t s,
                              hipEvent_t *ev);
hipError_t gm_launch_rg_infect(const GmDev &d, const GmPParams &pp, const GmSweepArgs &a, uint32_t parents_upper,
                               hipStream_t s, hipEvent_t *ev);
/* with several regions: the lists' prefixes and summed lengths from d.L.n */
hipError_t gm_launch_rg_prefix(const GmDev &d, hipStream_t s);
/* with several regions: nuclei the host placed, records scattered to their slots */
hipError_t gm_launch_rg_place(const GmDev &d, const uint32_t *slots, const GmRec *recs, uint32_t n, hipStream_t s);
hipError_t gm_launch_infect(const GmDev &d, const GmPParams &pp, const GmSweepArgs &a, uint32_t parents_upper,
                            hipStream_t s, hipEvent_t *ev);
/* start of a window: step index and Xcells; clears the window's accumulators */
hipError_t gm_launch_window_begin(const GmDev &d, uint32_t k, double xcells, hipStream_t s);
/* end of every step of a window: records the step's row, advances the step
 * index, and sets halt when the host must act before the next step */
hipError_t gm_launch_step_tail(const GmDev &d, const GmWindow &w, hipStream_t s);
/* cells left DEFERRED in a window, with their weights now resolved: replayed to
 * step a.k under the window's scalars and predicted */
hipError_t gm_launch_predict_deferred(const GmDev &d, const GmSweepArgs &a, uint32_t ndeferred, hipStream_t s);
hipError_t gm_launch_nucleate(const GmDev &d, const uint32_t *site, const double *nucP, uint32_t nrxg, hipStream_t s);
/* one block per step, between the re-anchoring and the sweep: the nuclei arriving
 * at the step the device is at, placed in pool order and predicted under a's scalars */
hipError_t gm_launch_arrivals(const GmDev &d, const GmSweepArgs &a, hipStream_t s);
hipError_t gm_launch_resolve_pending(const GmDev &d, const double *vals, uint32_t npend, hipStream_t s);
hipError_t gm_launch_stats_out(const GmDev &d, GmStats *out, hipStream_t s);
/* SCORE's view of slots [base, base+n) before step k, unpacked into per-field arrays */
struct GmFrontOut {
    uint8_t *act, *inf;
    uint32_t *cell, *defg, *rxg;
    double *frac, *P;
};
hipError_t gm_launch_materialize(const GmDev &d, uint32_t k, uint32_t base, uint32_t n, const GmFrontOut &o,
                                 hipStream_t s);
hipError_t gm_launch_dmap_init(const GmDev &d, hipStream_t s);
/* ompshar_defragmentation on the list whose slot 0 is global slot `base`:
 * first_inactive leaves in st->dfirst the first INACTIVE slot in [lo, hi), or hi;
 * move compacts the ACTIVE slots of (first, nnever) to [first, first+count) through
 * the scratch records trec/tcs (nnever-first-1 entries), leaves count in st->dcount
 * and freezes the ACTIVE records left behind at their value before step k. */
hipError_t gm_launch_defrag_first_inactive(const GmDev &d, uint32_t base, uint32_t lo, uint32_t hi, uint32_t k,
                                           hipStream_t s);
hipError_t gm_launch_defrag_move(const GmDev &d, uint32_t base, uint32_t first, uint32_t nnever, uint32_t k,
                                 GmRec *trec, gm_cs_t *tcs, hipStream_t s);
/* ompcrit_clarify_status_infectedcells: (cell, grain shown) of every ACTIVE record,
 * Inside then Border, in slot order, before step k; the count goes to st->dcount.
 * ixyz and gid hold nI+nB entries. */
hipError_t gm_launch_status(const GmDev &d, uint32_t nI, uint32_t nB, uint32_t k, double thr, uint32_t *ixyz,
                            uint32_t *gid, hipStream_t s);
/* the cell grid as SCORE holds it, from a copy of d.grid in `out`: every retired
 * cell set to ndefg+rxg from the log, every cell held by an active front record
 * marked INFECTED */
hipError_t gm_launch_overlay_infected(const GmDev &d, uint32_t nI, uint32_t nB, uint32_t k, uint32_t *out,
                                      hipStream_t s);
hipError_t gm_launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s);
hipError_t gm_launch_fill_u64(unsigned long long *p, unsigned long long v, size_t n, hipStream_t s);
hipError_t gm_launch_hash_insert(const GmDev &d, const unsigned long long *keys, const double *vals, uint32_t n,
                                 hipStream_t s);
/* after a window whose weight lookups spilled (st->novf): the DEFERRED cells
 * holding the spill marker look their pair up again in the (grown) cache,
 * inserting pending entries; clears st->npend and st->novf first */
/* The fixed device array of the entry point that allocates nothing
 * (gm_noalloc.hip): its address on `device`, which must be current, looked up
 * on the device's first use; `lock` holds the device's array for the caller
 * until it is released. */
hipError_t gm_scratch_acquire(int device, uint32_t **ptr, std::unique_lock<std::mutex> &lock);
hipError_t gm_launch_reinsert_spilled(const GmDev &d, const GmPParams &pp, uint32_t ndeferred, uint32_t nI, uint32_t nB,
                                      hipStream_t s);

#endif
