/*
 * gm_internal.h -- device layout and kernel launchers of the growth machine.
 *
 * One domain lives in HBM as:
 *   grid   uint32[N]   cell states: deformed id (< ndefg), ndefg+rxg (transformed) or INFECTED
 *   claim  uint32[N]   per-cell infection claim key, 0xFFFFFFFF when free
 *   front  SoA over capI+capB slots: Inside list slots [0,capI), Border list slots [capI,capI+capB)
 *          act u8, inf u8, cell u32, defg u32, rxg u32, frac f64, P f64
 * The two lists keep SCORE's slot numbering (SCORE_Kernel.cpp:3587-3874) so that
 * slot-level state can be compared with a serial run.
 */
#ifndef SCORE_GM_INTERNAL_H
#define SCORE_GM_INTERNAL_H

#include <hip/hip_runtime.h>
#include <stdint.h>

#define GM_INFECTED 4096002000u
#define GM_FREE_CLAIM 0xFFFFFFFFu
#define GM_PENDING_TAG 0x7FF8000000000000ull  /* quiet NaN; low 32 bits carry the P-cache slot */
#define GM_PENDING_MASK 0xFFF8000000000000ull

/* error flags raised on the device */
#define GM_ERRF_BORDER_FULL 0x1u
#define GM_ERRF_INSIDE_FULL 0x2u
#define GM_ERRF_PCACHE_FULL 0x4u

#define GM_TILE 64          /* slots per compaction tile (one wave) */
#define GM_BLOCK 256
#define GM_SCAN_BLOCKS 1024 /* fixed grid of the device-wide scans */

struct GmStats {
    uint32_t nI, nB;          /* nextSlotNeverActive of both lists */
    uint32_t sv;              /* ACTIVE slots seen by the sweep */
    uint32_t nfull;           /* parents (completed cells), Inside then Border */
    uint32_t nfullI;
    uint32_t totRecI, totRecB;/* INACTIVE slots seen by the sweep */
    uint32_t recLenI, recLenB;/* recycled slots available (capped) */
    uint32_t winI, winB;      /* successful infections into each list */
    uint32_t npend;           /* P-cache entries inserted this step, still unresolved */
    uint32_t err;             /* GM_ERRF_* */
    uint32_t pone;            /* set when a pair weight was short-circuited to 1.0 */
    unsigned long long rhomax_bits; /* max rho over deformed grains with cells left */
    uint32_t scan_tot[4];     /* totals written by the device-wide scans */
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
    uint8_t *act, *inf;
    uint32_t *cell, *defg, *rxg;
    double *frac, *P;
    /* per-step compaction */
    uint64_t *mfull, *mrec;     /* per tile masks */
    uint32_t *tcntF, *tcntR;    /* per tile counts, then exclusive offsets */
    uint32_t *full;             /* parents as global slot numbers */
    uint32_t *recI, *recB;      /* recycled slots (list-relative), ascending */
    uint32_t *pwinI, *pwinB;    /* per-parent winner masks over the 26 direction ranks */
    uint32_t *pcntI, *pcntB;    /* per-parent counts, then exclusive offsets */
    uint32_t tileCap, parentCap;
    /* grains */
    double *rho, *defq, *rxq;
    uint32_t *defcnt, *rxcnt;
    /* mobility-weight cache: open addressing on key rxg*ndefg+defg+1 */
    unsigned long long *hkeys;
    double *hvals;
    uint32_t hcap;
    uint32_t *pend_slot;
    unsigned long long *pend_key;
    /* scans */
    uint32_t *scan_bsum;        /* 2*GM_SCAN_BLOCKS */
    GmStats *st;
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

/* launchers (gm_kernels.hip); all asynchronous on `s` */
hipError_t gm_launch_begin_step(const GmDev &d, hipStream_t s);
hipError_t gm_launch_sweep(const GmDev &d, const GmSweepParams &sp, uint32_t nI, uint32_t nB, hipStream_t s);
hipError_t gm_launch_compact(const GmDev &d, uint32_t nI, uint32_t nB, hipStream_t s);
hipError_t gm_launch_infect(const GmDev &d, const GmPParams &pp, hipStream_t s);
hipError_t gm_launch_rhomax(const GmDev &d, hipStream_t s);
hipError_t gm_launch_nucleate(const GmDev &d, const uint32_t *site, const double *nucP, uint32_t nrxg, hipStream_t s);
hipError_t gm_launch_resolve_pending(const GmDev &d, const double *vals, uint32_t npend, hipStream_t s);
hipError_t gm_launch_resolve_front(const GmDev &d, uint32_t nI, uint32_t nB, hipStream_t s);
hipError_t gm_launch_fill_u32(uint32_t *p, uint32_t v, size_t n, hipStream_t s);
hipError_t gm_launch_fill_u64(unsigned long long *p, unsigned long long v, size_t n, hipStream_t s);
hipError_t gm_launch_fill_f64_bits(double *p, unsigned long long bits, size_t n, hipStream_t s);
hipError_t gm_launch_hash_insert(const GmDev &d, const unsigned long long *keys, const double *vals, uint32_t n,
                                 hipStream_t s);
hipError_t gm_launch_defrag(const GmDev &d, uint32_t list, uint32_t first, uint32_t n, uint32_t *tmp_u32,
                            uint8_t *tmp_u8, double *tmp_f64, hipStream_t s);

#endif
