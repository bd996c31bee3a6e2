/*
 * gm_host.cpp -- C ABI of the growth machine (include/score_gm.h).
 *
 * The host does what SCORE's master thread does between growth steps
 * (SCORE_Kernel.cpp:6688-6742): it advances t, re-evaluates T(t), G(T), the
 * mobilities and the adaptive time step with the same FP64 expressions and the
 * same glibc transcendentals, so every scalar is bit-identical to the
 * reference's.  It also evaluates mobility weights P that the device cannot
 * reproduce bit-exactly (glibc acos/pow/exp) when a new nucleus/deformed-grain
 * pair is met; the device caches them.  All per-cell work runs in
 * gm_kernels.hip.
 */
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/score_gm.h"
#include "gm_internal.h"
#include "score_orient.h"

namespace {

/* constants (SCORE_Defs_Physics.h:51-160, SCORE_Math.h:61) */
const double kBoltzmann = 1.3806488e-23;
const double kPi = 3.1415926535897932384626433832795;
const double kLagbHagb = ((15.0) / (180.0) * (kPi));
const double kMaxDev40_111 = ((10.0) / (180.0) * (kPi));
const double kInitialDeltaT = 1e-6;
const double kInfinite = 1e36;
const double kSmallHeatRate = 0.000277777777777777;
const double kSmallHeat = 1.0;
const double kSmallDraggingRate = 0.027777777; /* SMALL_DRAGGINGRATE, SCORE_Defs_Physics.h:84 */
const double kSmallRecoveryRate = 2.77e10;     /* SMALL_RECOVERYRATE, :83 */
const double kSmallRho = 1.0e11;               /* SMALL_RHO, :86 */
const double kSmallZenerForce = 100.0;         /* SMALL_ZENERFORCE, :87 */
const double kRhoWellAnnealed = 1e8;
const double kDefaultPmax = 0.0;
const double kDefaultSmallNumber = 1e-6;
const double kFullRecyclingAtt = 3.0;
const int64_t kDefaultNmax = 1000000;          /* DEFAULT_NMAX, SCORE_Defs_Physics.h:133 */

bool same_bits(double a, double b) { return memcmp(&a, &b, sizeof(double)) == 0; }

bool same_scalars(const GmSweepParams &a, const GmSweepParams &b) {
    return a.model == b.model && same_bits(a.mRH, b.mRH) && same_bits(a.mLAGB, b.mLAGB) &&
           same_bits(a.mHAGB, b.mHAGB) && same_bits(a.mGS, b.mGS) && same_bits(a.Gbhalfsq, b.Gbhalfsq) &&
           same_bits(a.zener, b.zener) && same_bits(a.dt, b.dt) && same_bits(a.inv_dx, b.inv_dx);
}

thread_local std::string g_last_error;

int fail(int code, const char *what) {
    g_last_error = what;
    return code;
}

int hip_fail(hipError_t e, const char *where) {
    g_last_error = std::string(where) + ": " + hipGetErrorString(e);
    return SCORE_GM_ERR_HIP;
}

#define HIPCHK(x)                                              \
    do {                                                       \
        hipError_t e_ = (x);                                   \
        if (e_ != hipSuccess) return hip_fail(e_, #x);         \
    } while (0)

/* ---------------------------------------------------------------- orientation math */

using score_orient::disori_angle;
using score_orient::disori_quat;

double rh_weight_of_theta(const score_gm_params &p, double theta) {
    return 1.0 - (p.rh_cut * exp(-1.0 * p.rh_trans * pow((theta / kLagbHagb), p.rh_exponent)));
}

/* calc_mobilityweight_rollettholm / _sebaldgottstein (SCORE_GrainBoundaryPhysics.cpp:50-135),
 * identical arithmetic to reg_calc_mobilityweight (SCORE_Kernel.cpp:2460-2527) */
double mobility_weight(const score_gm_params &p, const double *qrx, const double *qdef) {
    const double theta = disori_angle(qrx, qdef);
    if (p.mobility_model == SCORE_GM_MOB_ROLLETTHOLM) return rh_weight_of_theta(p, theta);
    if (theta <= kLagbHagb) return -1.0;
    double qdis[4] = {1.0, 0.0, 0.0, 0.0};
    disori_quat(qrx, qdef, qdis);
    const double sq3 = 1 / sqrt(3.0);
    const double oneNinth = 1.0 / 9.0;
    const double m40[4] = {cos(oneNinth * kPi), sq3 * sin(oneNinth * kPi), sq3 * sin(oneNinth * kPi),
                           sq3 * sin(oneNinth * kPi)};
    const double dev = disori_angle(qdis, m40);
    double w = 0.0;
    if (dev < kMaxDev40_111) {
        const double cc = cos(0.5 * kPi * dev / kMaxDev40_111);
        w = cc * cc;
    }
    return w;
}

/* Largest omega for which the Rollett-Holm weight is exactly 1.0 with margin:
 * the device returns P = 1.0 for omega <= this value and asks the host otherwise.
 * Returns -1 when no safe threshold exists (shortcut disabled). */
double rh_omega_threshold(const score_gm_params &p) {
    if (p.mobility_model != SCORE_GM_MOB_ROLLETTHOLM) return -1.0;
    if (!(p.rh_cut > 0.0 && p.rh_trans > 0.0 && p.rh_exponent > 0.0)) return -1.0;
    if (rh_weight_of_theta(p, kPi) != 1.0) return -1.0;
    double lo = 0.0, hi = kPi; /* invariant: weight(hi) == 1.0 */
    if (rh_weight_of_theta(p, lo) == 1.0) return 1.0; /* every pair has P == 1.0 */
    for (int it = 0; it < 200; it++) {
        double mid = 0.5 * (lo + hi);
        if (rh_weight_of_theta(p, mid) == 1.0) hi = mid;
        else lo = mid;
    }
    /* margin of 0.2 degree: the weight's deficit falls by a large factor over it */
    const double theta_hi = hi + 0.2 / 180.0 * kPi;
    if (theta_hi >= 2.0 * acos(0.85)) return -1.0; /* beyond the largest cubic disorientation */
    for (int i = 0; i <= 200000; i++) {
        double th = theta_hi + (kPi - theta_hi) * (double)i / 200000.0;
        if (rh_weight_of_theta(p, th) != 1.0) return -1.0;
    }
    double om = cos(0.5 * theta_hi);
    /* a few ulps below, so that 2*acos(om') >= theta_hi whatever the last-bit behaviour of acos */
    for (int i = 0; i < 4; i++) om = nextafter(om, 0.0);
    if (!(2.0 * acos(om) >= theta_hi)) return -1.0;
    return om;
}

}  // namespace

/* ------------------------------------------------------------------ handle */

struct score_gm_handle {
    int device = 0;
    hipStream_t stream = nullptr;
    hipMemPool_t pool = nullptr; /* the library's own pool on this device */
    GmDev d{};
    GmStats *hst = nullptr; /* pinned mirror of d.st (mapped, coherent), written by k_stats_out */
    GmStats *hst_dev = nullptr; /* its device address */
    uint32_t *hprog = nullptr;  /* pinned progress word (mapped, coherent), written by k_step_tail */
    uint32_t *hprog_dev = nullptr;
    score_gm_params p{};
    std::vector<double> sched_t, sched_T, defrag_at;
    std::vector<double> defq, rxq;
    double cellsize = 0.0, inv_dx = 0.0;
    double tprocessingend = 0.0;
    /* caHdl scalars */
    double t = 0.0, X = 0.0, Xcells = 0.0, dt = 0.0, T = 0.0, G = 0.0, b = 0.0, Gbhalfsq = 0.0;
    double mLAGB = 0.0, mHAGB = 0.0, mGS = 0.0, mRHHAGB = 0.0, pzmin = 0.0, rhomax = 0.0;
    double fr = 0.0; /* dispersoid f/r (myDragData.fr) */
    std::vector<double> disp_t, disp_fr;
    double Pmax = kDefaultPmax, regPmax = kDefaultPmax;
    uint32_t step = 0;
    uint32_t nI = 0, nB = 0;
    uint32_t nuclei = 0;
    int completed = 0;
    int err = 0;
    uint64_t sum_sv = 0;
    double loop_seconds = 0.0;
    GmPParams pp{};
    uint32_t hcount = 0; /* entries in the P cache */
    std::vector<score_gm_traj_row> traj;
    /* profiling */
    bool prof = false;
    hipEvent_t ev[GM_NK + 1] = {};
    score_gm_profile profc{};
    /* staging for pending P */
    std::vector<unsigned long long> pend_keys;
    double *hpend_vals = nullptr; /* pinned, mapped and coherent: k_resolve_pending reads it */
    double *hpend_dev = nullptr;  /* its device address */
    uint32_t pend_vals_cap = 0;
    /* scalars of the previous sweep: a change forces re-anchoring of rxFrac */
    GmSweepParams last_sp{};
    bool have_last_sp = false;
    uint32_t defrag_cnt = 0;     /* defragmentations done (reg_defrag_cnt) */
    uint32_t windows = 0;        /* launch windows run */
    bool force_reanchor = false; /* records moved: every cell is predicted again */
    /* deformed grains by descending rho and whether they still own cells */
    std::vector<double> rho_host;
    std::vector<double> rho0_host;     /* mydefgpool[].rho0 (Michalak-Paxton recovery) */
    bool rho_changed = false;          /* recovery changed rho: upload and re-anchor */
    std::vector<uint32_t> rho_order;
    std::vector<uint8_t> alive;
    size_t rho_top = 0;
    std::vector<uint32_t> zbuf;
    /* weights of the window's new pairs, in insertion order (steps in order; a
     * spilled step's re-inserted pairs last) */
    std::vector<double> pend_P;
    /* nuclei arriving after step 0: offsets per step (arr_off[k]..arr_off[k+1]) */
    std::vector<uint32_t> arr_off;
    /* grain-evolution log (mygrainevolution) */
    std::vector<double> glog_at;
    uint32_t glog_cnt = 0;
    std::vector<score_gm_grainlog_row> glog;
    std::vector<uint32_t> glog_counts;
    /* several memory regions: the region table and the lists' first slots and
     * capacities (2T lists, region r's Inside list 2r and Border list 2r+1) */
    std::vector<GmRegion> reg;
    std::vector<uint32_t> lbase, lcap, lrcap, lrbase;
};

namespace {

/* The library's own stream-ordered memory pool per device.  Freed blocks stay in
 * it while any handle of the library is alive on the device, so that handles
 * created next to running ones (batches, the request loop) reuse them without
 * a device-wide synchronisation; when the last handle on the device is
 * destroyed, the pool gives its unused memory back.  The device's default pool
 * and its settings are left alone. */
struct LibraryPool {
    std::mutex m;
    hipMemPool_t pool = nullptr;
    uint32_t handles = 0;
};
constexpr int kMaxDevices = 64;
LibraryPool g_pools[kMaxDevices];

int pool_acquire(int device, hipMemPool_t *out) {
    *out = nullptr;
    if (device < 0 || device >= kMaxDevices) return fail(SCORE_GM_ERR_NODEVICE, "device index out of range");
    LibraryPool &lp = g_pools[device];
    std::lock_guard<std::mutex> lk(lp.m);
    if (!lp.pool) {
        hipMemPoolProps props;
        memset(&props, 0, sizeof(props));
        props.allocType = hipMemAllocationTypePinned;
        props.handleTypes = hipMemHandleTypeNone;
        props.location.type = hipMemLocationTypeDevice;
        props.location.id = device;
        hipMemPool_t pool = nullptr;
        hipError_t e = hipMemPoolCreate(&pool, &props);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            return hip_fail(e, "hipMemPoolCreate");
        }
        uint64_t keep = UINT64_MAX;
        e = hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
        if (e != hipSuccess) {
            (void)hipMemPoolDestroy(pool);
            (void)hipGetLastError();
            return hip_fail(e, "hipMemPoolSetAttribute");
        }
        lp.pool = pool;
    }
    lp.handles++;
    *out = lp.pool;
    return 0;
}

/* after the handle's frees have completed (its stream synchronised) */
void pool_release(int device) {
    if (device < 0 || device >= kMaxDevices) return;
    LibraryPool &lp = g_pools[device];
    std::lock_guard<std::mutex> lk(lp.m);
    if (lp.handles == 0) return;
    if (--lp.handles == 0 && lp.pool) (void)hipMemPoolTrimTo(lp.pool, 0);
}

/* Pinned host memory that kernels read or write directly: mapped into the
 * device's address space and coherent (not cached by the device), so that
 * what a kernel stores is in host memory when the stream has been synchronised
 * and what the host stores before a launch is what the kernel reads. */
int pinned_mapped(score_gm_handle *h, void **host, void **dev, size_t bytes) {
    (void)h;
    *host = *dev = nullptr;
    hipError_t e = hipHostMalloc(host, bytes, hipHostMallocMapped | hipHostMallocCoherent);
    if (e == hipSuccess) {
        e = hipHostGetDevicePointer(dev, *host, 0);
        if (e != hipSuccess) {
            (void)hipHostFree(*host);
            *host = nullptr;
        }
    }
    if (e != hipSuccess) {
        *host = *dev = nullptr;
        (void)hipGetLastError();
        g_last_error = std::string("pinned host memory: ") + hipGetErrorString(e);
        return SCORE_GM_ERR_OOM;
    }
    return 0;
}

/* Device memory comes from the device's stream-ordered pool on the handle's
 * stream: a free does not wait for the whole device (as hipFree does), so
 * handles created and destroyed next to running ones do not stall them, and
 * freed blocks are reused by later handles. */
template <typename T>
int dalloc(score_gm_handle *h, T **p, size_t n) {
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMallocFromPoolAsync((void **)p, n * sizeof(T), h->pool, h->stream);
    if (e != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        g_last_error = std::string("hipMallocFromPoolAsync: ") + hipGetErrorString(e);
        return SCORE_GM_ERR_OOM;
    }
    return 0;
}

void dfree(score_gm_handle *h, void *p) {
    if (p) (void)hipFreeAsync(p, h->stream);
}

void free_device(score_gm_handle *h) {
    GmDev &d = h->d;
    void *ptrs[] = {d.grid, d.claim, d.rlog, d.dmap, d.cstep, d.rec, d.pslot, d.dlist,
                    d.par, d.recI, d.recB, d.mF, d.mR, d.pmI, d.pmB, d.bcnt, d.rho, d.defq,
                    d.rxq, d.defcnt, d.rxcnt, d.zlist, d.hkeys, d.hvals, d.pend_slot, d.pend_key, d.nflag, d.nI_, d.nB_,
                    d.scan_bsum, d.st, d.arr_ids, d.arr_off, d.nsite, d.nucP,
                    d.reg, d.pmH, d.L.base, d.L.cap, d.L.rcap, d.L.rbase, d.L.n, d.L.nold, d.L.rlen,
                    d.L.win, d.L.occ, d.L.tile, d.L.tile_s, d.L.par, d.L.rst, d.L.rbuf, d.L.rs, d.L.bstart, d.L.breg,
                    d.L.hoff, d.plist};
    for (void *q : ptrs) dfree(h, q);
    memset(&h->d, 0, sizeof(h->d));
    if (h->hst) (void)hipHostFree(h->hst);
    h->hst = nullptr;
    h->hst_dev = nullptr;
    if (h->hprog) (void)hipHostFree(h->hprog);
    h->hprog = nullptr;
    h->hprog_dev = nullptr;
    if (h->hpend_vals) (void)hipHostFree(h->hpend_vals);
    h->hpend_vals = nullptr;
    h->hpend_dev = nullptr;
    for (hipEvent_t &ev : h->ev)
        if (ev) { (void)hipEventDestroy(ev); ev = nullptr; }
    if (h->stream) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipStreamDestroy(h->stream);
    }
    h->stream = nullptr;
    if (h->pool) pool_release(h->device);
    h->pool = nullptr;
}

/* mathMethods::smallestYFactor (SCORE_Math.cpp:97-115): the y factor of the region
 * grid, the last of the listed primes that divides n before the next prime
 * reaches the cofactor */
uint32_t smallest_y_factor(uint32_t n) {
    static const uint32_t primes[32] = {1, 2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47,
                                        53, 59, 61, 67, 71, 73, 79, 83, 89, 97, 101, 103, 107, 109, 113, 127};
    uint32_t z = n, y = 1;
    for (uint32_t c = 0; c < 31; c++) {
        const uint32_t q = n / primes[c];
        if (q * primes[c] == n) {
            z = q;
            y = primes[c];
        }
        if (primes[c + 1] >= z) break;
    }
    return y;
}

int validate(const score_gm_domain *dom, const score_gm_params *prm) {
    if (!dom || !prm) return fail(SCORE_GM_ERR_ARG, "null domain or params");
    if (dom->nx < 3 || dom->ny < 3 || dom->nz < 3) return fail(SCORE_GM_ERR_ARG, "edge length below 3 cells");
    if (dom->nx > 1600 || dom->ny > 1600 || dom->nz > 1600)
        return fail(SCORE_GM_ERR_ARG, "edge length above 1600 cells (CA_DIMENSIONS_MAXIMUM)");
    if (!dom->cellgrid || !dom->ori_quat || !dom->defg_ori || !dom->defg_rho || dom->n_defg == 0)
        return fail(SCORE_GM_ERR_ARG, "missing domain arrays");
    if (dom->n_rxg && (!dom->rxg_ori || !dom->rxg_site)) return fail(SCORE_GM_ERR_ARG, "missing nucleus arrays");
    if (!(dom->cellsize > 0.0)) return fail(SCORE_GM_ERR_ARG, "cellsize must be positive");
    if (dom->n_defg >= 4096000000u || dom->n_rxg >= 4096000000u - dom->n_defg)
        return fail(SCORE_GM_ERR_ARG, "grain ids exceed CA_GRAINIDS_MAXIMUM");
    if (prm->mobility_model != SCORE_GM_MOB_ROLLETTHOLM && prm->mobility_model != SCORE_GM_MOB_SEBALDGOTTSTEIN)
        return fail(SCORE_GM_ERR_ARG, "unknown mobility model");
    if (prm->zener_mode < 0 || prm->zener_mode > 2) return fail(SCORE_GM_ERR_ARG, "unknown Zener drag mode");
    if (prm->zener_mode != 0) {
        if (prm->n_disp < (prm->zener_mode == 2 ? 2u : 1u) || !prm->disp_time || !prm->disp_fr)
            return fail(SCORE_GM_ERR_ARG, "Zener drag needs its f/r table (two rows for the time-dependent mode)");
        if (prm->zener_mode == 2) {
            /* SCORE reads the entry before the first one for times up to disp_time[0] */
            if (prm->disp_time[0] != 0.0) return fail(SCORE_GM_ERR_ARG, "time-dependent drag table must start at t = 0");
            for (uint32_t i = 1; i < prm->n_disp; i++)
                if (!(prm->disp_time[i] > prm->disp_time[i - 1]))
                    return fail(SCORE_GM_ERR_ARG, "drag table times must increase strictly");
        }
    }
    if (prm->recovery_mode < 0 || prm->recovery_mode > 4) return fail(SCORE_GM_ERR_ARG, "unknown recovery model");
    if (prm->recovery_mode == 4 && !(prm->rec_tau > 0.0))
        return fail(SCORE_GM_ERR_ARG, "exponential recovery needs a positive time constant");
    if (prm->n_grainlog != 0 && !prm->grainlog_at) return fail(SCORE_GM_ERR_ARG, "grain-evolution log points missing");
    if (prm->n_regions < 0 || prm->n_regions > 1024) return fail(SCORE_GM_ERR_ARG, "n_regions outside [0, 1024]");
    if (prm->n_regions > 1) {
        /* every region more than MINIMUM_REGION_SIZE cells thick (SCORE_GridPartitioning.cpp:87-93) */
        const uint32_t npy = smallest_y_factor((uint32_t)prm->n_regions), npz = (uint32_t)prm->n_regions / npy;
        if (dom->ny / npy <= 10 || dom->nz / npz <= 10)
            return fail(SCORE_GM_ERR_ARG, "regions must be more than 10 cells thick along y and z");
        /* SCORE's OpenMP path has no nucleation after step 0 (SCORE_Kernel.cpp:6632) */
        if (dom->rxg_step)
            for (uint32_t i = 0; i < dom->n_rxg; i++)
                if (dom->rxg_step[i] != 0)
                    return fail(SCORE_GM_ERR_UNSUPPORTED, "nuclei arriving after step 0 with more than one region");
    }
    if (prm->n_defrag != 0 && !prm->defrag_at) return fail(SCORE_GM_ERR_ARG, "defragmentation thresholds missing");
    if (prm->n_sched < 2 || !prm->sched_time || !prm->sched_temp)
        return fail(SCORE_GM_ERR_ARG, "annealing schedule needs at least two points");
    if (!(prm->maxfill > 0.0)) return fail(SCORE_GM_ERR_ARG, "maxfill must be positive");
    /* the range SCORE accepts for NMAX (SCORE_Kernel.cpp:274) */
    if (prm->nmax < 1 || prm->nmax > kDefaultNmax) return fail(SCORE_GM_ERR_ARG, "nmax outside [1, 1000000] (DEFAULT_NMAX)");
    uint64_t N = (uint64_t)dom->nx * dom->ny * dom->nz;
    for (uint32_t i = 0; i < dom->n_defg; i++)
        if (dom->defg_ori[i] >= dom->n_ori) return fail(SCORE_GM_ERR_ARG, "defg_ori out of range");
    for (uint32_t i = 0; i < dom->n_rxg; i++)
        if (dom->rxg_ori[i] >= dom->n_ori || dom->rxg_site[i] >= N)
            return fail(SCORE_GM_ERR_ARG, "nucleus orientation or site out of range");
    for (uint64_t i = 0; i < N; i++)
        if (dom->cellgrid[i] >= dom->n_defg) return fail(SCORE_GM_ERR_ARG, "cellgrid holds a non-deformed id");
    return 0;
}

/* The region grid of ompshar_ca2regions (SCORE_GridPartitioning.cpp:69-166):
 * npy = smallestYFactor(T) slabs along y, npz = T / npy along z, x not split;
 * region iz*npy + iy spans [iy*nry, iy*nry + nry - 1] (the last slab up to
 * ny - 1) and likewise along z.  Its neighbour table (ompshar_init_neighbortopology,
 * SCORE_Kernel.cpp:2028-2080, periodic on the region grid) and halos
 * (ompshar_init_haloregions, :2223-2457, origins as determine_origins,
 * :2134-2220: the neighbour's layer next to the region, a dummy where the
 * neighbour is the region itself).  Then the lists: region r's Inside list may
 * hold every cell of the region; its Border list and both recycling lists have
 * SCORE's sizes for the region (ompshar_init_myRXFrontInside/Border,
 * SCORE_Kernel.cpp:3587-3710). */
void build_regions(score_gm_handle *h, uint32_t T) {
    static const int8_t dir[27][3] = {
        {0, 0, 0},
        {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1},
        {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {1, 0, 1},
        {-1, -1, 0}, {-1, 1, 0}, {1, -1, 0}, {1, 1, 0},
        {0, -1, -1}, {0, -1, 1}, {0, 1, -1}, {0, 1, 1},
        {-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1}};
    GmDev &d = h->d;
    d.nreg = T;
    d.nlists = 2 * T;
    d.npy = smallest_y_factor(T);
    d.npz = T / d.npy;
    d.nry = d.ny / d.npy;
    d.nrz = d.nz / d.npz;
    h->reg.assign(T, GmRegion{});
    for (uint32_t iz = 0; iz < d.npz; iz++)
        for (uint32_t iy = 0; iy < d.npy; iy++) {
            GmRegion &R = h->reg[iz * d.npy + iy];
            R.lo[0] = 0;
            R.hi[0] = d.nx - 1;
            R.lo[1] = iy * d.nry;
            R.hi[1] = iy == d.npy - 1 ? d.ny - 1 : iy * d.nry + d.nry - 1;
            R.lo[2] = iz * d.nrz;
            R.hi[2] = iz == d.npz - 1 ? d.nz - 1 : iz * d.nrz + d.nrz - 1;
        }
    const uint32_t dims[3] = {d.nx, d.ny, d.nz};
    for (uint32_t r = 0; r < T; r++) {
        GmRegion &R = h->reg[r];
        const int iy = (int)(r % d.npy), iz = (int)(r / d.npy);
        for (int nb = 0; nb < 27; nb++) {
            int jy = iy + dir[nb][1], jz = iz + dir[nb][2];
            if (jy < 0) jy += (int)d.npy;
            if (jy >= (int)d.npy) jy -= (int)d.npy;
            if (jz < 0) jz += (int)d.npz;
            if (jz >= (int)d.npz) jz -= (int)d.npz;
            R.nbor[nb] = (uint32_t)jz * d.npy + (uint32_t)jy;
        }
        for (int c = 0; c < 27; c++) {
            GmHalo &H = R.halo[c];
            H.inregion = R.nbor[c];
            if (c == 0 || R.nbor[c] == r) continue;
            for (int a = 0; a < 3; a++) {
                int org;
                if (dir[c][a] < 0) {
                    org = (int)R.lo[a] - 1;
                    if (org < 0) org += (int)dims[a];
                } else if (dir[c][a] == 0) {
                    org = (int)R.lo[a];
                } else {
                    org = (int)R.hi[a] + 1;
                    if (org >= (int)dims[a]) org -= (int)dims[a];
                }
                H.org[a] = (uint16_t)org;
                H.ext[a] = (uint16_t)(dir[c][a] == 0 ? R.hi[a] - R.lo[a] + 1 : 1);
            }
        }
    }
    const uint32_t L = 2 * T;
    h->lcap.assign(L, 0);
    h->lrcap.assign(L, 0);
    h->lbase.assign(L + 1, 0);
    h->lrbase.assign(L + 1, 0);
    for (uint32_t r = 0; r < T; r++) {
        const GmRegion &R = h->reg[r];
        const uint32_t ex = R.hi[0] - R.lo[0] + 1, ey = R.hi[1] - R.lo[1] + 1, ez = R.hi[2] - R.lo[2] + 1;
        const double vol = (double)((uint64_t)ex * ey * ez);
        const uint32_t capB = 2 * ((ex * ey) + (ex * ez) + (ey * ez));
        h->lcap[2 * r] = ex * ey * ez;
        h->lcap[2 * r + 1] = capB;
        h->lrcap[2 * r] = (uint32_t)(h->p.rel_cell_caching * vol * h->p.maxfill * kFullRecyclingAtt);
        h->lrcap[2 * r + 1] = (uint32_t)((double)capB * h->p.maxfill * kFullRecyclingAtt);
    }
    for (uint32_t l = 0; l < L; l++) {
        h->lbase[l + 1] = h->lbase[l] + (h->lcap[l] + 63) / 64 * 64;
        h->lrbase[l + 1] = h->lrbase[l] + h->lrcap[l];
    }
}

/* Site-saturated nucleation with several regions (ompshar_sim_myCA_sitesaturatedNucleation,
 * SCORE_GrowthMachine.cpp:50-121; infect_JuvenileNucleation, :124-189): each
 * region places the nuclei whose site lies in its box, in pool order, appending
 * a record (rxFrac 1.0, infector 26) to its Inside list for a site strictly
 * inside the box, else to its Border list.  site0[k] is nucleus k's site or
 * GM_NO_SITE; a site met again holds a nucleus already and rejects it.  The
 * records go to the device with the lists' tables and lengths. */
int place_nuclei_regions(score_gm_handle *h, const score_gm_domain *dom, const std::vector<uint32_t> &site0,
                         const std::vector<double> &nucP) {
    GmDev &d = h->d;
    hipStream_t s = h->stream;
    const uint32_t L = d.nlists;
    std::vector<uint32_t> cnt(L, 0u), slots;
    std::vector<GmRec> recs;
    std::vector<uint64_t> seen((d.N + 63) / 64, 0ull);
    for (uint32_t k = 0; k < d.nrxg; k++) {
        const uint32_t w = site0[k];
        if (w == GM_NO_SITE || (seen[w >> 6] >> (w & 63) & 1ull)) continue;
        seen[w >> 6] |= 1ull << (w & 63);
        const uint32_t z = w / d.nxy, rem = w - z * d.nxy, y = rem / d.nx, x = rem - y * d.nx;
        uint32_t iy = y / d.nry, iz = z / d.nrz;
        if (iy >= d.npy) iy = d.npy - 1;
        if (iz >= d.npz) iz = d.npz - 1;
        const uint32_t r = iz * d.npy + iy;
        const GmRegion &R = h->reg[r];
        const bool inside = x > R.lo[0] && x < R.hi[0] && y > R.lo[1] && y < R.hi[1] && z > R.lo[2] && z < R.hi[2];
        const uint32_t l = 2 * r + (inside ? 0u : 1u);
        if (cnt[l] >= h->lcap[l]) return fail(SCORE_GM_ERR_CAPACITY, "Border front list full while placing nuclei");
        GmRec rec;
        memset(&rec, 0, sizeof(rec));
        rec.frac0 = 1.0;
        rec.s0 = 0;
        rec.inf = 26;
        rec.P = nucP[k];
        rec.cell = w;
        rec.defg = dom->cellgrid[w];
        rec.rxg = k;
        slots.push_back(h->lbase[l] + cnt[l]);
        recs.push_back(rec);
        cnt[l]++;
    }
    uint32_t *d_slots = nullptr;
    GmRec *d_recs = nullptr;
    int rc;
    if ((rc = dalloc(h, &d_slots, slots.size())) || (rc = dalloc(h, &d_recs, recs.size()))) {
        dfree(h, d_slots);
        return rc;
    }
    auto up = [&]() -> int {
        HIPCHK(hipMemcpyAsync(d.reg, h->reg.data(), sizeof(GmRegion) * d.nreg, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.L.base, h->lbase.data(), 4ull * (L + 1), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.L.cap, h->lcap.data(), 4ull * L, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.L.rcap, h->lrcap.data(), 4ull * L, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.L.rbase, h->lrbase.data(), 4ull * (L + 1), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.L.n, cnt.data(), 4ull * L, hipMemcpyHostToDevice, s));
        if (!slots.empty()) {
            HIPCHK(hipMemcpyAsync(d_slots, slots.data(), 4ull * slots.size(), hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(d_recs, recs.data(), sizeof(GmRec) * recs.size(), hipMemcpyHostToDevice, s));
        }
        HIPCHK(gm_launch_rg_place(d, d_slots, d_recs, (uint32_t)slots.size(), s));
        HIPCHK(gm_launch_rg_prefix(d, s));
        /* the host vectors must outlive the asynchronous copies */
        HIPCHK(hipStreamSynchronize(s));
        return 0;
    };
    rc = up();
    dfree(h, d_slots);
    dfree(h, d_recs);
    return rc;
}

/* ------------------------------------------------------------- physics scalars */

/* update_temperature (SCORE_Kernel.cpp:5318-5337) */
void update_temperature(score_gm_handle *h) {
    const size_t n = h->sched_t.size();
    size_t i = 0;
    while (i < n && h->sched_t[i] < h->t) i++;
    if (i == 0) { h->T = h->sched_T[0]; return; }
    if (i >= n) { h->T = h->sched_T[n - 1]; return; }
    h->T = h->sched_T[i - 1] +
           ((h->sched_T[i] - h->sched_T[i - 1]) / (h->sched_t[i] - h->sched_t[i - 1])) * (h->t - h->sched_t[i - 1]);
}

/* get_shearmodulus / get_burgersvector (SCORE_GeneralPhysics.cpp:65-90), update_atomisticproperties (SCORE_Kernel.cpp:5340-5352) */
void update_atomistic(score_gm_handle *h) {
    const double T = h->T;
    h->G = h->p.G0 * (-1.38e-9 * (T * T * T) + 1.21e-6 * (T * T) - 4.58e-4 * T + 1.00);
    h->b = h->p.b0;
    h->Gbhalfsq = 0.5 * h->G * (h->b * h->b);
}

/* update_intrinsicmobilities (SCORE_Kernel.cpp:5355-5368) */
void update_mobilities(score_gm_handle *h) {
    const double kT = kBoltzmann * h->T;
    h->mLAGB = h->p.lagb_m0 * exp(-1.0 * h->p.lagb_hact / kT);
    h->mHAGB = h->p.hagb_m0 * exp(-1.0 * h->p.hagb_hact / kT);
    h->mGS = h->p.gs_m0 * exp(-1.0 * h->p.gs_hact / kT);
    h->mRHHAGB = h->p.rh_hagb_m0 * exp(-1.0 * h->p.rh_hagb_hact / kT);
}

/* get_dtmax_instantslope_temp (SCORE_Kernel.cpp:5561-5591) */
double dt_temp(const score_gm_handle *h) {
    const size_t n = h->sched_t.size();
    size_t i = 0;
    while (i < n && h->sched_t[i] < h->t) i++;
    if (i == 0) return kInitialDeltaT;
    if (i < n) {
        double dTdt = (fabs(h->sched_T[i] - h->sched_T[i - 1]) / (h->sched_t[i] - h->sched_t[i - 1]));
        if (dTdt <= kSmallHeatRate) dTdt = kSmallHeatRate;
        double dtmax = (kSmallHeat / dTdt);
        const double dtintvl = h->sched_t[i] - h->sched_t[i - 1];
        if (dtmax >= dtintvl) dtmax = dtintvl;
        return dtmax;
    }
    return kInfinite;
}

/* update_microchemistry (SCORE_Kernel.cpp:5370-5405): f/r of the dispersoids,
 * constant (the first entry) or linear in time; past the table's end f/r keeps
 * its last value.  SCORE's prologue does not call it, so the first step runs
 * without drag. */
void update_microchemistry(score_gm_handle *h) {
    const int mode = h->p.zener_mode;
    if (mode == 0) {
        h->fr = 0.0;
        h->pzmin = 0.0;
        return;
    }
    if (mode == 1) {
        h->fr = h->disp_fr[0];
        h->pzmin = h->p.zener_alpha * h->p.zener_gamma * h->fr;
        return;
    }
    const size_t n = h->disp_t.size();
    size_t i = 0;
    while (i < n && h->disp_t[i] < h->t) i++;
    if (i == 0) h->fr = 0.0; /* t <= 0 (a negative time step): SCORE reads before its table here */
    else if (i < n) {
        const double dfrdt = ((h->disp_fr[i] - h->disp_fr[i - 1]) / (h->disp_t[i] - h->disp_t[i - 1]));
        h->fr = h->disp_fr[i - 1] + dfrdt * (h->t - h->disp_t[i - 1]);
    }
    h->pzmin = h->p.zener_alpha * h->p.zener_gamma * h->fr;
}

/* get_zener (SCORE_Kernel.cpp:5448-5456) */
double zener_force(const score_gm_handle *h) {
    if (h->p.zener_mode == 0) return 0.0;
    return (1.0 * h->p.zener_alpha * h->p.zener_gamma * h->fr);
}

/* get_dtmax_instantslope_zener (SCORE_Kernel.cpp:5639-5666) */
double dt_zener(const score_gm_handle *h) {
    if (h->p.zener_mode != 2) return kInfinite;
    const size_t n = h->disp_t.size();
    size_t i = 0;
    double dtmax = kInitialDeltaT;
    while (i < n && h->disp_t[i] < h->t) i++;
    if (i == 0) return dtmax;
    if (i < n) {
        double dpzdt = h->p.zener_alpha * h->p.zener_gamma *
                       (fabs(h->disp_fr[i] - h->disp_fr[i - 1]) / (h->disp_t[i] - h->disp_t[i - 1]));
        if (fabs(dpzdt) <= kSmallDraggingRate) dpzdt = kSmallDraggingRate;
        dtmax = (kSmallZenerForce / dpzdt);
        return dtmax;
    }
    return kInfinite;
}

/* update_deformedsubstructure (SCORE_Kernel.cpp:5408-5437): the recovery models
 * of SCORE_RecoveryModels.cpp, after the mobilities and before the new time step;
 * mode 1 uses the step just done (nes_networkgrowthmodel_vacancycorediff, :50-80),
 * mode 2 changes nothing (nes_networkgrowthmodel_solutedrag, :83-86), mode 3
 * scales rho0 (nes_michalakpaxton, :89-125) */
void update_deformedsubstructure(score_gm_handle *h) {
    const int mode = h->p.recovery_mode;
    if (mode == 4) {
        /* exponential recovery, not a SCORE model: rho = rho0 * exp(-t / tau) */
        const double factor = exp(-h->t / h->p.rec_tau);
        for (size_t g = 0; g < h->rho_host.size(); g++) h->rho_host[g] = h->rho0_host[g] * factor;
        h->rho_changed = true;
        return;
    }
    if (mode == 1) {
        double B = score_orient::kPi * h->p.rec_c3 * h->p.rec_c5 * h->p.rec_vacancy_d0 *
                   exp(-1.0 * h->p.rec_vacancy_hact / (kBoltzmann * h->T));
        B /= h->p.rec_kappa2;
        const double exponent = (h->b * h->b) / (kBoltzmann * h->T);
        const double expterm = exp(h->p.rec_alpha3 * h->p.rec_kappa2 * h->G * h->b * exponent);
        const double currdt = h->dt;
        for (double &rho : h->rho_host) {
            const double r_t = 1.0 / pow(rho, 0.5);
            const double Bx = (B / r_t) * expterm;
            const double r_tdt = (Bx * currdt) + r_t;
            rho = 1.0 / (r_tdt * r_tdt);
        }
        h->rho_changed = true;
    } else if (mode == 3) {
        const double f1 = 0.25, f2 = 0.75, tau2 = 0.3605, tau5 = 5.2824;
        const double currtime_minutes = h->t / 60.0;
        const double factor = f1 * pow((1 + (currtime_minutes / tau2)), -1.0) + f2 * pow((1 + (currtime_minutes / tau5)), -0.5);
        for (size_t g = 0; g < h->rho_host.size(); g++) h->rho_host[g] = h->rho0_host[g] * (factor * factor);
        h->rho_changed = true;
    }
}

/* get_dtmax_instantslope_rho (SCORE_Kernel.cpp:5594-5636) */
double dt_rho(const score_gm_handle *h) {
    const int mode = h->p.recovery_mode;
    double drdtmax, dtmax;
    if (mode == 1) {
        double Bx = score_orient::kPi * h->p.rec_c3 * h->p.rec_c5 * h->p.rec_vacancy_d0 *
                    exp(-1.0 * h->p.rec_vacancy_hact / (kBoltzmann * h->T));
        Bx = Bx / (h->p.rec_kappa2 * (1.0 / pow(h->rhomax, 0.5)));
        const double exponent = (h->b * h->b) / (kBoltzmann * h->T);
        drdtmax = Bx * exp(h->p.rec_alpha3 * h->p.rec_kappa2 * h->G * h->b * exponent);
        if (drdtmax < (1.0 / pow(kSmallRecoveryRate, 0.5))) drdtmax = (1.0 / pow(kSmallRecoveryRate, 0.5));
        dtmax = (1.0 / pow(kSmallRho, 0.5)) / drdtmax;
        return dtmax;
    }
    if (mode == 2) {
        const double Bsa = h->p.rec_c3 * h->b * h->p.rec_solute_d0 * exp(-1.0 * h->p.rec_solute_hact / (kBoltzmann * h->T));
        const double exponent = (h->b * h->b) / (kBoltzmann * h->T);
        const double ls = h->p.rec_solute_ls * pow(h->p.rec_solute_conc, (-2.0 / 3.0));
        drdtmax = Bsa * exp(h->p.rec_alpha3 * ls * h->G * h->b * exponent / (1.0 / (h->rhomax * h->rhomax)));
        if (drdtmax < (1.0 / pow(kSmallRecoveryRate, 0.5))) drdtmax = (1.0 / pow(kSmallRecoveryRate, 0.5));
        dtmax = (1.0 / pow(kSmallRho, 0.5)) / drdtmax;
        return dtmax;
    }
    return kInfinite;
}

/* get_dtmax_instantslope_cells (SCORE_Kernel.cpp:5519-5558) */
double dt_cells(const score_gm_handle *h) {
    const double kT = kBoltzmann * h->T;
    if (h->p.mobility_model == SCORE_GM_MOB_ROLLETTHOLM) {
        const double mmax = h->p.rh_hagb_m0 * exp(-1.0 * h->p.rh_hagb_hact / kT);
        const double vmax = mmax * ((h->Gbhalfsq * h->rhomax) - h->pzmin);
        return ((h->p.maxfill * h->cellsize) / vmax);
    }
    const double mgs = h->p.gs_m0 * exp(-1.0 * h->p.gs_hact / kT);
    const double mhagb = h->p.hagb_m0 * exp(-1.0 * h->p.hagb_hact / kT);
    double m_max = (h->Pmax * mgs) + ((1.0 - h->Pmax) * mhagb);
    if (h->Pmax < (0.0 - kDefaultSmallNumber)) m_max = h->p.lagb_m0 * exp(-1.0 * h->p.lagb_hact / kT);
    const double v_max = m_max * ((h->Gbhalfsq * h->rhomax) - h->pzmin);
    return ((h->p.maxfill * h->cellsize) / v_max);
}

/* get_dtmax_minimum (SCORE_Kernel.cpp:5687-5723); the nucleation rule is
 * INFINITE for site saturation */
double dt_minimum(const score_gm_handle *h) {
    double mn = kInfinite;
    const double a = dt_temp(h);
    if (a < mn) mn = a;
    const double r = dt_rho(h);
    if (r < mn) mn = r;
    const double z = dt_zener(h);
    if (z < mn) mn = z;
    const double c = dt_cells(h);
    if (c < mn) mn = c;
    return mn;
}

/* update_myrhomax (SCORE_RecoveryModels.cpp:146-165): grains sorted by rho
 * descending (update_mydgoptimize_rho, :128-135); the first one that still owns
 * cells gives rhomax, floored at RHOMAX_WELLANNEALED */
void update_rhomax(score_gm_handle *h) {
    while (h->rho_top < h->rho_order.size() && !h->alive[h->rho_order[h->rho_top]]) h->rho_top++;
    double r = kRhoWellAnnealed;
    if (h->rho_top < h->rho_order.size()) {
        const double v = h->rho_host[h->rho_order[h->rho_top]];
        if (v >= r) r = v;
    }
    h->rhomax = r;
}

/* termination test (SCORE_Kernel.cpp:6619); the main loop itself ends after
 * DEFAULT_NMAX iterations (:6615) */
bool done(const score_gm_handle *h) {
    return (h->Xcells >= (double)h->d.N) || (h->X > h->p.xmax) || ((int64_t)h->step > h->p.nmax) ||
           (h->t > h->tprocessingend) || (h->t > h->p.tmax) || ((int64_t)h->step >= kDefaultNmax);
}

uint32_t pow2_at_least(uint64_t v) {
    uint64_t c = 1;
    while (c < v) c <<= 1;
    return (uint32_t)std::min<uint64_t>(c, 1ull << 31);
}

/* nuclei arriving at the start of step k (after step 0) */
uint32_t arrivals_at(const score_gm_handle *h, uint32_t k) {
    return (k >= 1 && (size_t)k + 1 < h->arr_off.size()) ? h->arr_off[k + 1] - h->arr_off[k] : 0u;
}

/* per-step lists, tile masks and per-parent masks sized for the current
 * occupancy and the next step's arrivals (to either list) */
int ensure_step_capacity(score_gm_handle *h) {
    GmDev &d = h->d;
    const uint32_t na = arrivals_at(h, h->step);
    /* with several regions each list is swept in whole 8-tile chunks */
    const uint32_t tiles = d.nreg > 1 ? (h->nI + h->nB) / 64 + GM_SW_TPC * (d.nlists + 1)
                                      : gm_tiles_inside(h->nI + na) + (h->nB + na + 63) / 64 + GM_SW_TPC;
    if (tiles > d.tileCap) {
        uint32_t nc = std::max<uint32_t>(tiles, d.tileCap + d.tileCap / 2);
        dfree(h, d.mF); dfree(h, d.mR);
        d.mF = d.mR = nullptr;
        int rc;
        if ((rc = dalloc(h, &d.mF, nc)) || (rc = dalloc(h, &d.mR, nc))) return rc;
        d.tileCap = nc;
    }
    const uint32_t parents = h->nI + h->nB + na;
    if (parents > d.parentCap) {
        uint32_t nc = std::max<uint32_t>(parents, d.parentCap + d.parentCap / 2);
        dfree(h, d.pmI); dfree(h, d.pmB); dfree(h, d.pslot); dfree(h, d.par); dfree(h, d.pmH); dfree(h, d.plist);
        d.pmI = d.pmB = d.pslot = d.pmH = d.plist = nullptr;
        d.par = nullptr;
        d.parentCap = 0;
        int rc;
        if ((rc = dalloc(h, &d.pmI, nc)) || (rc = dalloc(h, &d.pmB, nc)) || (rc = dalloc(h, &d.pslot, nc)) ||
            (rc = dalloc(h, &d.par, (size_t)GM_PAR_PLANES * nc)) || (d.nreg > 1 && ((rc = dalloc(h, &d.pmH, nc)) || (rc = dalloc(h, &d.plist, nc)))))
            return rc;
        d.parentCap = nc;
    }
    return 0;
}

/* Evaluate the pending pairs with glibc and store them in the device cache. */
int resolve_pending(score_gm_handle *h, const GmStats &st) {
    const uint32_t npend = st.npend;
    if (npend == 0) return 0;
    GmDev &d = h->d;
    /* the keys: inline in the step's stats, or copied when there are more */
    const unsigned long long *keys = st.pend_keys;
    if (npend > GM_PEND_INLINE) {
        h->pend_keys.resize(npend);
        HIPCHK(hipMemcpyAsync(h->pend_keys.data(), d.pend_key, sizeof(unsigned long long) * npend,
                              hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        keys = h->pend_keys.data();
    }
    if (npend > h->pend_vals_cap) {
        /* the staging buffer is idle here: the stream was synchronised after the step */
        const uint32_t nc = std::max<uint32_t>(std::max<uint32_t>(npend, 2 * h->pend_vals_cap), 1024);
        if (h->hpend_vals) (void)hipHostFree(h->hpend_vals);
        h->hpend_vals = nullptr;
        h->hpend_dev = nullptr;
        h->pend_vals_cap = 0;
        int rc = pinned_mapped(h, (void **)&h->hpend_vals, (void **)&h->hpend_dev, sizeof(double) * nc);
        if (rc) return rc;
        h->pend_vals_cap = nc;
    }
    for (uint32_t i = 0; i < npend; i++) {
        const unsigned long long key = keys[i] - 1ull;
        const uint32_t rxg = (uint32_t)(key / d.ndefg), g = (uint32_t)(key % d.ndefg);
        const double P = mobility_weight(h->p, &h->rxq[4 * (size_t)rxg], &h->defq[4 * (size_t)g]);
        h->hpend_vals[i] = P;
        h->pend_P.push_back(P); /* folded into myMobilityWeightMax at the step that met the pair */
    }
    /* stream-ordered: the kernel reads the values from the mapped buffer and
     * stores them into the cache before the next step's kernels read it; the
     * buffer is not rewritten before the next window's sync */
    HIPCHK(gm_launch_resolve_pending(d, h->hpend_dev, npend, h->stream));
    h->hcount += npend;
    return 0;
}

int step_error(uint32_t err);
int list_lengths(score_gm_handle *h, std::vector<uint32_t> &n);

/* Grow the cache (doubling) until it is at most half full with `room` more
 * entries: markers in the front are resolved first, then every entry is
 * re-inserted into the larger table. */
int grow_pcache(score_gm_handle *h, uint64_t room = 0) {
    GmDev &d = h->d;
    if (((uint64_t)h->hcount + room) * 2 <= d.hcap) return 0;
    const GmFrontOut none{};
    if (d.nreg > 1) {
        std::vector<uint32_t> ln;
        const int rc = list_lengths(h, ln);
        if (rc) return rc;
        for (uint32_t l = 0; l < d.nlists; l++) HIPCHK(gm_launch_materialize(d, h->step, h->lbase[l], ln[l], none, h->stream));
    } else {
        HIPCHK(gm_launch_materialize(d, h->step, 0, h->nI, none, h->stream));
        HIPCHK(gm_launch_materialize(d, h->step, d.capI, h->nB, none, h->stream));
    }
    std::vector<unsigned long long> keys(d.hcap);
    std::vector<double> vals(d.hcap);
    HIPCHK(hipMemcpyAsync(keys.data(), d.hkeys, sizeof(unsigned long long) * d.hcap, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipMemcpyAsync(vals.data(), d.hvals, sizeof(double) * d.hcap, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    std::vector<unsigned long long> k2;
    std::vector<double> v2;
    for (uint32_t i = 0; i < d.hcap; i++)
        if (keys[i]) { k2.push_back(keys[i]); v2.push_back(vals[i]); }
    uint64_t ncap64 = (uint64_t)d.hcap * 2u;
    while (((uint64_t)h->hcount + room) * 2 > ncap64) ncap64 *= 2u;
    if (ncap64 > (1ull << 31)) return fail(SCORE_GM_ERR_OOM, "mobility-weight cache beyond 2^31 entries");
    const uint32_t ncap = (uint32_t)ncap64;
    dfree(h, d.hkeys); dfree(h, d.hvals); dfree(h, d.pend_slot); dfree(h, d.pend_key);
    d.hkeys = nullptr; d.hvals = nullptr; d.pend_slot = nullptr; d.pend_key = nullptr;
    int rc;
    if ((rc = dalloc(h, &d.hkeys, ncap)) || (rc = dalloc(h, &d.hvals, ncap)) || (rc = dalloc(h, &d.pend_slot, ncap)) ||
        (rc = dalloc(h, &d.pend_key, ncap)))
        return rc;
    d.hcap = ncap;
    HIPCHK(gm_launch_fill_u64(d.hkeys, 0ull, ncap, h->stream));
    HIPCHK(gm_launch_fill_u64((unsigned long long *)d.hvals, GM_PENDING_TAG, ncap, h->stream));
    unsigned long long *dk = nullptr;
    double *dv = nullptr;
    if ((rc = dalloc(h, &dk, k2.size())) || (rc = dalloc(h, &dv, v2.size()))) { dfree(h, dk); return rc; }
    HIPCHK(hipMemcpyAsync(dk, k2.data(), sizeof(unsigned long long) * k2.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(hipMemcpyAsync(dv, v2.data(), sizeof(double) * v2.size(), hipMemcpyHostToDevice, h->stream));
    HIPCHK(gm_launch_hash_insert(d, dk, dv, (uint32_t)k2.size(), h->stream));
    HIPCHK(gm_launch_stats_out(d, h->hst_dev, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    dfree(h, dk);
    dfree(h, dv);
    if (h->hst->err & GM_ERRF_PCACHE_FULL) return h->err = step_error(h->hst->err);
    return 0;
}

/* The window's weight lookups spilled (no room near a key's home in the cache):
 * grow the cache for them, let the spilled cells look their pairs up again and
 * resolve the new pairs with glibc, until none spills.  Runs after the window's
 * own pending weights are resolved and before its deferred cells are predicted.
 * The stats mirror is rewritten here; the caller keeps what it needs of it. */
int resolve_spilled(score_gm_handle *h, uint32_t nspilled, uint32_t ndeferred) {
    GmDev &d = h->d;
    for (int round = 0; nspilled; round++) {
        if (round == 8) return fail(SCORE_GM_ERR_STATE, "internal error: weight lookups keep spilling");
        int rc = grow_pcache(h, nspilled);
        if (rc) return rc;
        HIPCHK(gm_launch_reinsert_spilled(d, h->pp, ndeferred, h->nI, h->nB, h->stream));
        HIPCHK(gm_launch_stats_out(d, h->hst_dev, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        if ((rc = resolve_pending(h, *h->hst))) return rc;
        nspilled = h->hst->novf;
    }
    return grow_pcache(h);
}

/* with several regions: the lists' lengths from the device, and back with their
 * prefixes and sums (d.st->nI, d.st->nB, mirrored into h->nI, h->nB) */
int list_lengths(score_gm_handle *h, std::vector<uint32_t> &n) {
    n.assign(h->d.nlists, 0u);
    HIPCHK(hipMemcpyAsync(n.data(), h->d.L.n, 4ull * h->d.nlists, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int set_list_lengths(score_gm_handle *h, const std::vector<uint32_t> &n) {
    HIPCHK(hipMemcpyAsync(h->d.L.n, n.data(), 4ull * h->d.nlists, hipMemcpyHostToDevice, h->stream));
    HIPCHK(gm_launch_rg_prefix(h->d, h->stream));
    HIPCHK(gm_launch_stats_out(h->d, h->hst_dev, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    h->nI = h->hst->nI;
    h->nB = h->hst->nB;
    return 0;
}

/* ompshar_defragmentation (SCORE_Kernel.cpp:3877-3934) of the list whose slot
 * 0 is global slot `base`: capacity `cap`, nextSlotNeverActive `nnever`. */
int defragment_list(score_gm_handle *h, uint32_t base, uint32_t cap, uint32_t &nnever) {
    GmDev &d = h->d;
    hipStream_t s = h->stream;
    /* the first INACTIVE slot; the reference's search runs on past nnever while
     * it meets ACTIVE records left behind by an earlier defragmentation */
    HIPCHK(gm_launch_defrag_first_inactive(d, base, 0, nnever, h->step, s));
    HIPCHK(gm_launch_stats_out(d, h->hst_dev, s));
    HIPCHK(hipStreamSynchronize(s));
    uint32_t first = h->hst->dfirst;
    if (first >= nnever && nnever < cap) {
        HIPCHK(gm_launch_defrag_first_inactive(d, base, nnever, cap, h->step, s));
        HIPCHK(gm_launch_stats_out(d, h->hst_dev, s));
        HIPCHK(hipStreamSynchronize(s));
        first = h->hst->dfirst;
    }
    if (first + 1u >= nnever) {
        nnever = first;
        return 0;
    }
    const uint32_t n = nnever - first - 1u;
    GmRec *trec = nullptr;
    gm_cs_t *tcs = nullptr;
    int rc;
    if ((rc = dalloc(h, &trec, n)) || (rc = dalloc(h, &tcs, n))) {
        dfree(h, trec);
        return rc;
    }
    hipError_t e;
    if ((e = gm_launch_defrag_move(d, base, first, nnever, h->step, trec, tcs, s)) != hipSuccess ||
        (e = gm_launch_stats_out(d, h->hst_dev, s)) != hipSuccess ||
        (e = hipStreamSynchronize(s)) != hipSuccess) {
        dfree(h, trec);
        dfree(h, tcs);
        return hip_fail(e, "defragmentation");
    }
    dfree(h, trec);
    dfree(h, tcs);
    nnever = first + h->hst->dcount;
    return 0;
}

/* the per-step scalars the sweep uses, from the handle's current state */
GmSweepParams sweep_params(const score_gm_handle *h) {
    GmSweepParams sp;
    memset(&sp, 0, sizeof(sp));
    sp.model = h->p.mobility_model;
    sp.mRH = h->mRHHAGB; sp.mLAGB = h->mLAGB; sp.mHAGB = h->mHAGB; sp.mGS = h->mGS;
    sp.Gbhalfsq = h->Gbhalfsq; sp.zener = zener_force(h); sp.dt = h->dt; sp.inv_dx = h->inv_dx;
    return sp;
}

/* t += dt and the scalar updates of the master section (SCORE_Kernel.cpp:6725-6734) */
void advance_scalars(score_gm_handle *h) {
    h->t += h->dt;
    update_temperature(h);
    update_atomistic(h);
    update_microchemistry(h);
    update_mobilities(h);
    update_deformedsubstructure(h);
    h->dt = dt_minimum(h);
}

/* Steps that may follow the current one in a launch window: the termination
 * tests on t and the step count must not fire and the scalars must stay
 * bit-identical, found by running the scalar updates ahead on a copy; and a
 * cell infected in the window must not be able to complete before its
 * deferred prediction at the window's end.  Only the Rollett-Holm model
 * qualifies: its time step does not depend on the weights of new pairs. */
uint32_t window_length(score_gm_handle *h, uint32_t wmax, const GmSweepParams &sp) {
    if (h->prof || wmax <= 1 || h->p.mobility_model != SCORE_GM_MOB_ROLLETTHOLM) return 1;
    /* recovery that changes rho re-anchors every step */
    if (h->p.recovery_mode == 1 || h->p.recovery_mode == 3 || h->p.recovery_mode == 4) return 1;
    /* a new cell starts at most at 0.0682 times the largest increment and takes
     * at least imin additions to reach 1.0; the increment is at most maxfill
     * because the time step is bounded by maxfill*dx/vmax and P <= 1 */
    const double incmax = h->p.maxfill * (1.0 + 1e-9);
    const double imin = ceil((1.0 - 0.0682 * incmax) / incmax);
    uint32_t w = std::min<uint32_t>(wmax, GM_WMAX);
    if (imin - 1.0 < (double)w) w = imin >= 2.0 ? (uint32_t)(imin - 1.0) : 1u;
    if (w <= 1) return 1;
    const double t0 = h->t, T0 = h->T, G0 = h->G, b0 = h->b, gb0 = h->Gbhalfsq, dt0 = h->dt, pz0 = h->pzmin, fr0 = h->fr;
    const double m0 = h->mLAGB, m1 = h->mHAGB, m2 = h->mGS, m3 = h->mRHHAGB;
    const uint32_t step0 = h->step;
    uint32_t n = 1;
    while (n < w) {
        advance_scalars(h);
        h->step++;
        /* a step at a multiple of the horizon re-anchors, so it starts a window */
        if (done(h) || !same_scalars(sweep_params(h), sp) || h->step % h->d.horizon == 0) break;
        n++;
    }
    h->t = t0; h->T = T0; h->G = G0; h->b = b0; h->Gbhalfsq = gb0; h->dt = dt0; h->pzmin = pz0; h->fr = fr0;
    h->mLAGB = m0; h->mHAGB = m1; h->mGS = m2; h->mRHHAGB = m3;
    h->step = step0;
    return n;
}

/* the smallest Xcells at which the host acts: X > XMAX, Xcells >= N, or the next
 * defragmentation threshold */
double window_xstop(const score_gm_handle *h) {
    const double N = (double)h->d.N;
    double c = floor(h->p.xmax * N);
    if (c < 0.0) c = 0.0;
    while (c > 0.0 && (c / N) > h->p.xmax) c -= 1.0;
    while (!((c / N) > h->p.xmax) && c <= N) c += 1.0;
    double x = std::min(c, N);
    if (h->defrag_cnt < h->defrag_at.size()) x = std::min(x, h->defrag_at[h->defrag_cnt]);
    if (h->glog_cnt < h->glog_at.size()) x = std::min(x, h->glog_at[h->glog_cnt]);
    return x;
}

int step_error(uint32_t err) {
    if (err & GM_ERRF_BORDER_FULL)
        return fail(SCORE_GM_ERR_CAPACITY, "Border front list full (SCORE would write out of bounds)");
    if (err & GM_ERRF_INSIDE_FULL) return fail(SCORE_GM_ERR_CAPACITY, "Inside front list full");
    if (err & GM_ERRF_MISSED) return fail(SCORE_GM_ERR_STATE, "internal error: missed completion");
    if (err & GM_ERRF_KEYS)
        return fail(SCORE_GM_ERR_UNSUPPORTED,
                    "more cells completed in one step than 32-bit infection keys hold (2^27; 2^23 with several regions)");
    if (err & GM_ERRF_RLOG_FULL) return fail(SCORE_GM_ERR_CAPACITY, "retirement log full (cells retired twice)");
    if (err & GM_ERRF_HALO) return fail(SCORE_GM_ERR_STATE, "internal error: an infection into another region found no halo");
    return fail(SCORE_GM_ERR_STATE, "internal error: mobility-weight cache entry lost while growing it");
}

/* log_grainevo (SCORE_OnTheFlyCharacterization.cpp:478-530): time before t
 * advances, X, the log counter, step, Sv and the cell count of every grain
 * (deformed grains first), read from the device after the step */
int log_grainevo(score_gm_handle *h, uint32_t sv) {
    GmDev &d = h->d;
    const size_t ng = (size_t)d.ndefg + d.nrxg;
    const size_t off = h->glog_counts.size();
    h->glog_counts.resize(off + ng);
    HIPCHK(hipMemcpyAsync(h->glog_counts.data() + off, d.defcnt, sizeof(uint32_t) * d.ndefg, hipMemcpyDeviceToHost,
                          h->stream));
    if (d.nrxg)
        HIPCHK(hipMemcpyAsync(h->glog_counts.data() + off + d.ndefg, d.rxcnt, sizeof(uint32_t) * d.nrxg,
                              hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    score_gm_grainlog_row g;
    memset(&g, 0, sizeof(g));
    g.t = h->t;
    g.X = h->X;
    g.logcnt = h->glog_cnt;
    g.step = h->step;
    g.sv = sv;
    g.ngrains = (uint32_t)ng;
    h->glog.push_back(g);
    h->glog_cnt++;
    return 0;
}

/* Iterations of SCORE's main loop (SCORE_Kernel.cpp:6615-6747), up to wmax of
 * them launched back to back as one window.  Returns 1 when the termination
 * test fired before any step, 0 after *nsteps >= 1 steps, < 0 on error. */
/* steps queued ahead of the device within a window */
#define GM_LEAD 1u

/* Waits until the device has finished `need` steps of the window or halted it,
 * reading the progress word k_step_tail writes into mapped host memory; false
 * when the window halted.  After a second without getting there it stops
 * waiting (the window's end reports whatever held the device up). */
bool wait_steps(score_gm_handle *h, uint32_t need) {
    const auto t0 = std::chrono::steady_clock::now();
    for (uint32_t it = 1;; it++) {
        const uint32_t v = __atomic_load_n(h->hprog, __ATOMIC_ACQUIRE);
        if (v & 1u) return false;
        if ((v >> 1) >= need) return true;
        if ((it & 63u) == 0u) {
            std::this_thread::yield();
            if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(1)) return true;
        } else {
            __builtin_ia32_pause();
        }
    }
}

int step_window(score_gm_handle *h, uint32_t wmax, uint32_t *nsteps) {
    *nsteps = 0;
    if (h->err) return h->err;
    if (h->completed) return 1;
    if (done(h)) { h->completed = 1; return 1; }
    GmDev &d = h->d;
    int rc;
    h->pend_P.clear();
    /* on-the-fly defragmentation at the Xcells thresholds (SCORE_Kernel.cpp:6625-6630) */
    if (h->defrag_cnt < h->defrag_at.size() && h->Xcells >= h->defrag_at[h->defrag_cnt]) {
        if (d.nreg > 1) {
            /* every region defragments both its lists (the threshold test is the same for all) */
            std::vector<uint32_t> n;
            if ((rc = list_lengths(h, n))) return h->err = rc;
            for (uint32_t l = 0; l < d.nlists; l++)
                if ((rc = defragment_list(h, h->lbase[l], h->lcap[l], n[l]))) return h->err = rc;
            if ((rc = set_list_lengths(h, n))) return h->err = rc;
        } else {
            if ((rc = defragment_list(h, 0u, d.capI, h->nI)) || (rc = defragment_list(h, d.capI, d.capB, h->nB)))
                return h->err = rc;
            /* the device keeps the list lengths in its stats */
            hipError_t e2;
            if ((e2 = gm_launch_fill_u32(&d.st->nI, h->nI, 1, h->stream)) != hipSuccess ||
                (e2 = gm_launch_fill_u32(&d.st->nB, h->nB, 1, h->stream)) != hipSuccess)
                return h->err = hip_fail(e2, "defragmentation");
        }
        h->defrag_cnt++;
        h->force_reanchor = true;
    }
    /* every cell is predicted again at each multiple of the horizon, so that no
     * stored completion step lies a horizon or more ahead */
    if (h->step > 0 && h->step % d.horizon == 0) h->force_reanchor = true;
    rc = ensure_step_capacity(h);
    if (rc) return h->err = rc;
    GmSweepArgs a;
    memset(&a, 0, sizeof(a));
    a.sp = sweep_params(h);
    const GmSweepParams sp = a.sp;
    a.k = h->step;
    a.reanchor = (h->force_reanchor || !h->have_last_sp || !same_scalars(sp, h->last_sp)) ? 1u : 0u;
    a.stepcap = (uint32_t)std::min<int64_t>(std::max<int64_t>(h->p.nmax, 0), (int64_t)GM_STEP_MAX - 1);
    /* recovery that changes rho re-anchors every cell at every step, so a
     * prediction only has to tell whether the cell completes in the step itself */
    a.next_only = (h->p.recovery_mode == 1 || h->p.recovery_mode == 3 || h->p.recovery_mode == 4) ? 1u : 0u;
    const bool prof = h->prof;
    const uint32_t slots0 = h->nI + h->nB;
    const uint32_t W = window_length(h, wmax, sp);
    /* grids: exact for one step, with room for growth over a window (the kernels
     * take the lengths from the device and cover any grid) */
    const uint64_t capIe = d.nreg > 1 ? 0x7FFFFFFFull : d.capI, capBe = d.nreg > 1 ? 0x7FFFFFFFull : d.capB;
    const uint32_t nIe = W == 1 ? h->nI : (uint32_t)std::min<uint64_t>(capIe, h->nI + h->nI / 2 + 4096u);
    const uint32_t nBe = W == 1 ? h->nB : (uint32_t)std::min<uint64_t>(capBe, h->nB + h->nB / 2 + 4096u);
    a.nI = nIe;
    a.nB = nBe;
    GmWindow w;
    memset(&w, 0, sizeof(w));
    w.xstop = window_xstop(h);
    w.top_grain = h->rho_top < h->rho_order.size() ? h->rho_order[h->rho_top] : 0xFFFFFFFFu;
    w.pend_limit = (uint32_t)std::min<uint64_t>(0xFFFFFFFFull, (uint64_t)d.hcap / 2 > h->hcount ? (uint64_t)d.hcap / 2 - h->hcount : 0ull);
    w.defer = W > 1 ? 1u : 0u;
    hipEvent_t *ev = prof ? h->ev : nullptr;
    const auto th0 = std::chrono::steady_clock::now();
    hipError_t e;
    /* The steps are launched GM_LEAD ahead of the device: before step j is
     * queued the device has finished j - GM_LEAD of them, so that after a step
     * that halts the window (later steps would return at once) no more are
     * queued.  Steps not launched are exactly those that would have done
     * nothing, whatever the host sees in time. */
    const bool paced = !prof && W > GM_LEAD + 1;
    if (paced) __atomic_store_n(h->hprog, 0u, __ATOMIC_RELAXED);
    if ((e = gm_launch_window_begin(d, h->step, h->Xcells, h->stream)) != hipSuccess)
        return h->err = hip_fail(e, "growth step");
    for (uint32_t j = 0; j < W; j++) {
        if (paced && j > GM_LEAD && !wait_steps(h, j - GM_LEAD)) break;
        GmSweepArgs aj = a;
        if (j) aj.reanchor = 0;
        if (d.nreg > 1) {
            if ((e = gm_launch_rg_sweep(d, aj, std::max(slots0, nIe + nBe), h->stream, ev)) != hipSuccess ||
                (e = gm_launch_rg_infect(d, h->pp, aj, std::max(slots0, nIe + nBe), h->stream, ev)) != hipSuccess ||
                (e = gm_launch_step_tail(d, w, h->stream)) != hipSuccess)
                return h->err = hip_fail(e, "growth step");
            continue;
        }
        if ((e = gm_launch_sweep(d, aj, h->stream, ev)) != hipSuccess ||
            (e = gm_launch_infect(d, h->pp, aj, std::max(slots0, nIe + nBe), h->stream, ev)) != hipSuccess ||
            (e = gm_launch_step_tail(d, w, h->stream)) != hipSuccess)
            return h->err = hip_fail(e, "growth step");
    }
    if ((e = gm_launch_stats_out(d, h->hst_dev, h->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(h->stream)) != hipSuccess)
        return h->err = hip_fail(e, "growth step");
    h->last_sp = sp;
    h->have_last_sp = true;
    h->force_reanchor = false;
    /* read in place: only the rows and inline ids the window filled are touched,
     * and nothing rewrites the mirror before the next window's end */
    const GmStats &st = *h->hst;
    const uint32_t nexec = st.nexec;
    if (st.err) return h->err = step_error(st.err);
    if (nexec < 1 || nexec > W) return h->err = fail(SCORE_GM_ERR_STATE, "internal error: window step count");
    if (prof) {
        /* one step per window when profiling */
        const GmStepRow &r0 = st.rows[0];
        score_gm_profile &pf = h->profc;
        for (int i = 0; i < GM_NK; i++) {
            float ms;
            if ((e = hipEventElapsedTime(&ms, h->ev[i], h->ev[i + 1])) != hipSuccess)
                return h->err = hip_fail(e, "hipEventElapsedTime");
            pf.ms_kernel[i] += ms;
        }
        float ms;
        if ((e = hipEventElapsedTime(&ms, h->ev[0], h->ev[GM_NK])) != hipSuccess)
            return h->err = hip_fail(e, "hipEventElapsedTime");
        const uint32_t sv0 = slots0 + r0.nplaced - (r0.totRecI + r0.totRecB);
        pf.ms_step += ms;
        pf.steps++;
        pf.sum_sv += sv0;
        pf.slots_scanned += slots0;
        pf.n_full += (uint64_t)r0.nfullI + r0.nfullB;
        pf.n_recycled += (uint64_t)std::min(r0.winI, r0.recLenI) + std::min(r0.winB, r0.recLenB);
        pf.n_attempts += r0.nattempt;
        pf.n_new += (uint64_t)r0.winI + r0.winB;
        /* increments computed when the cells were infected (re-anchored cells are counted separately) */
        pf.n_predicted += (uint64_t)r0.winI + r0.winB;
        if (a.reanchor) {
            pf.n_reanchor_steps++;
            pf.n_reanchored += sv0;
        }
        pf.pending_resolved += r0.npend;
    }
    /* the window's new pairs, evaluated with glibc and stored in the device cache */
    if ((rc = resolve_pending(h, st))) return h->err = rc;
    {
        const GmStepRow &last = st.rows[nexec - 1];
        h->nI = last.nI;
        h->nB = last.nB;
    }
    if ((rc = grow_pcache(h))) return h->err = rc;
    if (st.novf) {
        /* the window ended after the step whose lookups spilled; the deferred
         * count and the rows stay as the window left them on the device */
        const uint32_t ndeferred = st.ndeferred;
        if ((rc = resolve_spilled(h, st.novf, ndeferred))) return h->err = rc;
        if (h->hst->ndeferred != ndeferred) return h->err = fail(SCORE_GM_ERR_STATE, "internal error: deferred cells");
    }
    if (st.ndeferred) {
        /* cells that waited for a weight: predicted under the window's scalars at
         * the step after the window, queued before the host's master section so
         * that the device works while the host does */
        GmSweepArgs ad = a;
        ad.sp = sp;
        ad.k = h->step + nexec;
        ad.nI = h->nI;
        ad.nB = h->nB;
        if ((e = gm_launch_predict_deferred(d, ad, st.ndeferred, h->stream)) != hipSuccess)
            return h->err = hip_fail(e, "deferred predictions");
    }
    /* master section of every step (SCORE_Kernel.cpp:6688-6742) */
    uint32_t prev_slots = slots0;
    size_t pend_off = 0;
    for (uint32_t j = 0; j < nexec; j++) {
        const GmStepRow &row = st.rows[j];
        const uint32_t nfull = row.nfullI + row.nfullB;
        /* the sweep saw the slots occupied before the step and the nuclei placed at its start */
        const uint32_t sv = prev_slots + row.nplaced - (row.totRecI + row.totRecB);
        prev_slots = row.nI + row.nB;
        score_gm_traj_row r;
        memset(&r, 0, sizeof(r));
        r.step = h->step;
        r.t = h->t;
        r.dt = h->dt;
        r.temperature = h->T;
        /* reg_myMobilityWeightMax of the step: the nuclei placed at its start, the
         * new pairs its infections met (short-circuited to 1.0 or resolved with
         * glibc; the pairs a spilled step met are resolved last) */
        h->nuclei += row.nplaced;
        if (row.pkey) {
            const double Pn = gm_pkey_value(row.pkey);
            if (Pn >= h->regPmax) h->regPmax = Pn;
        }
        if (row.pone && 1.0 >= h->regPmax) h->regPmax = 1.0;
        const size_t pend_end = j + 1 == nexec ? h->pend_P.size() : std::min(h->pend_P.size(), pend_off + row.npend);
        for (size_t i = pend_off; i < pend_end; i++)
            if (h->pend_P[i] >= h->regPmax) h->regPmax = h->pend_P[i];
        pend_off = pend_end;
        double mymobmax = kDefaultPmax;
        if (h->regPmax >= mymobmax) mymobmax = h->regPmax;
        h->Pmax = mymobmax;
        h->Xcells = h->Xcells + ((double)nfull);
        h->X = (h->Xcells / ((double)d.N));
        h->sum_sv += sv;
        if (j + 1 == nexec) {
            /* grains that lost their last cell leave the rhomax search; before the
             * last step of a window none of them was the one that sets rhomax */
            if (st.nzeroed) {
                const uint32_t *ids = st.zeroed_ids;
                if (st.nzeroed > GM_ZEROED_INLINE) {
                    h->zbuf.resize(st.nzeroed);
                    if ((e = hipMemcpyAsync(h->zbuf.data(), d.zlist, sizeof(uint32_t) * std::min(st.nzeroed, d.ndefg),
                                            hipMemcpyDeviceToHost, h->stream)) != hipSuccess ||
                        (e = hipStreamSynchronize(h->stream)) != hipSuccess)
                        return h->err = hip_fail(e, "exhausted grains");
                    ids = h->zbuf.data();
                }
                for (uint32_t i = 0; i < std::min(st.nzeroed, d.ndefg); i++) h->alive[ids[i]] = 0;
            }
            if (st.cwrap) {
                /* a count went below zero and wrapped: cellcount > 0 again, as in SCORE */
                std::vector<uint32_t> cnt(d.ndefg);
                if ((e = hipMemcpyAsync(cnt.data(), d.defcnt, sizeof(uint32_t) * d.ndefg, hipMemcpyDeviceToHost,
                                        h->stream)) != hipSuccess ||
                    (e = hipStreamSynchronize(h->stream)) != hipSuccess)
                    return h->err = hip_fail(e, "deformed grain counts");
                for (uint32_t g = 0; g < d.ndefg; g++) h->alive[g] = cnt[g] > 0 ? 1 : 0;
                h->rho_top = 0;
            }
        }
        if (h->p.recovery_mode != 0) {
            /* update_mydgoptimize_rho (SCORE_RecoveryModels.cpp:128-135): re-sorted by the
             * current rho; a stable sort leaves an ordered list as it is, so an
             * order that recovery kept (every rho scaled alike) is not sorted again */
            auto desc = [&](uint32_t a, uint32_t b) { return h->rho_host[a] > h->rho_host[b]; };
            if (!std::is_sorted(h->rho_order.begin(), h->rho_order.end(), desc))
                std::stable_sort(h->rho_order.begin(), h->rho_order.end(), desc);
            h->rho_top = 0;
        }
        update_rhomax(h);
        r.sv = sv;
        r.nfull_inside = row.nfullI;
        r.nfull_border = row.nfullB;
        r.nslots_inside = row.nI;
        r.nslots_border = row.nB;
        r.ninfected = row.winI + row.winB;
        r.npending = row.npend;
        r.X = h->X;
        r.xcells = h->Xcells;
        r.rhomax = h->rhomax;
        r.pmax = h->Pmax;
        r.nplaced = row.nplaced;
        /* log_OUTPUT_Diehl (SCORE_OnTheFlyCharacterization.cpp:646-651): the window
         * ends at the step whose Xcells reaches the next log point, so the device's
         * counts are this step's */
        if (h->glog_cnt < h->glog_at.size() && h->Xcells >= h->glog_at[h->glog_cnt]) {
            if (j + 1 != nexec) return h->err = fail(SCORE_GM_ERR_STATE, "internal error: grain log inside a window");
            if ((rc = log_grainevo(h, sv))) return h->err = rc;
        }
        advance_scalars(h);
        h->step++;
        h->traj.push_back(r);
        if (j + 1 < nexec && !same_scalars(sweep_params(h), sp))
            return h->err = fail(SCORE_GM_ERR_STATE, "internal error: scalars changed inside a launch window");
    }
    if (h->rho_changed) {
        /* recovered dislocation densities for the next step (after the
         * window's own predictions); every cell's increment changes, so it re-anchors */
        if ((e = hipMemcpyAsync(d.rho, h->rho_host.data(), sizeof(double) * d.ndefg, hipMemcpyHostToDevice,
                                h->stream)) != hipSuccess)
            return h->err = hip_fail(e, "recovered dislocation densities");
        h->rho_changed = false;
        h->force_reanchor = true;
    }
    *nsteps = nexec;
    h->windows++;
    if (prof) h->profc.ms_host += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - th0).count();
    return 0;
}

}  // namespace

/* ------------------------------------------------------------------- ABI */

extern "C" {

int score_gm_abi_version(void) { return SCORE_GM_ABI_VERSION; }

const char *score_gm_strerror(int status) {
    switch (status) {
        case SCORE_GM_OK: return "ok";
        case SCORE_GM_ERR_ARG: return g_last_error.empty() ? "invalid argument" : g_last_error.c_str();
        case SCORE_GM_ERR_OOM: return g_last_error.empty() ? "out of memory" : g_last_error.c_str();
        case SCORE_GM_ERR_HIP: return g_last_error.empty() ? "HIP error" : g_last_error.c_str();
        case SCORE_GM_ERR_CAPACITY: return g_last_error.empty() ? "front list capacity exceeded" : g_last_error.c_str();
        case SCORE_GM_ERR_UNSUPPORTED: return g_last_error.empty() ? "unsupported option" : g_last_error.c_str();
        case SCORE_GM_ERR_STATE: return g_last_error.empty() ? "invalid state" : g_last_error.c_str();
        case SCORE_GM_ERR_NODEVICE: return "no usable gfx950 device";
        default: return "unknown status";
    }
}

int score_gm_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    int usable = 0;
    for (int i = 0; i < n; i++) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, i) == hipSuccess && strncmp(prop.gcnArchName, "gfx950", 6) == 0) usable++;
    }
    return usable;
}

int score_gm_create(const score_gm_domain *dom, const score_gm_params *prm, int device, score_gm_handle **out) {
    if (!out) return fail(SCORE_GM_ERR_ARG, "null output handle");
    *out = nullptr;
    int rc = validate(dom, prm);
    if (rc) return rc;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        (void)hipGetLastError();
        return fail(SCORE_GM_ERR_NODEVICE, "no HIP device");
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess || strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(SCORE_GM_ERR_NODEVICE, "device is not gfx950");
    HIPCHK(hipSetDevice(device));

    score_gm_handle *h = new (std::nothrow) score_gm_handle();
    if (!h) return fail(SCORE_GM_ERR_OOM, "handle");
    h->device = device;
    h->p = *prm;
    h->sched_t.assign(prm->sched_time, prm->sched_time + prm->n_sched);
    h->sched_T.assign(prm->sched_temp, prm->sched_temp + prm->n_sched);
    h->p.sched_time = h->sched_t.data();
    h->p.sched_temp = h->sched_T.data();
    if (prm->n_defrag) h->defrag_at.assign(prm->defrag_at, prm->defrag_at + prm->n_defrag);
    h->p.defrag_at = h->defrag_at.data();
    if (prm->zener_mode != 0) {
        h->disp_t.assign(prm->disp_time, prm->disp_time + prm->n_disp);
        h->disp_fr.assign(prm->disp_fr, prm->disp_fr + prm->n_disp);
    }
    h->p.disp_time = h->disp_t.data();
    h->p.disp_fr = h->disp_fr.data();
    if (prm->n_grainlog) h->glog_at.assign(prm->grainlog_at, prm->grainlog_at + prm->n_grainlog);
    h->p.grainlog_at = h->glog_at.data();
    h->cellsize = dom->cellsize;
    h->inv_dx = (1.0 / dom->cellsize);
    /* init_processing (SCORE_Kernel.cpp:5888-5901) */
    h->tprocessingend = 0.0;
    for (double tt : h->sched_t)
        if (tt >= h->tprocessingend) h->tprocessingend = tt;

    GmDev &d = h->d;
    d.nx = dom->nx; d.ny = dom->ny; d.nz = dom->nz; d.nxy = dom->nx * dom->ny;
    d.N = (uint64_t)dom->nx * dom->ny * dom->nz;
    d.ndefg = dom->n_defg; d.nrxg = dom->n_rxg;
    /* list capacities (ompshar_init_myRXFrontInside/Border, SCORE_Kernel.cpp:3587-3710):
     * the Inside list may grow up to N here (rounded up to whole 64-slot tiles, so
     * that the Border slots start on a tile); the Border list and both recycling
     * lists are fixed */
    d.nreg = 1;
    d.nlists = 2;
    if (prm->n_regions > 1) {
        /* several regions: per-region lists (build_regions); the one-region sizes stay 0 */
        build_regions(h, (uint32_t)prm->n_regions);
    } else {
        d.capI = (uint32_t)((d.N + 63) / 64 * 64);
        d.capB = 2 * ((d.nx * d.ny) + (d.nx * d.nz) + (d.ny * d.nz));
        d.recCapI = (uint32_t)(prm->rel_cell_caching * (double)d.N * prm->maxfill * kFullRecyclingAtt);
        d.recCapB = (uint32_t)((double)d.capB * prm->maxfill * kFullRecyclingAtt);
    }
    const bool multi = d.nreg > 1;

    h->defq.resize(4 * (size_t)d.ndefg);
    h->rxq.resize(4 * (size_t)d.nrxg + 4);
    for (uint32_t i = 0; i < d.ndefg; i++)
        memcpy(&h->defq[4 * (size_t)i], &dom->ori_quat[4 * (size_t)dom->defg_ori[i]], 4 * sizeof(double));
    for (uint32_t i = 0; i < d.nrxg; i++)
        memcpy(&h->rxq[4 * (size_t)i], &dom->ori_quat[4 * (size_t)dom->rxg_ori[i]], 4 * sizeof(double));

    /* initial grain counts after solve_REPLACE_CA_STRUCTURE (SCORE_Kernel.cpp:6390-6420) */
    std::vector<uint32_t> defcnt(d.ndefg, 0u);
    for (uint64_t i = 0; i < d.N; i++) defcnt[dom->cellgrid[i]]++;

    /* nucleus weights with glibc (calc_mobilityweight, SCORE_GrowthMachine.cpp:152); a
     * nucleus keeps the deformed grain of its site until it arrives, since a cell
     * changes its grain only by transforming, which rejects the nucleus.  The
     * site-saturated set (step 0): a site already taken by an earlier nucleus is
     * rejected (:130-138).  Later arrivals are grouped by step in pool order. */
    std::vector<double> nucP(d.nrxg + 1);
    std::vector<uint32_t> site0(d.nrxg + 1, GM_NO_SITE);
    std::vector<uint32_t> arr_ids;
    {
        std::vector<uint64_t> seen((d.N + 63) / 64, 0ull);
        uint32_t last = 0;
        for (uint32_t k = 0; k < d.nrxg; k++) {
            const uint32_t w = dom->rxg_site[k];
            const uint32_t g = dom->cellgrid[w];
            nucP[k] = mobility_weight(h->p, &h->rxq[4 * (size_t)k], &h->defq[4 * (size_t)g]);
            const uint32_t ks = dom->rxg_step ? dom->rxg_step[k] : 0u;
            if (ks > last) last = ks;
            if (ks != 0) continue;
            site0[k] = w;
            if (!(seen[w >> 6] >> (w & 63) & 1ull)) {
                seen[w >> 6] |= 1ull << (w & 63);
                h->nuclei++;
                if (nucP[k] >= h->regPmax) h->regPmax = nucP[k];
            }
        }
        if (last) {
            if (last >= GM_STEP_MAX) {
                delete h;
                return fail(SCORE_GM_ERR_ARG, "nucleus arrival step out of range");
            }
            h->arr_off.assign((size_t)last + 2, 0u);
            for (uint32_t k = 0; k < d.nrxg; k++) h->arr_off[dom->rxg_step[k] + 1]++;
            for (uint32_t k = 0; k <= last; k++) h->arr_off[k + 1] += h->arr_off[k];
            std::vector<uint32_t> fillp(h->arr_off.begin(), h->arr_off.end() - 1);
            arr_ids.resize(d.nrxg);
            for (uint32_t k = 0; k < d.nrxg; k++) arr_ids[fillp[dom->rxg_step[k]]++] = k;
            d.arr_last = last;
        }
    }

    h->pp.model = prm->mobility_model;
    h->pp.omega_hi = rh_omega_threshold(h->p);
    h->pp.fac = 1.0 / sqrt(2.0);

    /* slots plus a margin of one sweep chunk (512 slots) */
    const size_t capTot = multi ? (size_t)h->lbase[d.nlists] + 512 : (size_t)d.capI + (d.capB + 63) / 64 * 64 + 512;
    d.hcap = pow2_at_least(std::max<uint64_t>(65536, 4ull * d.nrxg));
    d.dlistCap = (uint32_t)std::min<uint64_t>(d.N, 1u << 20);
    /* completion steps further ahead than the horizon are re-predicted at its
     * multiples; SCORE_GM_HORIZON lowers it (tests use it to make that happen) */
    d.horizon = GM_CS_HORIZON_MAX;
    if (const char *ev = getenv("SCORE_GM_HORIZON")) {
        char *end = nullptr;
        const long v = strtol(ev, &end, 10);
        if (end == ev || *end != '\0' || v < 16 || v > (long)GM_CS_HORIZON_MAX) {
            delete h;
            return fail(SCORE_GM_ERR_ARG, "SCORE_GM_HORIZON must be an integer in [16, 16384]");
        }
        d.horizon = (uint32_t)v;
    }
    d.dnbx = (d.nx + 15) / 16;
    d.dnby = (d.ny + 7) / 8;
    d.dnbz = (d.nz + 7) / 8;
    d.tileCap = (uint32_t)std::max<uint64_t>(1024, (d.N / 16) / 64);
    d.parentCap = (uint32_t)std::max<uint64_t>(65536, d.N / 16);
    if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
        (void)hipGetLastError();
        delete h;
        return fail(SCORE_GM_ERR_HIP, "hipStreamCreate");
    }
    if ((rc = pool_acquire(device, &h->pool))) {
        free_device(h);
        delete h;
        return rc;
    }
    /* every cell retires once; a cell revived by a defragmentation can retire
     * again, at most once per front slot of the Border list beyond that */
    uint64_t capBsum = d.capB;
    if (multi)
        for (uint32_t r = 0; r < d.nreg; r++) capBsum += h->lcap[2 * r + 1];
    d.rlogCap = d.N + capBsum + 4096u;
    if ((rc = dalloc(h, &d.grid, d.N)) || (rc = dalloc(h, &d.claim, d.N)) ||
        (rc = dalloc(h, &d.rlog, d.rlogCap)) ||
        (rc = dalloc(h, &d.cstep, capTot)) ||
        (rc = dalloc(h, &d.rec, capTot)) ||
        (rc = dalloc(h, &d.dlist, d.dlistCap)) ||
        (rc = dalloc(h, &d.dmap, (size_t)d.dnbx * d.dnby * d.dnbz * 32)) || (rc = dalloc(h, &d.pslot, d.parentCap)) ||
        (rc = dalloc(h, &d.par, (size_t)GM_PAR_PLANES * d.parentCap)) || (rc = dalloc(h, &d.recI, d.recCapI)) || (rc = dalloc(h, &d.recB, d.recCapB)) ||
        (rc = dalloc(h, &d.mF, d.tileCap)) || (rc = dalloc(h, &d.mR, d.tileCap)) || (rc = dalloc(h, &d.pmI, d.parentCap)) ||
        (rc = dalloc(h, &d.pmB, d.parentCap)) || (rc = dalloc(h, &d.bcnt, GM_BCNT_ROWS * GM_MAX_BLOCKS)) ||
        (rc = dalloc(h, &d.rho, d.ndefg)) || (rc = dalloc(h, &d.defq, 4 * (size_t)d.ndefg)) ||
        (rc = dalloc(h, &d.rxq, 4 * (size_t)d.nrxg + 4)) || (rc = dalloc(h, &d.defcnt, d.ndefg)) ||
        (rc = dalloc(h, &d.rxcnt, (size_t)d.nrxg + 1)) || (rc = dalloc(h, &d.zlist, d.ndefg)) ||
        (rc = dalloc(h, &d.hkeys, d.hcap)) ||
        (rc = dalloc(h, &d.hvals, d.hcap)) || (rc = dalloc(h, &d.pend_slot, d.hcap)) || (rc = dalloc(h, &d.pend_key, d.hcap)) ||
        (rc = dalloc(h, &d.nflag, (size_t)d.nrxg + 1)) || (rc = dalloc(h, &d.nI_, (size_t)d.nrxg + 1)) ||
        (rc = dalloc(h, &d.nB_, (size_t)d.nrxg + 1)) || (rc = dalloc(h, &d.scan_bsum, 2 * GM_SCAN_BLOCKS)) ||
        (rc = dalloc(h, &d.st, 1))) {
        free_device(h);
        delete h;
        return rc;
    }
    if (multi) {
        const uint32_t L = d.nlists;
        if ((rc = dalloc(h, &d.reg, d.nreg)) || (rc = dalloc(h, &d.pmH, d.parentCap)) ||
            (rc = dalloc(h, &d.plist, d.parentCap)) || (rc = dalloc(h, &d.L.bstart, d.nreg + 1)) ||
            (rc = dalloc(h, &d.L.breg, GM_MAX_BLOCKS)) ||
            (rc = dalloc(h, &d.L.base, L + 1)) || (rc = dalloc(h, &d.L.cap, L)) || (rc = dalloc(h, &d.L.rcap, L)) ||
            (rc = dalloc(h, &d.L.rbase, L + 1)) || (rc = dalloc(h, &d.L.n, L)) || (rc = dalloc(h, &d.L.nold, L)) ||
            (rc = dalloc(h, &d.L.rlen, L)) || (rc = dalloc(h, &d.L.win, L)) || (rc = dalloc(h, &d.L.occ, L + 1)) ||
            (rc = dalloc(h, &d.L.tile, L + 1)) || (rc = dalloc(h, &d.L.tile_s, L + 1)) ||
            (rc = dalloc(h, &d.L.par, L + 1)) || (rc = dalloc(h, &d.L.rst, L + 1)) ||
            (rc = dalloc(h, &d.L.rbuf, h->lrbase[L])) || (rc = dalloc(h, &d.L.rs, (size_t)(d.nreg + 1) * GM_RROWS)) ||
            (rc = dalloc(h, &d.L.hoff, (size_t)d.nreg * GM_NHALO))) {
            free_device(h);
            delete h;
            return rc;
        }
    }
    if (d.arr_last &&
        ((rc = dalloc(h, &d.arr_ids, d.nrxg)) || (rc = dalloc(h, &d.arr_off, h->arr_off.size())) ||
         (rc = dalloc(h, &d.nsite, d.nrxg)) || (rc = dalloc(h, &d.nucP, d.nrxg)))) {
        free_device(h);
        delete h;
        return rc;
    }
    if ((rc = pinned_mapped(h, (void **)&h->hst, (void **)&h->hst_dev, sizeof(GmStats))) ||
        (rc = pinned_mapped(h, (void **)&h->hprog, (void **)&h->hprog_dev, 64))) {
        free_device(h);
        delete h;
        return rc;
    }
    d.prog = h->hprog_dev;
    uint32_t *d_site = nullptr;
    double *d_nucP = nullptr;
    auto setup = [&]() -> int {
        hipStream_t s = h->stream;
        HIPCHK(hipMemcpyAsync(d.grid, dom->cellgrid, sizeof(uint32_t) * d.N, hipMemcpyHostToDevice, s));
        HIPCHK(gm_launch_fill_u32(d.claim, GM_FREE_CLAIM, d.N, s));
        HIPCHK(gm_launch_dmap_init(d, s));
        /* two INACTIVE codes per word (capTot is a multiple of 64) */
        HIPCHK(gm_launch_fill_u32((uint32_t *)d.cstep, 0xFFFFFFFFu, capTot / 2, s));
        HIPCHK(gm_launch_fill_u64((unsigned long long *)d.rec, 0ull, capTot * (sizeof(GmRec) / 8), s));
        HIPCHK(hipMemcpyAsync(d.rho, dom->defg_rho, sizeof(double) * d.ndefg, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.defq, h->defq.data(), sizeof(double) * 4 * d.ndefg, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.rxq, h->rxq.data(), sizeof(double) * h->rxq.size(), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(d.defcnt, defcnt.data(), sizeof(uint32_t) * d.ndefg, hipMemcpyHostToDevice, s));
        HIPCHK(gm_launch_fill_u32(d.rxcnt, 0u, (size_t)d.nrxg + 1, s));
        HIPCHK(gm_launch_fill_u64(d.hkeys, 0ull, d.hcap, s));
        HIPCHK(gm_launch_fill_u64((unsigned long long *)d.hvals, GM_PENDING_TAG, d.hcap, s));
        HIPCHK(gm_launch_fill_u32((uint32_t *)d.st, 0u, sizeof(GmStats) / 4, s));
        int r2;
        if ((r2 = dalloc(h, &d_site, (size_t)d.nrxg + 1)) || (r2 = dalloc(h, &d_nucP, (size_t)d.nrxg + 1))) return r2;
        if (d.nrxg) {
            HIPCHK(hipMemcpyAsync(d_site, site0.data(), sizeof(uint32_t) * d.nrxg, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(d_nucP, nucP.data(), sizeof(double) * d.nrxg, hipMemcpyHostToDevice, s));
        }
        if (d.arr_last) {
            HIPCHK(hipMemcpyAsync(d.arr_ids, arr_ids.data(), sizeof(uint32_t) * d.nrxg, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(d.arr_off, h->arr_off.data(), sizeof(uint32_t) * h->arr_off.size(),
                                  hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(d.nsite, dom->rxg_site, sizeof(uint32_t) * d.nrxg, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(d.nucP, nucP.data(), sizeof(double) * d.nrxg, hipMemcpyHostToDevice, s));
        }
        if (multi) {
            if ((r2 = place_nuclei_regions(h, dom, site0, nucP))) return r2;
        } else {
            HIPCHK(gm_launch_nucleate(d, d_site, d_nucP, d.nrxg, s));
        }
        HIPCHK(gm_launch_stats_out(d, h->hst_dev, s));
        HIPCHK(hipStreamSynchronize(s));
        return 0;
    };
    rc = setup();
    dfree(h, d_site);
    dfree(h, d_nucP);
    dfree(h, d.nflag); dfree(h, d.nI_); dfree(h, d.nB_);
    d.nflag = d.nI_ = d.nB_ = nullptr;
    if (rc == 0 && h->hst->err) rc = fail(SCORE_GM_ERR_CAPACITY, "Border front list full while placing nuclei");
    if (rc) {
        free_device(h);
        delete h;
        return rc;
    }
    h->nI = h->hst->nI;
    h->nB = h->hst->nB;

    /* solve_RXGROWTH prologue (SCORE_Kernel.cpp:6501-6530): scalars and the first dt */
    h->rho_host.assign(dom->defg_rho, dom->defg_rho + d.ndefg);
    h->rho0_host = h->rho_host;
    h->rho_order.resize(d.ndefg);
    h->alive.resize(d.ndefg);
    for (uint32_t g = 0; g < d.ndefg; g++) {
        h->rho_order[g] = g;
        h->alive[g] = defcnt[g] > 0 ? 1 : 0;
    }
    std::stable_sort(h->rho_order.begin(), h->rho_order.end(),
                     [&](uint32_t a, uint32_t b) { return h->rho_host[a] > h->rho_host[b]; });
    h->rho_top = 0;
    update_rhomax(h);
    h->t = 0.0; h->X = 0.0; h->Xcells = 0.0; h->step = 0;
    update_temperature(h);
    update_atomistic(h);
    update_mobilities(h);
    h->Pmax = kDefaultPmax;
    h->dt = dt_minimum(h);
    /* myMobilityWeightMax after nucleation (SCORE_Kernel.cpp:6598-6605) */
    if (h->regPmax >= h->Pmax) h->Pmax = h->regPmax;
    *out = h;
    return SCORE_GM_OK;
}

int score_gm_step(score_gm_handle *h, uint32_t max_steps, int32_t *completed) {
    if (!h) return fail(SCORE_GM_ERR_ARG, "null handle");
    if (hipSetDevice(h->device) != hipSuccess) return hip_fail(hipGetLastError(), "hipSetDevice");
    auto t0 = std::chrono::steady_clock::now();
    int rc = 0;
    for (uint32_t k = 0; k < max_steps;) {
        uint32_t n = 0;
        rc = step_window(h, max_steps - k, &n);
        if (rc != 0) break;
        k += n;
    }
    h->loop_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (completed) *completed = h->completed;
    return rc < 0 ? rc : SCORE_GM_OK;
}

int score_gm_get_result(const score_gm_handle *h, score_gm_result *res) {
    if (!h || !res) return fail(SCORE_GM_ERR_ARG, "null argument");
    memset(res, 0, sizeof(*res));
    res->t = h->t;
    res->X = h->X;
    res->xcells = h->Xcells;
    res->step = h->step;
    res->completed = h->completed;
    res->sum_sv = h->sum_sv;
    res->nuclei_placed = h->nuclei;
    res->windows = h->windows;
    res->pmax = h->Pmax;
    res->loop_seconds = h->loop_seconds;
    return h->err;
}

int score_gm_run(score_gm_handle *h, score_gm_result *res) {
    if (!h) return fail(SCORE_GM_ERR_ARG, "null handle");
    int32_t completed = 0;
    int rc = 0;
    while (!completed) {
        rc = score_gm_step(h, 1u << 20, &completed);
        if (rc) break;
    }
    if (res) score_gm_get_result(h, res);
    return rc;
}

int score_gm_get_grainmap(score_gm_handle *h, uint32_t *out, int32_t encoding) {
    if (!h || !out) return fail(SCORE_GM_ERR_ARG, "null argument");
    if (encoding != SCORE_GM_MAP_CELLGRID && encoding != SCORE_GM_MAP_RAW2GID)
        return fail(SCORE_GM_ERR_ARG, "unknown encoding");
    HIPCHK(hipSetDevice(h->device));
    /* the grid with the front's cells marked infected, assembled in a device buffer */
    uint32_t *tmp = nullptr;
    int rc = dalloc(h, &tmp, h->d.N);
    if (rc) return rc;
    auto copy = [&]() -> int {
        HIPCHK(hipMemcpyAsync(tmp, h->d.grid, sizeof(uint32_t) * h->d.N, hipMemcpyDeviceToDevice, h->stream));
        HIPCHK(gm_launch_overlay_infected(h->d, h->nI, h->nB, h->step, tmp, h->stream));
        HIPCHK(hipMemcpyAsync(out, tmp, sizeof(uint32_t) * h->d.N, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(hipStreamSynchronize(h->stream));
        return 0;
    };
    rc = copy();
    dfree(h, tmp);
    if (rc) return rc;
    if (encoding == SCORE_GM_MAP_RAW2GID)
        for (uint64_t i = 0; i < h->d.N; i++) out[i] = out[i] == GM_INFECTED ? 1u : 2u + out[i];
    return 0;
}

int score_gm_get_counts(score_gm_handle *h, uint32_t *defg_counts, uint32_t *rxg_counts) {
    if (!h) return fail(SCORE_GM_ERR_ARG, "null handle");
    HIPCHK(hipSetDevice(h->device));
    if (defg_counts)
        HIPCHK(hipMemcpyAsync(defg_counts, h->d.defcnt, sizeof(uint32_t) * h->d.ndefg, hipMemcpyDeviceToHost, h->stream));
    if (rxg_counts && h->d.nrxg)
        HIPCHK(hipMemcpyAsync(rxg_counts, h->d.rxcnt, sizeof(uint32_t) * h->d.nrxg, hipMemcpyDeviceToHost, h->stream));
    HIPCHK(hipStreamSynchronize(h->stream));
    return 0;
}

int score_gm_get_grainlog(const score_gm_handle *h, score_gm_grainlog_row *rows, uint32_t *counts, size_t cap,
                          size_t *n) {
    if (!h || !n) return fail(SCORE_GM_ERR_ARG, "null argument");
    *n = h->glog.size();
    const size_t m = std::min(cap, h->glog.size());
    const size_t ng = (size_t)h->d.ndefg + h->d.nrxg;
    if (rows && m) memcpy(rows, h->glog.data(), m * sizeof(*rows));
    if (counts && m) memcpy(counts, h->glog_counts.data(), m * ng * sizeof(uint32_t));
    return 0;
}

int score_gm_get_trajectory(const score_gm_handle *h, score_gm_traj_row *rows, size_t cap, size_t *n) {
    if (!h || !n) return fail(SCORE_GM_ERR_ARG, "null argument");
    *n = h->traj.size();
    size_t m = std::min(cap, h->traj.size());
    if (rows && m) memcpy(rows, h->traj.data(), m * sizeof(*rows));
    return 0;
}

int score_gm_get_front(score_gm_handle *h, int32_t list, uint32_t cap, uint32_t *n, uint8_t *activity,
                       uint8_t *infector, uint32_t *cell, double *rxfrac, double *P, uint32_t *defg, uint32_t *rxg) {
    if (!h || !n) return fail(SCORE_GM_ERR_ARG, "null argument");
    if (list < 0 || (uint32_t)list >= h->d.nlists) return fail(SCORE_GM_ERR_ARG, "unknown list");
    HIPCHK(hipSetDevice(h->device));
    uint32_t cnt = list == SCORE_GM_LIST_INSIDE ? h->nI : h->nB;
    uint32_t base = list == SCORE_GM_LIST_INSIDE ? 0u : h->d.capI;
    if (h->d.nreg > 1) {
        std::vector<uint32_t> ln;
        int rc = list_lengths(h, ln);
        if (rc) return rc;
        cnt = ln[list];
        base = h->lbase[list];
    }
    *n = cnt;
    const uint32_t m = std::min(cap, cnt);
    hipStream_t s = h->stream;
    if (m) {
        /* activity and rxFrac as SCORE holds them before the next step, P without pending markers */
        char *buf = nullptr;
        const size_t bytes = (size_t)m * (1 + 1 + 4 + 4 + 4 + 8 + 8);
        int rc = dalloc(h, &buf, bytes);
        if (rc) return rc;
        GmFrontOut o;
        o.frac = reinterpret_cast<double *>(buf);
        o.P = o.frac + m;
        o.cell = reinterpret_cast<uint32_t *>(o.P + m);
        o.defg = o.cell + m;
        o.rxg = o.defg + m;
        o.act = reinterpret_cast<uint8_t *>(o.rxg + m);
        o.inf = o.act + m;
        auto copy = [&]() -> int {
            HIPCHK(gm_launch_materialize(h->d, h->step, base, m, o, s));
            if (activity) HIPCHK(hipMemcpyAsync(activity, o.act, m, hipMemcpyDeviceToHost, s));
            if (rxfrac) HIPCHK(hipMemcpyAsync(rxfrac, o.frac, 8ull * m, hipMemcpyDeviceToHost, s));
            if (infector) HIPCHK(hipMemcpyAsync(infector, o.inf, m, hipMemcpyDeviceToHost, s));
            if (cell) HIPCHK(hipMemcpyAsync(cell, o.cell, 4ull * m, hipMemcpyDeviceToHost, s));
            if (P) HIPCHK(hipMemcpyAsync(P, o.P, 8ull * m, hipMemcpyDeviceToHost, s));
            if (defg) HIPCHK(hipMemcpyAsync(defg, o.defg, 4ull * m, hipMemcpyDeviceToHost, s));
            if (rxg) HIPCHK(hipMemcpyAsync(rxg, o.rxg, 4ull * m, hipMemcpyDeviceToHost, s));
            HIPCHK(hipStreamSynchronize(s));
            return 0;
        };
        rc = copy();
        dfree(h, buf);
        return rc;
    }
    return 0;
}

}  // extern "C"

namespace {

/* ompcrit_clarify_status_infectedcells into the device buffers stage_ixyz and
 * stage_gid (nI+nB entries each), then min(cap, count) entries to the caller */
int status_staged(score_gm_handle *h, double rxthreshold, uint32_t cap, uint32_t *n, uint32_t *ixyz, uint32_t *gid,
                  uint32_t *stage_ixyz, uint32_t *stage_gid) {
    hipStream_t s = h->stream;
    HIPCHK(gm_launch_status(h->d, h->nI, h->nB, h->step, rxthreshold, stage_ixyz, stage_gid, s));
    HIPCHK(gm_launch_stats_out(h->d, h->hst_dev, s));
    HIPCHK(hipStreamSynchronize(s));
    const uint32_t cnt = h->hst->dcount;
    *n = cnt;
    const uint32_t k = std::min(cap, cnt);
    if (k && ixyz) HIPCHK(hipMemcpyAsync(ixyz, stage_ixyz, 4ull * k, hipMemcpyDeviceToHost, s));
    if (k && gid) HIPCHK(hipMemcpyAsync(gid, stage_gid, 4ull * k, hipMemcpyDeviceToHost, s));
    HIPCHK(hipStreamSynchronize(s));
    return 0;
}

}  // namespace

extern "C" {

int score_gm_get_status(score_gm_handle *h, double rxthreshold, uint32_t cap, uint32_t *n, uint32_t *ixyz,
                        uint32_t *gid) {
    if (!h || !n) return fail(SCORE_GM_ERR_ARG, "null argument");
    if (h->err) return h->err;
    HIPCHK(hipSetDevice(h->device));
    const uint32_t m = h->nI + h->nB;
    uint32_t *buf = nullptr;
    int rc = dalloc(h, &buf, 2ull * std::max<uint32_t>(m, 1u));
    if (rc) return rc;
    rc = status_staged(h, rxthreshold, cap, n, ixyz, gid, buf, buf + m);
    dfree(h, buf);
    return rc;
}

int score_gm_get_status_noalloc(score_gm_handle *h, double rxthreshold, uint32_t cap, uint32_t *n, uint32_t *ixyz,
                                uint32_t *gid) {
    if (!h || !n) return fail(SCORE_GM_ERR_ARG, "null argument");
    if (h->err) return h->err;
    if ((uint64_t)h->nI + h->nB > SCORE_GM_STATUS_NOALLOC_MAX)
        return fail(SCORE_GM_ERR_CAPACITY,
                    "front has more occupied slots than the fixed status array holds (SCORE_GM_STATUS_NOALLOC_MAX)");
    HIPCHK(hipSetDevice(h->device));
    uint32_t *stage = nullptr;
    std::unique_lock<std::mutex> lock;
    HIPCHK(gm_scratch_acquire(h->device, &stage, lock));
    return status_staged(h, rxthreshold, cap, n, ixyz, gid, stage, stage + SCORE_GM_STATUS_NOALLOC_MAX);
}

namespace {

/* one domain of a batch, create to destroy; a domain that does not fit into
 * device memory retries each time another one finishes */
struct BatchGate {
    std::mutex m;
    std::condition_variable cv;
    uint32_t inflight = 0; /* domains taken, including those waiting for memory */
    uint32_t waiting = 0;
    uint64_t finished = 0;
};

int batch_one(score_gm_batch_item &it, int32_t device, BatchGate &gate) {
    score_gm_handle *h = nullptr;
    int rc;
    for (;;) {
        uint64_t seen;
        {
            std::lock_guard<std::mutex> lk(gate.m);
            seen = gate.finished;
        }
        rc = score_gm_create(it.domain, it.params, device, &h);
        if (rc != SCORE_GM_ERR_OOM) break;
        /* wait for a running domain to give its memory back; with none running
         * besides the waiting ones, it does not fit */
        std::unique_lock<std::mutex> lk(gate.m);
        if (gate.finished == seen && gate.inflight - gate.waiting <= 1) break;
        gate.waiting++;
        gate.cv.wait(lk, [&] { return gate.finished != seen; });
        gate.waiting--;
    }
    if (rc) return rc;
    rc = score_gm_run(h, &it.result);
    if (!rc && (it.defg_counts || it.rxg_counts)) rc = score_gm_get_counts(h, it.defg_counts, it.rxg_counts);
    if (!rc && it.grainmap) rc = score_gm_get_grainmap(h, it.grainmap, it.map_encoding);
    score_gm_destroy(h);
    return rc;
}

}  // namespace

int score_gm_batch_run(score_gm_batch_item *items, uint32_t n, int32_t device, uint32_t max_inflight) {
    if (n && !items) return fail(SCORE_GM_ERR_ARG, "null items");
    for (uint32_t i = 0; i < n; i++) {
        items[i].status = SCORE_GM_ERR_STATE;
        memset(&items[i].result, 0, sizeof(items[i].result));
        items[i].seconds = 0.0;
        items[i].queued_seconds = 0.0;
    }
    const uint32_t nt = std::min<uint32_t>(n, max_inflight ? max_inflight : 8u);
    std::atomic<uint32_t> next{0};
    BatchGate gate;
    const auto tcall = std::chrono::steady_clock::now();
    auto worker = [&]() {
        for (;;) {
            const uint32_t i = next.fetch_add(1);
            if (i >= n) return;
            {
                std::lock_guard<std::mutex> lk(gate.m);
                gate.inflight++;
            }
            const auto t0 = std::chrono::steady_clock::now();
            items[i].queued_seconds = std::chrono::duration<double>(t0 - tcall).count();
            items[i].status = batch_one(items[i], device, gate);
            items[i].seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            {
                std::lock_guard<std::mutex> lk(gate.m);
                gate.inflight--;
                gate.finished++;
            }
            gate.cv.notify_all();
        }
    };
    std::vector<std::thread> pool;
    try {
        for (uint32_t t = 1; t < nt; t++) pool.emplace_back(worker);
    } catch (...) {
        /* fewer threads than asked: the ones running take every item */
    }
    worker();
    for (std::thread &t : pool) t.join();
    for (uint32_t i = 0; i < n; i++)
        if (items[i].status != SCORE_GM_OK) return items[i].status;
    return SCORE_GM_OK;
}

int score_gm_footprint(const score_gm_domain *dom, const score_gm_params *prm, uint64_t *bytes) {
    if (!bytes) return fail(SCORE_GM_ERR_ARG, "null argument");
    int rc = validate(dom, prm);
    if (rc) return rc;
    const uint64_t N = (uint64_t)dom->nx * dom->ny * dom->nz;
    uint64_t capB = 2ull * ((uint64_t)dom->nx * dom->ny + (uint64_t)dom->nx * dom->nz + (uint64_t)dom->ny * dom->nz);
    uint64_t capTot = (N + 63) / 64 * 64 + (capB + 63) / 64 * 64 + 512;
    uint64_t rbuf = (uint64_t)((double)N * prm->rel_cell_caching * prm->maxfill * kFullRecyclingAtt) +
                    (uint64_t)((double)capB * prm->maxfill * kFullRecyclingAtt);
    const uint64_t bricks = (uint64_t)((dom->nx + 15) / 16) * ((dom->ny + 7) / 8) * ((dom->nz + 7) / 8);
    const uint64_t hcap = pow2_at_least(std::max<uint64_t>(65536, 4ull * dom->n_rxg));
    const uint64_t front = std::max<uint64_t>(65536, N / 4);
    uint64_t b = 0;
    if (prm->n_regions > 1) {
        /* the lists of every region (build_regions), halo masks and tables */
        const uint32_t T = (uint32_t)prm->n_regions, npy = smallest_y_factor(T), npz = T / npy;
        const uint64_t nry = dom->ny / npy, nrz = dom->nz / npz, ex = dom->nx;
        capB = 0;
        capTot = 512;
        rbuf = 0;
        for (uint32_t iz = 0; iz < npz; iz++)
            for (uint32_t iy = 0; iy < npy; iy++) {
                const uint64_t ey = iy == npy - 1 ? dom->ny - iy * nry : nry, ez = iz == npz - 1 ? dom->nz - iz * nrz : nrz;
                const uint64_t vol = ex * ey * ez, cb = 2 * (ex * ey + ex * ez + ey * ez);
                capB += cb;
                capTot += (vol + 63) / 64 * 64 + (cb + 63) / 64 * 64;
                rbuf += (uint64_t)(prm->rel_cell_caching * (double)vol * prm->maxfill * kFullRecyclingAtt) +
                        (uint64_t)((double)cb * prm->maxfill * kFullRecyclingAtt);
            }
        b += 4 * front + (sizeof(GmRegion) + 4ull * (GM_RROWS + GM_NHALO + 26)) * (T + 1) + 16ull * GM_SW_TPC * 2 * T;
    }
    b += 8 * N;                                   /* grid, claim */
    b += 8 * (N + capB + 4096);                   /* retirement log */
    b += 32 * bricks;                             /* deformed bitmap */
    b += (sizeof(gm_cs_t) + sizeof(GmRec)) * capTot; /* completion steps, records */
    b += (4 + 16 * GM_PAR_PLANES + 8) * front;    /* parents: slots, data, winner masks */
    b += 4 * std::min<uint64_t>(N, 1u << 20);     /* deferred-cell list */
    b += 16 * (capTot / 64 + GM_SW_TPC);          /* tile masks */
    b += 4 * rbuf;                                /* recycled slots */
    b += 4 * GM_BCNT_ROWS * GM_MAX_BLOCKS;
    b += 52ull * dom->n_defg + 52ull * (dom->n_rxg + 1);
    if (dom->rxg_step) b += 20ull * dom->n_rxg;   /* arrivals: ids, sites, weights */
    b += 28 * hcap;                               /* weight cache: keys, values, pending slots and keys */
    b += sizeof(GmStats) + 8 * GM_SCAN_BLOCKS;
    *bytes = b;
    return 0;
}

struct score_gm_server {
    int32_t device = 0;
    uint64_t budget = 0;
    std::mutex m;
    std::condition_variable cv_work, cv_done;
    struct Req {
        score_gm_batch_item *item;
        uint64_t ticket, bytes;
        std::chrono::steady_clock::time_point submitted;
    };
    std::deque<Req> queue;
    std::vector<uint8_t> done;      /* by ticket */
    uint64_t in_use = 0;            /* footprints of the running requests */
    uint32_t running = 0;
    uint64_t finished = 0;
    bool stopping = false;
    std::vector<std::thread> pool;
    BatchGate gate;                 /* retries after an allocation failure despite the budget */

    void work() {
        for (;;) {
            Req r;
            {
                std::unique_lock<std::mutex> lk(m);
                cv_work.wait(lk, [&] {
                    return (stopping && queue.empty()) ||
                           (!queue.empty() && (running == 0 || in_use + queue.front().bytes <= budget));
                });
                if (queue.empty()) return;
                r = queue.front();
                queue.pop_front();
                running++;
                in_use += r.bytes;
            }
            const auto t0 = std::chrono::steady_clock::now();
            r.item->queued_seconds = std::chrono::duration<double>(t0 - r.submitted).count();
            {
                std::lock_guard<std::mutex> lk(gate.m);
                gate.inflight++;
            }
            r.item->status = batch_one(*r.item, device, gate);
            r.item->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            {
                std::lock_guard<std::mutex> lk(gate.m);
                gate.inflight--;
                gate.finished++;
            }
            gate.cv.notify_all();
            {
                std::lock_guard<std::mutex> lk(m);
                running--;
                in_use -= r.bytes;
                finished++;
                done[r.ticket] = 1;
            }
            cv_done.notify_all();
            cv_work.notify_all();
        }
    }
};

int score_gm_server_create(int32_t device, uint32_t max_inflight, uint64_t hbm_budget, score_gm_server **out) {
    if (!out) return fail(SCORE_GM_ERR_ARG, "null output");
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) {
        (void)hipGetLastError();
        return fail(SCORE_GM_ERR_NODEVICE, "no such device");
    }
    if (hipSetDevice(device) != hipSuccess) return hip_fail(hipGetLastError(), "hipSetDevice");
    if (!hbm_budget) {
        size_t fr = 0, tot = 0;
        HIPCHK(hipMemGetInfo(&fr, &tot));
        hbm_budget = (uint64_t)((double)tot * 0.9);
    }
    score_gm_server *srv = new (std::nothrow) score_gm_server;
    if (!srv) return fail(SCORE_GM_ERR_OOM, "server");
    srv->device = device;
    srv->budget = hbm_budget;
    const uint32_t nt = max_inflight ? max_inflight : 8u;
    try {
        for (uint32_t t = 0; t < nt; t++) srv->pool.emplace_back([srv] { srv->work(); });
    } catch (...) {
        if (srv->pool.empty()) {
            delete srv;
            return fail(SCORE_GM_ERR_OOM, "server threads");
        }
    }
    *out = srv;
    return 0;
}

int score_gm_server_submit(score_gm_server *srv, score_gm_batch_item *item, uint64_t *ticket) {
    if (!srv || !item || !ticket) return fail(SCORE_GM_ERR_ARG, "null argument");
    uint64_t bytes = 0;
    int rc = score_gm_footprint(item->domain, item->params, &bytes);
    item->status = SCORE_GM_ERR_STATE;
    memset(&item->result, 0, sizeof(item->result));
    item->seconds = 0.0;
    item->queued_seconds = 0.0;
    {
        std::lock_guard<std::mutex> lk(srv->m);
        if (srv->stopping) return fail(SCORE_GM_ERR_STATE, "server is shutting down");
        *ticket = srv->done.size();
        if (rc) {
            /* refused before it is queued: complete at once with its error */
            item->status = rc;
            srv->done.push_back(1);
            return rc;
        }
        srv->done.push_back(0);
        srv->queue.push_back({item, *ticket, bytes, std::chrono::steady_clock::now()});
    }
    srv->cv_work.notify_all();
    return 0;
}

int score_gm_server_wait(score_gm_server *srv, uint64_t ticket, int32_t timeout_ms) {
    if (!srv) return fail(SCORE_GM_ERR_ARG, "null server");
    std::unique_lock<std::mutex> lk(srv->m);
    if (ticket >= srv->done.size()) return fail(SCORE_GM_ERR_ARG, "unknown ticket");
    auto ready = [&] { return srv->done[ticket] != 0; };
    if (timeout_ms < 0) {
        srv->cv_done.wait(lk, ready);
        return 1;
    }
    return srv->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready) ? 1 : 0;
}

int score_gm_server_destroy(score_gm_server *srv) {
    if (!srv) return 0;
    {
        std::lock_guard<std::mutex> lk(srv->m);
        srv->stopping = true;
    }
    srv->cv_work.notify_all();
    for (std::thread &t : srv->pool) t.join();
    delete srv;
    return 0;
}

int score_gm_set_profiling(score_gm_handle *h, int32_t enable) {
    if (!h) return fail(SCORE_GM_ERR_ARG, "null handle");
    HIPCHK(hipSetDevice(h->device));
    if (enable) {
        for (hipEvent_t &ev : h->ev)
            if (!ev) HIPCHK(hipEventCreate(&ev));
        memset(&h->profc, 0, sizeof(h->profc));
    }
    h->prof = enable != 0;
    return 0;
}

int score_gm_get_profile(const score_gm_handle *h, score_gm_profile *out) {
    if (!h || !out) return fail(SCORE_GM_ERR_ARG, "null argument");
    *out = h->profc;
    return 0;
}

void score_gm_destroy(score_gm_handle *h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    free_device(h);
    delete h;
}

}  // extern "C"
