/*
 * score_gm.h -- C ABI of the MI355X growth machine.
 *
 * The library replaces the body of SCORE's recrystallization growth loop on
 * one domain, caHdl::solve_RXGROWTH (SCORE_Kernel.h:1081, SCORE_Kernel.cpp:6497-6792).
 * The caller hands over the domain exactly as SCORE holds it after
 * caHdl::solve_REPLACE_CA_STRUCTURE (SCORE_Kernel.cpp:6353-6445): the cell grid
 * holds deformed-grain ids, the pools hold orientation indices, dislocation
 * densities and nucleus sites.  The library places the site-saturated nuclei,
 * runs the step loop on the GPU and returns the grain map, the grain counts and
 * the per-step trajectory.
 *
 * All functions return 0 on success and a negative SCORE_GM_ERR_* code on
 * failure; nothing throws across the ABI.  Host buffers are caller-owned and
 * copied; no caller pointer is retained after a call returns.  A handle is used
 * by one host thread at a time; distinct handles are independent.
 */
#ifndef SCORE_GM_H
#define SCORE_GM_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SCORE_GM_ABI_VERSION 11

/* status codes */
#define SCORE_GM_OK                 0
#define SCORE_GM_ERR_ARG           -1  /* invalid argument or domain */
#define SCORE_GM_ERR_OOM           -2  /* device or host allocation failed */
#define SCORE_GM_ERR_HIP           -3  /* HIP runtime error */
#define SCORE_GM_ERR_CAPACITY      -4  /* a fixed-size front list overflowed (SCORE would write out of bounds) */
#define SCORE_GM_ERR_UNSUPPORTED   -5  /* option this library does not implement */
#define SCORE_GM_ERR_STATE         -6  /* call not valid in the handle's current state */
#define SCORE_GM_ERR_NODEVICE      -7  /* no usable gfx950 device */

/* mobility models, same numbering as SCORE_Defs_Functionality.h:200-201 */
#define SCORE_GM_MOB_SEBALDGOTTSTEIN 1
#define SCORE_GM_MOB_ROLLETTHOLM     2

/* grain-map encodings for score_gm_get_grainmap */
#define SCORE_GM_MAP_CELLGRID   0  /* raw cell-grid values: defg id, ndefg+rxg id, or 4096002000 (infected) */
#define SCORE_GM_MAP_RAW2GID    1  /* 1 = infected, else 2+gid (SCORE_Io_Characterization.cpp:583-676) */

/* sentinel, SCORE_Defs_Functionality.h:95 */
#define SCORE_GM_CURRENTLY_INFECTED 4096002000u

/* front lists, SCORE's myRXFrontInside / myRXFrontBorder (SCORE_Kernel.h:1403-1585) */
#define SCORE_GM_LIST_INSIDE 0
#define SCORE_GM_LIST_BORDER 1

/* Domain after solve_REPLACE_CA_STRUCTURE (SCORE_Kernel.cpp:6353-6445). */
typedef struct score_gm_domain {
    uint32_t nx, ny, nz;          /* edge lengths in cells along RD (x), TD (y), ND (z) */
    double cellsize;              /* m (myCAGeometry.cellsize) */
    const uint32_t *cellgrid;     /* nx*ny*nz deformed-grain ids, index x + y*nx + z*nx*ny */
    uint32_t n_ori;               /* orientation pool (myoripool) */
    const double *ori_quat;       /* n_ori*4: q0,q1,q2,q3 per orientation */
    uint32_t n_defg;              /* deformed-grain pool (mydefgpool) */
    const uint32_t *defg_ori;     /* n_defg orientation indices (caori) */
    const double *defg_rho;       /* n_defg dislocation densities, m^-2 (rho) */
    uint32_t n_rxg;               /* nucleus pool (myrxgpool) */
    const uint32_t *rxg_ori;      /* n_rxg orientation indices (caori) */
    const uint32_t *rxg_site;     /* n_rxg global linear sites (nucsite) */
    /* n_rxg arrival steps, or NULL for site saturation (every nucleus at step 0).
     * Not in SCORE's OpenMP path (SCORE_Kernel.cpp:519-522): a nucleus of step
     * k > 0 is placed at the start of step k, where the reference marks the
     * injection point of time-dependent nucleation (:6632), in pool order, by the
     * same rule as the site-saturated ones (infect_JuvenileNucleation,
     * SCORE_GrowthMachine.cpp:124-189: rejected unless its cell is still deformed,
     * appended to the Inside or Border list, rxFrac 1.0). */
    const uint32_t *rxg_step;
} score_gm_domain;

/* Physics and integrator settings (caHdl::init_parameter, SCORE_Kernel.cpp:5747-5880). */
typedef struct score_gm_params {
    int32_t mobility_model;       /* SCORE_GM_MOB_* */
    int32_t zener_mode;           /* dispersoid drag: 0 none, 1 constant (first f/r of the table), 2 f/r linear in time */
    int32_t recovery_mode;        /* 0 none, 1 Nes vacancy-core diffusion, 2 Nes solute drag, 3 Michalak-Paxton */
    /* OMP_NUM_THREADS of the SCORE run to reproduce (0 or 1: one region).  SCORE
     * splits the domain into that many memory regions (ompshar_ca2regions,
     * SCORE_GridPartitioning.cpp:69-166): npy = smallestYFactor(n) slabs along y
     * times n / npy along z, each with its own Inside and Border lists; an
     * infection that crosses into another region is parked in a halo and taken up
     * by that region after the step (SCORE_Kernel.cpp:4494-4555, :4777-4872), so
     * the outcome depends on n.  With n > 1 each region must be more than 10 cells
     * thick along y and z (MINIMUM_REGION_SIZE, SCORE_GridPartitioning.cpp:93);
     * at most 1024 regions. */
    int32_t n_regions;
    /* Rollett-Holm (MobilityModel 2) */
    double rh_hagb_m0;            /* m^4/(J s) */
    double rh_hagb_hact;          /* J (eV * 1.602176565e-19, as SCORE_Kernel.cpp:579) */
    double rh_cut, rh_trans, rh_exponent;
    /* Sebald-Gottstein (MobilityModel 1) */
    double lagb_m0, lagb_hact, hagb_m0, hagb_hact, gs_m0, gs_hact;   /* m^4/(J s), J */
    /* material */
    double G0;                    /* ZeroKelvinShearModulusG0, Pa */
    double b0;                    /* ZeroCelsiusBurgersVector, m */
    /* annealing schedule: time (s, strictly increasing from 0) and temperature (K) */
    uint32_t n_sched;
    uint32_t reserved1;
    const double *sched_time;
    const double *sched_temp;
    /* integrator */
    double maxfill;               /* MAXFILLPERSTEP */
    double xmax;                  /* XMAX */
    double tmax;                  /* TIMEMAX, s */
    int64_t nmax;                 /* NMAX */
    double rel_cell_caching;      /* InitialCacheSizeRXCells */
    double rel_cell_recaching;    /* ReCacheSizeRXCells */
    /* on-the-fly defragmentation (ompshar_defragmentation, SCORE_Kernel.cpp:3877-3934):
     * thresholds in cells, compared with Xcells at the start of a step, at most one
     * per step (SCORE_Kernel.cpp:6625-6630).  SCORE multiplies its X log points by
     * the box volume (SCORE_OnTheFlyCharacterization.cpp:56-57); pass those products. */
    uint32_t n_defrag;
    uint32_t reserved2;
    const double *defrag_at;
    /* Zener drag p_z = zener_alpha * zener_gamma * f/r (get_zener, update_microchemistry,
     * get_dtmax_instantslope_zener; SCORE_Kernel.cpp:5370-5456, :5639-5666): the table
     * EvoMicrochemistry of times (s) and f/r (1/m); mode 2 needs at least two rows,
     * strictly increasing in time from t = 0 */
    double zener_alpha;           /* ZenerAlpha */
    double zener_gamma;           /* ZenerGamma, J/m^2 */
    uint32_t n_disp;
    uint32_t reserved3;
    const double *disp_time;
    const double *disp_fr;
    /* recovery of the deformed grains' dislocation density after every step
     * (update_deformedsubstructure, SCORE_Kernel.cpp:5408-5437; SCORE_RecoveryModels.cpp:50-125)
     * and its time-step rule (get_dtmax_instantslope_rho, SCORE_Kernel.cpp:5594-5636):
     * mode 1 grows the network radius 1/sqrt(rho) of every grain, mode 2 (as in SCORE)
     * changes no rho but limits the step, mode 3 scales rho0 by a fixed function of time */
    double rec_vacancy_d0;        /* RecoveryVacancyDiffPreexp */
    double rec_vacancy_hact;      /* RecoveryVacancyDiffHact, J */
    double rec_solute_d0;         /* RecoverySoluteDiffPreexp */
    double rec_solute_hact;       /* RecoverySoluteDiffHact, J */
    double rec_solute_ls;         /* RecoverySoluteLsPropFactor */
    double rec_solute_conc;       /* RecoverySoluteConcentration */
    double rec_alpha3, rec_kappa2, rec_c3, rec_c5;   /* RecoveryParameterAlpha3, Kappa2, C3, C5 */
    /* recovery_mode 4, not in SCORE: rho = rho0 * exp(-t / rec_tau) for every
     * deformed grain, applied where SCORE applies its models; no time-step rule */
    double rec_tau;               /* s */
    /* grain-evolution log points (log_grainevo, SCORE_OnTheFlyCharacterization.cpp:478-530,
     * called from log_OUTPUT_Diehl, :646-651): thresholds in cells (SCORE's X log
     * points times the box volume, :56-57); in the master section of the first
     * step whose Xcells reaches the next threshold, before t advances, the cell
     * count of every grain is logged (at most one log per step) */
    uint32_t n_grainlog;
    uint32_t reserved4;
    const double *grainlog_at;
} score_gm_params;

/* One row per executed growth step (the values SCORE's master section holds,
 * SCORE_Kernel.cpp:6688-6742).  t and dt are the time at the start of the step
 * and the increment used in it; X, xcells and rhomax are the values after it. */
typedef struct score_gm_traj_row {
    uint32_t step;
    uint32_t sv;                  /* ACTIVE front cells advanced in the step (Sv) */
    uint32_t nfull_inside;        /* cells that completed, Inside list */
    uint32_t nfull_border;        /* cells that completed, Border list */
    uint32_t nslots_inside;       /* nextSlotNeverActiveRXFrontInside after the step */
    uint32_t nslots_border;       /* nextSlotNeverActiveRXFrontBorder after the step */
    uint32_t ninfected;           /* new infections in the step */
    uint32_t npending;            /* mobility weights resolved on the host after the step */
    double t;
    double dt;
    double X;
    double xcells;
    double temperature;
    double rhomax;
    double pmax;                  /* myMobilityWeightMax after the step (log_rxfrontstats localPmax) */
    uint32_t nplaced;             /* nuclei placed at the start of the step (arrivals after step 0) */
    uint32_t reserved;
} score_gm_traj_row;

/* One grain-evolution log entry (loginfo_grainevo_ca, SCORE_Kernel.h:642-652). */
typedef struct score_gm_grainlog_row {
    double t;                     /* localtime: time at the start of the logged step */
    double X;                     /* localX after it */
    uint32_t logcnt;              /* locallogcnt */
    uint32_t step;                /* localtstep */
    uint32_t sv;                  /* localSv */
    uint32_t ngrains;             /* nlocaldata = n_defg + n_rxg */
} score_gm_grainlog_row;

typedef struct score_gm_result {
    double t;                     /* tsimend */
    double X;
    double xcells;
    uint32_t step;                /* stepsimend */
    int32_t completed;            /* 1 once the termination test of SCORE_Kernel.cpp:6619 fired */
    uint64_t sum_sv;              /* sum of Sv over all steps (cell updates) */
    uint32_t nuclei_placed;       /* nmynuclei */
    uint32_t windows;             /* launch windows the steps ran in (one host round trip each) */
    double pmax;                  /* myMobilityWeightMax as tracked by the host */
    double loop_seconds;          /* host wall time spent inside step/run calls */
} score_gm_result;

/* Kernels of one growth step, in launch order (index into ms_kernel). */
#define SCORE_GM_K_REANCHOR 0    /* steps whose scalars changed: every cell advanced and predicted again */
#define SCORE_GM_K_SWEEP 1       /* calcGrowthStepInside/Border over the occupied slots */
#define SCORE_GM_K_SWEEP_SCAN 2  /* list offsets of the sweep's blocks */
#define SCORE_GM_K_COMPACT 3     /* FullRX list */
#define SCORE_GM_K_PROPOSE 4     /* parent data; every infection attempt claims its target */
#define SCORE_GM_K_DECIDE 5      /* winners per parent */
#define SCORE_GM_K_DECIDE_SCAN 6 /* ranks of the winners' blocks */
#define SCORE_GM_K_RECYCLE 7     /* the recycled slots that are handed out */
#define SCORE_GM_K_COMMIT 8      /* new front records with their predicted completion steps, retirement of the parents */
#define SCORE_GM_NKERNELS 9

/* Per-kernel device time and work counters, the counterpart of SCORE's per-thread
 * phase timers (omp_log_rxfrontstats, SCORE_Kernel.cpp:5019-5063; written to
 * ThreadProfilingGrowth.csv when OutputThreadProfiling is set).  Times are HIP
 * events recorded on the handle's stream between the kernels of every step;
 * counters are sums over the profiled steps. */
typedef struct score_gm_profile {
    uint64_t steps;               /* steps recorded while profiling was on */
    double ms_kernel[SCORE_GM_NKERNELS];
    double ms_step;               /* device time from the first to the last kernel of each step */
    double ms_host;               /* host wall time of the profiled steps */
    uint64_t sum_sv;              /* ACTIVE slots advanced (Sv) */
    uint64_t slots_scanned;       /* occupied slots swept (nextSlotNeverActive of both lists) */
    uint64_t n_full;              /* completed cells, the FullRX list lengths */
    uint64_t n_recycled;          /* recycled slots handed out (recycling-list entries written) */
    uint64_t n_attempts;          /* infection attempts that found a deformed target */
    uint64_t n_new;               /* infections committed */
    uint64_t n_predicted;         /* new cells given their increment and completion step when infected */
    uint64_t n_reanchor_steps;    /* steps whose scalars changed: every cell re-anchored */
    uint64_t n_reanchored;        /* slots re-anchored */
    uint64_t pending_resolved;    /* mobility weights resolved on the host */
} score_gm_profile;

typedef struct score_gm_handle score_gm_handle;

/* ABI version of the loaded library. */
int score_gm_abi_version(void);

/* Human-readable text for a status code. */
const char *score_gm_strerror(int status);

/* Number of usable gfx950 devices (0 when none). */
int score_gm_device_count(void);

/* Copy the domain to device `device`, allocate the front lists, place the
 * site-saturated nuclei (ompshar_sim_myCA_sitesaturatedNucleation,
 * SCORE_GrowthMachine.cpp:50-189) and compute the first time step
 * (SCORE_Kernel.cpp:6497-6530).  Replaces the set-up part of solve_RXGROWTH. */
int score_gm_create(const score_gm_domain *dom, const score_gm_params *prm,
                    int device, score_gm_handle **out);

/* Advance by up to `max_steps` growth steps of the main loop
 * (SCORE_Kernel.cpp:6615-6747).  *completed is set to 1 once the termination
 * test fires.  Replaces iterations of solve_RXGROWTH's loop. */
int score_gm_step(score_gm_handle *h, uint32_t max_steps, int32_t *completed);

/* Run to completion (XMAX, TIMEMAX, NMAX, end of schedule or all cells
 * transformed) and fill *res.  Replaces the whole of solve_RXGROWTH. */
int score_gm_run(score_gm_handle *h, score_gm_result *res);

/* Current scalars (tsimend, X, Xcells, step, ...). */
int score_gm_get_result(const score_gm_handle *h, score_gm_result *res);

/* Grain map in the requested SCORE_GM_MAP_* encoding, nx*ny*nz values. */
int score_gm_get_grainmap(score_gm_handle *h, uint32_t *out, int32_t encoding);

/* Cell counts per deformed grain (mydefgpool[].cellcount, n_defg values) and
 * per nucleus (myrxgpool[].cellcount, n_rxg values). Either pointer may be NULL. */
int score_gm_get_counts(score_gm_handle *h, uint32_t *defg_counts, uint32_t *rxg_counts);

/* Per-step trajectory: copies min(cap, rows recorded) rows and sets *n to the
 * number of rows recorded. */
int score_gm_get_trajectory(const score_gm_handle *h, score_gm_traj_row *rows,
                            size_t cap, size_t *n);

/* The grain-evolution log (caHdl::mygrainevolution): copies min(cap, recorded)
 * entries in logging order and sets *n to the number recorded.  counts, unless
 * NULL, receives for each copied entry its n_defg deformed-grain counts, then its
 * n_rxg nucleus counts (localdatabucket). */
int score_gm_get_grainlog(const score_gm_handle *h, score_gm_grainlog_row *rows, uint32_t *counts, size_t cap,
                          size_t *n);

/* Slot-level copy of one front list (SCORE_GM_LIST_*), slots [0, *n) where *n
 * is nextSlotNeverActive of that list (SCORE_Kernel.h:124-137 record):
 * activity, infector, linear cell index, rxFrac, P, deformed and nucleus ids.
 * Each array may be NULL; otherwise it must hold `cap` entries. */
int score_gm_get_front(score_gm_handle *h, int32_t list, uint32_t cap, uint32_t *n,
                       uint8_t *activity, uint8_t *infector, uint32_t *cell,
                       double *rxfrac, double *P, uint32_t *defg, uint32_t *rxg);

/* Grain assignment of the cells on the front, as
 * caHdl::ompcrit_clarify_status_infectedcells(rxthreshold) collects it
 * (SCORE_Io_Characterization.cpp:925-997): for every ACTIVE record, Inside list
 * then Border list in slot order, the linear cell index and the grain id the
 * cell shows (n_defg + nucleus id once rxFrac >= rxthreshold, else its deformed
 * grain id).  SCORE calls it with RXFRACTION_THRESHOLD = 0.5.  Sets *n to the
 * number of entries and copies min(cap, *n) of them; either array may be NULL. */
int score_gm_get_status(score_gm_handle *h, double rxthreshold, uint32_t cap, uint32_t *n,
                        uint32_t *ixyz, uint32_t *gid);

/* Largest number of occupied front slots (nextSlotNeverActive of both lists
 * together) that score_gm_get_status_noalloc accepts. */
#define SCORE_GM_STATUS_NOALLOC_MAX (1u << 22)

/* score_gm_get_status without any device memory allocated by the call.  No
 * counterpart in the reference: it is for callers that cannot allocate device
 * memory (code under graph capture, tight serving loops, harnesses that forbid
 * it).  The entries are staged in a fixed device array of the library
 * (SCORE_GM_STATUS_NOALLOC_MAX entries per column, 32 MiB, held on every GPU the
 * library runs on until the process ends); a handle with more occupied slots is
 * refused with SCORE_GM_ERR_CAPACITY.  Calls on one GPU share that array and are
 * serialised by the library; the array keeps the last call's contents, of which
 * only the entries a call writes are returned. */
int score_gm_get_status_noalloc(score_gm_handle *h, double rxthreshold, uint32_t cap, uint32_t *n,
                                uint32_t *ixyz, uint32_t *gid);

/* One domain of a batch: what ensembleHdl::SIMULATE_myCAs (SCORE_Kernel.cpp:1693-1757)
 * hands to solve_RXGROWTH for each caHdl of the rank, and what it reads back. */
typedef struct score_gm_batch_item {
    const score_gm_domain *domain; /* in */
    const score_gm_params *params; /* in */
    uint32_t *defg_counts;         /* out: n_defg entries, or NULL */
    uint32_t *rxg_counts;          /* out: n_rxg entries, or NULL */
    uint32_t *grainmap;            /* out: nx*ny*nz entries, or NULL */
    int32_t map_encoding;          /* SCORE_GM_MAP_* of grainmap */
    int32_t status;                /* out: SCORE_GM_OK or this domain's error */
    score_gm_result result;        /* out */
    double seconds;                /* out: wall time of this domain, create to destroy */
    double queued_seconds;         /* out: from submission (or the batch call) to its start */
} score_gm_batch_item;

/* Runs every domain of `items` to completion on one device, as the rank's loop
 * over its domains does (SCORE_Kernel.cpp:1693-1757), with up to max_inflight
 * domains (0 selects 8) in flight at once, each on its own handle, stream and
 * host thread, taken in index order.  A domain whose device memory does not fit
 * next to the ones in flight waits for one of them to finish.  Returns
 * SCORE_GM_OK when every item succeeded, else the status of the lowest failed
 * item; every item's own status is set either way. */
int score_gm_batch_run(score_gm_batch_item *items, uint32_t n, int32_t device, uint32_t max_inflight);

/* Device memory one domain needs at most, in bytes: the arrays score_gm_create
 * allocates plus the per-step lists at the size a front of N/4 cells needs. */
int score_gm_footprint(const score_gm_domain *dom, const score_gm_params *prm, uint64_t *bytes);

/* A request loop over domains that arrive while others run (continuous
 * batching): each submitted domain starts as soon as a worker is free and its
 * footprint fits into the memory budget next to the domains running (in
 * submission order; a domain alone is always started), runs on its own handle
 * and stream, and is then released.  SCORE has no counterpart; it serves the
 * same call, solve_RXGROWTH, to a stream of requests. */
typedef struct score_gm_server score_gm_server;

/* max_inflight 0 selects 8; hbm_budget 0 selects 90% of the device's memory. */
int score_gm_server_create(int32_t device, uint32_t max_inflight, uint64_t hbm_budget, score_gm_server **out);

/* Queues one domain.  The item and every array it points to must stay valid
 * until the request is complete; its outputs are written as in
 * score_gm_batch_run.  *ticket identifies the request for score_gm_server_wait. */
int score_gm_server_submit(score_gm_server *srv, score_gm_batch_item *item, uint64_t *ticket);

/* Waits up to timeout_ms (negative: no limit) for a request; returns 1 once it
 * is complete (its item's status says how it ended), 0 on time-out. */
int score_gm_server_wait(score_gm_server *srv, uint64_t ticket, int32_t timeout_ms);

/* Completes every queued request, then releases the workers. */
int score_gm_server_destroy(score_gm_server *srv);

/* Switch per-phase event timing on (1) or off (0); switching on clears the
 * counters.  Replaces SCORE's OutputThreadProfiling switch (SCORE_Kernel.cpp:6677-6683). */
int score_gm_set_profiling(score_gm_handle *h, int32_t enable);

/* Counters accumulated since profiling was switched on. */
int score_gm_get_profile(const score_gm_handle *h, score_gm_profile *out);

/* Release all device memory of the handle (omp_cleanupMemoryUsedForGrowthMachine,
 * SCORE_Kernel.cpp:5066-5133). NULL is accepted. */
void score_gm_destroy(score_gm_handle *h);

#ifdef __cplusplus
}
#endif

#endif /* SCORE_GM_H */
