/*
 * score_oracle.c -- TEST INFRASTRUCTURE ONLY.
 *
 * A serial CPU restatement of SCORE's growth machine on one domain
 * (caHdl::solve_RXGROWTH, SCORE_Kernel.cpp:6497-6792) with the semantics of a
 * run with OMP_NUM_THREADS = params.n_regions: the domain split into that many
 * memory regions, each with an Inside and a Border front list, recycle-then-append
 * slot assignment, first-come infection in list order and fixed direction order;
 * an infection into another region is parked in a halo and taken up by that
 * region after the step.  The regions run in parallel in SCORE but touch only
 * their own cells, lists and halos between barriers, so one region after another
 * here gives the same result.  The GPU library is tested
 * against it; it is never linked into the library, and only tests/, the smoke
 * entry point and the CPU-baseline leg of bench.py load it.
 *
 * Each function names the reference lines it follows.  It consumes the same
 * score_gm_domain / score_gm_params structs as the library (include/score_gm.h)
 * so both receive byte-identical inputs.
 *
 * Arithmetic: FP64, compiled with -ffp-contract=off and without -march so that
 * no FMA is formed, matching SCORE's GCC -O3 build (CMakeLists.txt:23,89-91).
 * Transcendentals come from the host glibc (acos, pow, exp, cos, sin, sqrt), as
 * in the reference.
 *
 * Parity of this file against the reference is pinned by tests/test_oracle_*.py
 * (closed forms, brute force over the cubic group, invariants, hand-derived
 * first steps, stable compaction, and the statistics of the reference's shipped
 * example/jmak run).  The disorientation behind every mobility weight
 * (oracle_disori_angle, oracle_disori_quat) is pinned bit for bit against the
 * reference's own SCORE_Math.cpp, built into oracle/_ref (tests/test_ref_pins.py);
 * the weight's outer expression is the reference's one-line formula on that
 * angle with the same glibc exp/pow/cos, checked there and by its closed forms.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "score_gm.h"

/* ---- constants (SCORE_Defs_Functionality.h:83-101, SCORE_Defs_Physics.h:51-160, SCORE_Math.h:54-70) ---- */
#define O_CURRENTLY_INFECTED 4096002000u
#define O_INVALID_ADDRESS 4294967290u
#define O_KBOLTZMAN (1.3806488e-23)
#define O_PI (3.1415926535897932384626433832795)
#define O_RHO_RX (1.0e11)
#define O_DOUBLE_ACCURACY (5.0e-16)
#define O_INITIAL_DELTAT (1e-6)
#define O_INFINITE (1e36)
#define O_SMALL_HEATRATE (0.000277777777777777)
#define O_SMALL_HEAT (1.0)
#define O_SMALL_DRAGGINGRATE (0.027777777)
#define O_SMALL_ZENERFORCE (100.0)
#define O_SMALL_RECOVERYRATE (2.77e10)
#define O_SMALL_RHO (1.0e11)
#define O_RHOMAX_WELLANNEALED (1e8)
#define O_DEFAULT_PMAX (0.0)
#define O_DEFAULT_SMALL_NUMBER (1e-6)
#define O_MAXDISORI_LAGB2HAGB ((15.0) / (180.0) * (O_PI))
#define O_MAXDISORI_TO_40DEG111 ((10.0) / (180.0) * (O_PI))
#define O_FULLRXCACHING_ATT_FACTOR (1.5)
#define O_FULLRECYCLING_ATT_FACTOR (3.0)
#define O_DEFAULT_NMAX (1000000)   /* DEFAULT_NMAX, SCORE_Defs_Physics.h:133 */
#define O_MINIMUM_REGION_SIZE (10) /* SCORE_Defs_Functionality.h:85 */
#define O_NO_HALO (0xFFFFFFFFu)
#define O_FGEOFACE (1.0000)
#define O_FGEOEDGE (0.7957)
#define O_FGEODIAG (0.6494)

/* shape factors indexed by infector code (SCORE_Kernel.cpp:4051-4053) */
static const double o_shape[27] = {
    O_FGEODIAG, O_FGEOEDGE, O_FGEODIAG, O_FGEOEDGE, O_FGEOFACE, O_FGEOEDGE, O_FGEODIAG, O_FGEOEDGE, O_FGEODIAG,
    O_FGEOEDGE, O_FGEOFACE, O_FGEOEDGE, O_FGEOFACE, O_FGEOFACE, O_FGEOEDGE, O_FGEOFACE, O_FGEOEDGE,
    O_FGEODIAG, O_FGEOEDGE, O_FGEODIAG, O_FGEOEDGE, O_FGEOFACE, O_FGEOEDGE, O_FGEODIAG, O_FGEOEDGE, O_FGEODIAG, (1.0)};

/* The 26 infection attempts in call order (SCORE_Kernel.cpp:4611-4638 and
 * :4719-4748): offset, infector code, carry class (1 face, 2 edge, 3 corner). */
typedef struct { int dx, dy, dz, code, carry; } o_dir;
static const o_dir o_dirs[26] = {
    {-1, -1, 0, 16, 2}, {0, -1, 0, 15, 1}, {+1, -1, 0, 14, 2}, {-1, 0, 0, 13, 1}, {+1, 0, 0, 12, 1},
    {-1, +1, 0, 11, 2}, {0, +1, 0, 10, 1}, {+1, +1, 0, 9, 2},
    {-1, +1, -1, 19, 3}, {0, +1, -1, 18, 2}, {+1, +1, -1, 17, 3}, {-1, -1, -1, 25, 3}, {0, -1, -1, 24, 2},
    {+1, -1, -1, 23, 3}, {-1, 0, -1, 22, 2}, {0, 0, -1, 21, 1}, {+1, 0, -1, 20, 2},
    {-1, -1, +1, 8, 3}, {0, -1, +1, 7, 2}, {+1, -1, +1, 6, 3}, {-1, 0, +1, 5, 2}, {0, 0, +1, 4, 1},
    {+1, 0, +1, 3, 2}, {-1, +1, +1, 2, 3}, {0, +1, +1, 1, 2}, {+1, +1, +1, 0, 3}};

/* front record, the 32-byte `cell` of SCORE_Kernel.h:124-137 */
typedef struct {
    uint8_t act;
    uint8_t inf;
    int16_t x, y, z;
    double frac;
    double P;
    uint32_t defg;
    uint32_t rxg;
} o_cell;

/* one front list with its recycling and FullRX bookkeeping (SCORE_Kernel.cpp:3587-3710) */
typedef struct {
    o_cell *c;
    uint32_t ntotal;        /* allocated slots */
    uint32_t nnever;        /* nextSlotNeverActive */
    int growable;           /* Inside grows (:3795-3843); Border is fixed (:3847-3874) */
    uint32_t *rec;
    uint32_t rec_cap;       /* ntotalRecyclingList, fixed */
    uint32_t rec_next;      /* nextSlotThatBecomesRecycled */
    uint32_t rec_first_not; /* firstNotRecycledYet */
    uint32_t *full;
    uint32_t full_cap;
    uint32_t full_n;        /* nextSlotToFullRX */
    uint32_t sv;
    uint32_t nupdated;
} o_list;

/* halo of a region towards one neighbouring region (ompshar_init_haloregions,
 * SCORE_Kernel.cpp:2223-2457; ompshar_initialize_halo, :2085-2127): the layer of
 * the neighbour's cells next to this region; an infection this region directs at
 * such a cell is parked in the halo, first come first served, and the neighbour
 * applies it after the step (ompshar_synchronize_haloregions_rx, :4777-4872).
 * A halo cell's status is never reset. */
typedef struct {
    uint32_t ext[3], org[3];   /* extent (all 0: no halo) and origin, global coordinates */
    uint32_t inregion;         /* the region the halo lies in */
    uint8_t *status;           /* infection parked here at some step */
    uint8_t *inf;
    int16_t *xyz;
    double *carry;
    uint32_t *rxg;
    uint32_t *refs;            /* theHaloRefs of this step, nextFreeHaloRefs entries */
    uint32_t nref;
} o_halo;

/* one memory region (ompshar_ca2regions, SCORE_GridPartitioning.cpp:69-166) with
 * its own front lists, neighbour table and halos */
typedef struct {
    uint32_t lo[3], hi[3];     /* inclusive box */
    o_list in, bd;
    uint32_t nbor[27];         /* mynborregions ids in SCORE's order */
    uint32_t nborhalo[27];     /* mynborhalos (ompshar_link_haloregions): halo index or O_NO_HALO */
    o_halo halo[27];
} o_region;

uint32_t oracle_smallest_y_factor(uint32_t n);

typedef struct oracle {
    uint32_t nx, ny, nz, nxy;
    uint64_t N;
    double cellsize, inv_dx;
    uint32_t *grid;
    uint32_t n_defg, n_rxg;
    double *defq, *rxq, *rho, *rho0;
    uint32_t *site;
    uint32_t *defcnt, *rxcnt;
    score_gm_params p;
    double *sched_t, *sched_T, *defrag_at;
    double *disp_t, *disp_fr;  /* EvoMicrochemistry: time, f/r */
    double fr;                 /* myDragData.fr */
    double tprocessingend;
    o_region *reg;             /* the regions of OMP_NUM_THREADS = nreg threads */
    uint32_t nreg, npy, npz;
    /* caHdl scalars */
    double t, X, Xcells, dt, T, G, b, Gbhalfsq;
    double mLAGB, mHAGB, mGS, mRHHAGB, pzmin, rhomax, Pmax, reg_Pmax;
    uint32_t step;
    uint32_t nuclei;
    uint32_t defrag_cnt;
    int completed;
    int err;
    uint64_t sum_sv;
    score_gm_traj_row *traj;
    size_t ntraj, captraj;
    uint32_t last_infected;
    /* nuclei arriving after step 0: pool ids ordered by arrival step, and the
     * offset of each step's first one (arr_off[k] .. arr_off[k+1]) */
    uint32_t *arr_ids, *arr_off;
    uint32_t arr_last;      /* last arrival step (0: site saturation only) */
    /* grain-evolution log (mygrainevolution) */
    double *glog_at;
    uint32_t glog_cnt;
    score_gm_grainlog_row *glog;
    uint32_t *glog_counts;  /* n_defg + n_rxg per row */
    size_t nglog, capglog;
} oracle;

/* ---------------- orientation math ---------------- */

/* Hamilton product, multiply_quat_degraef (SCORE_Math.cpp:230-239) with P_EIJK = +1 */
static void o_qmul(const double *p, const double *q, double *r) {
    r[0] = p[0] * q[0] - p[1] * q[1] - p[2] * q[2] - p[3] * q[3];
    r[1] = q[0] * p[1] + p[0] * q[1] + (p[2] * q[3] - p[3] * q[2]);
    r[2] = q[0] * p[2] + p[0] * q[2] + (-(p[1] * q[3]) + p[3] * q[1]);
    r[3] = q[0] * p[3] + p[0] * q[3] + (p[1] * q[2] - p[2] * q[1]);
}

/* The six Grimmer candidate quaternions of the misorientation r
 * (disori_angle_oriquat_cubic_imm, SCORE_Math.cpp:592-602). */
static void o_grimmer(const double *r, double r0[6][4]) {
    double a = r[0], b = r[1], c = r[2], d = r[3];
    double fac = 1.0 / sqrt(2.0);
    r0[0][0] = (a + b) * fac; r0[0][1] = (a - b) * fac; r0[0][2] = (c + d) * fac; r0[0][3] = (c - d) * fac;
    r0[1][0] = (a + c) * fac; r0[1][1] = (a - c) * fac; r0[1][2] = (b + d) * fac; r0[1][3] = (b - d) * fac;
    r0[2][0] = (a + d) * fac; r0[2][1] = (a - d) * fac; r0[2][2] = (b + c) * fac; r0[2][3] = (b - c) * fac;
    r0[3][0] = (a + b + c + d) * 0.5; r0[3][1] = (a + b - c - d) * 0.5; r0[3][2] = (a - b + c - d) * 0.5; r0[3][3] = (a - b - c + d) * 0.5;
    r0[4][0] = (a + b + c - d) * 0.5; r0[4][1] = (a + b - c + d) * 0.5; r0[4][2] = (a - b + c + d) * 0.5; r0[4][3] = (a - b - c - d) * 0.5;
    r0[5][0] = a; r0[5][1] = b; r0[5][2] = c; r0[5][3] = d;
}

/* misorientation q^-1 p (SCORE_Math.cpp:564-572) */
static void o_misori(const double *p, const double *q, double *r) {
    double qm1[4];
    qm1[0] = q[0];
    qm1[1] = -1.0 * q[1];
    qm1[2] = -1.0 * q[2];
    qm1[3] = -1.0 * q[3];
    o_qmul(qm1, p, r);
}

/* largest |component| over the Grimmer candidates (SCORE_Math.cpp:604-608) */
double oracle_disori_omega(const double *p, const double *q) {
    double r[4], r0[6][4];
    o_misori(p, q, r);
    o_grimmer(r, r0);
    double omega = 0.0;
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 4; j++)
            if (fabs(r0[i][j]) > omega) omega = fabs(r0[i][j]);
    return omega;
}

/* disori_angle_oriquat_cubic_imm (SCORE_Math.cpp:558-622), radians */
double oracle_disori_angle(const double *p, const double *q) {
    double omega = oracle_disori_omega(p, q);
    if (omega < (1.0 - O_DOUBLE_ACCURACY)) return 2.0 * acos(omega);
    return 0.0;
}

/* disori_fromoriquat_cubic_imm (SCORE_Math.cpp:1067-1131) */
void oracle_disori_quat(const double *p, const double *q, double *quat) {
    double r[4], r0[6][4];
    o_misori(p, q, r);
    o_grimmer(r, r0);
    int mi = 0;
    double mx = 0.0;
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 4; j++)
            if (fabs(r0[i][j]) > mx) { mx = fabs(r0[i][j]); mi = i; }
    double rq[4] = {fabs(r0[mi][0]), fabs(r0[mi][1]), fabs(r0[mi][2]), fabs(r0[mi][3])};
    for (int i = 0; i < 3; i++) /* ascending sort of four values */
        for (int j = 0; j < 3 - i; j++)
            if (rq[j] > rq[j + 1]) { double tmp = rq[j]; rq[j] = rq[j + 1]; rq[j + 1] = tmp; }
    quat[0] = rq[3]; quat[1] = rq[2]; quat[2] = rq[1]; quat[3] = rq[0];
}

/* Mobility weight P of a nucleus/deformed pair: Rollett-Holm sigmoid
 * (SCORE_Kernel.cpp:2490-2496, SCORE_GrainBoundaryPhysics.cpp:110-135) or
 * Sebald-Gottstein classes (SCORE_Kernel.cpp:2498-2526, SCORE_GrainBoundaryPhysics.cpp:59-107). */
double oracle_mobility_weight(const score_gm_params *p, const double *qrx, const double *qdef) {
    double theta = oracle_disori_angle(qrx, qdef);
    if (p->mobility_model == SCORE_GM_MOB_ROLLETTHOLM) {
        return 1.0 - (p->rh_cut * exp(-1.0 * p->rh_trans * pow((theta / O_MAXDISORI_LAGB2HAGB), p->rh_exponent)));
    }
    if (theta <= O_MAXDISORI_LAGB2HAGB) return -1.0;
    double qdis[4] = {1.0, 0.0, 0.0, 0.0};
    oracle_disori_quat(qrx, qdef, qdis);
    double maxDev40_111 = O_MAXDISORI_TO_40DEG111;
    double _sqrt3 = 1 / sqrt(3.0);
    double oneNinth = 1.0 / 9.0;
    double m40_111[4] = {cos(oneNinth * O_PI), _sqrt3 * sin(oneNinth * O_PI), _sqrt3 * sin(oneNinth * O_PI),
                         _sqrt3 * sin(oneNinth * O_PI)};
    double dev = oracle_disori_angle(qdis, m40_111);
    double w = 0.0;
    if (dev < maxDev40_111) {
        double cc = cos(0.5 * O_PI * dev / maxDev40_111);
        w = cc * cc;
    }
    return w;
}

/* ---------------- per-step physics scalars ---------------- */

/* update_temperature (SCORE_Kernel.cpp:5318-5337) */
static void o_update_temperature(oracle *o) {
    uint32_t i = 0, n = o->p.n_sched;
    while (i < n && o->sched_t[i] < o->t) i++;
    if (i == 0) { o->T = o->sched_T[0]; return; }
    if (i >= n) { o->T = o->sched_T[n - 1]; return; }
    o->T = o->sched_T[i - 1] + ((o->sched_T[i] - o->sched_T[i - 1]) / (o->sched_t[i] - o->sched_t[i - 1])) * (o->t - o->sched_t[i - 1]);
}

/* get_shearmodulus, get_burgersvector (SCORE_GeneralPhysics.cpp:65-90);
 * update_atomisticproperties (SCORE_Kernel.cpp:5340-5352) */
static void o_update_atomistic(oracle *o) {
    double T = o->T;
    o->G = o->p.G0 * (-1.38e-9 * (T * T * T) + 1.21e-6 * (T * T) - 4.58e-4 * T + 1.00);
    o->b = o->p.b0;
    o->Gbhalfsq = 0.5 * o->G * (o->b * o->b);
}

/* update_intrinsicmobilities (SCORE_Kernel.cpp:5355-5368) */
static void o_update_mobilities(oracle *o) {
    double kT = O_KBOLTZMAN * o->T;
    o->mLAGB = o->p.lagb_m0 * exp(-1.0 * o->p.lagb_hact / kT);
    o->mHAGB = o->p.hagb_m0 * exp(-1.0 * o->p.hagb_hact / kT);
    o->mGS = o->p.gs_m0 * exp(-1.0 * o->p.gs_hact / kT);
    o->mRHHAGB = o->p.rh_hagb_m0 * exp(-1.0 * o->p.rh_hagb_hact / (O_KBOLTZMAN * o->T));
}

/* get_dtmax_instantslope_temp (SCORE_Kernel.cpp:5561-5591) */
static double o_dt_temp(const oracle *o) {
    uint32_t i = 0, n = o->p.n_sched;
    while (i < n && o->sched_t[i] < o->t) i++;
    if (i == 0) return O_INITIAL_DELTAT;
    if (i < n) {
        double dTdt = (fabs(o->sched_T[i] - o->sched_T[i - 1]) / (o->sched_t[i] - o->sched_t[i - 1]));
        if (dTdt <= O_SMALL_HEATRATE) dTdt = O_SMALL_HEATRATE;
        double dtmax = (O_SMALL_HEAT / dTdt);
        double dtintvl = o->sched_t[i] - o->sched_t[i - 1];
        if (dtmax >= dtintvl) dtmax = dtintvl;
        return dtmax;
    }
    return O_INFINITE;
}

/* get_dtmax_instantslope_cells (SCORE_Kernel.cpp:5519-5558) */
/* update_microchemistry (SCORE_Kernel.cpp:5370-5405): f/r constant (first entry)
 * or linear in time; past the table's end it keeps its last value */
static void o_update_microchemistry(oracle *o) {
    if (o->p.zener_mode == 0) {
        o->fr = 0.0;
        o->pzmin = 0.0;
        return;
    }
    if (o->p.zener_mode == 1) {
        o->fr = o->disp_fr[0];
        o->pzmin = o->p.zener_alpha * o->p.zener_gamma * o->fr;
        return;
    }
    uint32_t i = 0, n = o->p.n_disp;
    while (i < n && o->disp_t[i] < o->t) i++;
    if (i == 0) o->fr = 0.0; /* t <= 0: the reference reads before its table here */
    else if (i < n) {
        double dfrdt = ((o->disp_fr[i] - o->disp_fr[i - 1]) / (o->disp_t[i] - o->disp_t[i - 1]));
        o->fr = o->disp_fr[i - 1] + dfrdt * (o->t - o->disp_t[i - 1]);
    }
    o->pzmin = o->p.zener_alpha * o->p.zener_gamma * o->fr;
}

/* get_zener (SCORE_Kernel.cpp:5448-5456) */
static double o_get_zener(const oracle *o) {
    if (o->p.zener_mode == 0) return 0.0;
    return (1.0 * o->p.zener_alpha * o->p.zener_gamma * o->fr);
}

/* get_dtmax_instantslope_zener (SCORE_Kernel.cpp:5639-5666) */
static double o_dt_zener(const oracle *o) {
    if (o->p.zener_mode != 2) return O_INFINITE;
    uint32_t i = 0, n = o->p.n_disp;
    double dtmax = O_INITIAL_DELTAT;
    while (i < n && o->disp_t[i] < o->t) i++;
    if (i == 0) return dtmax;
    if (i < n) {
        double dpzdt = o->p.zener_alpha * o->p.zener_gamma *
                       (fabs(o->disp_fr[i] - o->disp_fr[i - 1]) / (o->disp_t[i] - o->disp_t[i - 1]));
        if (fabs(dpzdt) <= O_SMALL_DRAGGINGRATE) dpzdt = O_SMALL_DRAGGINGRATE;
        dtmax = (O_SMALL_ZENERFORCE / dpzdt);
        return dtmax;
    }
    return O_INFINITE;
}

/* update_deformedsubstructure (SCORE_Kernel.cpp:5408-5437) with the recovery
 * models of SCORE_RecoveryModels.cpp: Nes vacancy-core diffusion (:50-80) grows
 * r = rho^-1/2 of every grain over the step just done; solute drag (:83-86)
 * changes nothing; Michalak-Paxton (:89-125) scales rho0 by a function of t */
static void o_update_deformedsubstructure(oracle *o) {
    if (o->p.recovery_mode == 4) {
        /* exponential recovery (this build's own model, not in SCORE):
         * rho = rho0 * exp(-t / tau) for every deformed grain */
        const double factor = exp(-o->t / o->p.rec_tau);
        for (uint32_t dg = 0; dg < o->n_defg; dg++) o->rho[dg] = o->rho0[dg] * factor;
        return;
    }
    if (o->p.recovery_mode == 1) {
        double B = O_PI * o->p.rec_c3 * o->p.rec_c5 * o->p.rec_vacancy_d0 *
                   exp(-1.0 * o->p.rec_vacancy_hact / (O_KBOLTZMAN * o->T));
        B /= o->p.rec_kappa2;
        double exponent = (o->b * o->b) / (O_KBOLTZMAN * o->T);
        double expterm = exp(o->p.rec_alpha3 * o->p.rec_kappa2 * o->G * o->b * exponent);
        double currdt = o->dt;
        for (uint32_t dg = 0; dg < o->n_defg; dg++) {
            double r_t = 1.0 / pow(o->rho[dg], 0.5);
            double Bx = (B / r_t) * expterm;
            double r_tdt = (Bx * currdt) + r_t;
            o->rho[dg] = 1.0 / (r_tdt * r_tdt);
        }
    } else if (o->p.recovery_mode == 3) {
        double f1 = 0.25, f2 = 0.75, tau2 = 0.3605, tau5 = 5.2824;
        double currtime_minutes = o->t / 60.0;
        double factor = f1 * pow((1 + (currtime_minutes / tau2)), -1.0) + f2 * pow((1 + (currtime_minutes / tau5)), -0.5);
        for (uint32_t dg = 0; dg < o->n_defg; dg++) o->rho[dg] = o->rho0[dg] * (factor * factor);
    }
}

/* get_dtmax_instantslope_rho (SCORE_Kernel.cpp:5594-5636) */
static double o_dt_rho(const oracle *o) {
    double drdtmax, dtmax;
    if (o->p.recovery_mode == 1) {
        double Bx = O_PI * o->p.rec_c3 * o->p.rec_c5 * o->p.rec_vacancy_d0 *
                    exp(-1.0 * o->p.rec_vacancy_hact / (O_KBOLTZMAN * o->T));
        Bx = Bx / (o->p.rec_kappa2 * (1.0 / pow(o->rhomax, 0.5)));
        double exponent = (o->b * o->b) / (O_KBOLTZMAN * o->T);
        drdtmax = Bx * exp(o->p.rec_alpha3 * o->p.rec_kappa2 * o->G * o->b * exponent);
        if (drdtmax < (1.0 / pow(O_SMALL_RECOVERYRATE, 0.5))) drdtmax = (1.0 / pow(O_SMALL_RECOVERYRATE, 0.5));
        dtmax = (1.0 / pow(O_SMALL_RHO, 0.5)) / drdtmax;
        return dtmax;
    }
    if (o->p.recovery_mode == 2) {
        double Bsa = o->p.rec_c3 * o->b * o->p.rec_solute_d0 * exp(-1.0 * o->p.rec_solute_hact / (O_KBOLTZMAN * o->T));
        double exponent = (o->b * o->b) / (O_KBOLTZMAN * o->T);
        double ls = o->p.rec_solute_ls * pow(o->p.rec_solute_conc, (-2.0 / 3.0));
        drdtmax = Bsa * exp(o->p.rec_alpha3 * ls * o->G * o->b * exponent / (1.0 / (o->rhomax * o->rhomax)));
        if (drdtmax < (1.0 / pow(O_SMALL_RECOVERYRATE, 0.5))) drdtmax = (1.0 / pow(O_SMALL_RECOVERYRATE, 0.5));
        dtmax = (1.0 / pow(O_SMALL_RHO, 0.5)) / drdtmax;
        return dtmax;
    }
    return O_INFINITE;
}

static double o_dt_cells(const oracle *o) {
    double kT = O_KBOLTZMAN * o->T;
    if (o->p.mobility_model == SCORE_GM_MOB_ROLLETTHOLM) {
        double mmax = o->p.rh_hagb_m0 * exp(-1.0 * o->p.rh_hagb_hact / (O_KBOLTZMAN * o->T));
        double vmax = mmax * ((o->Gbhalfsq * o->rhomax) - o->pzmin);
        return ((o->p.maxfill * o->cellsize) / vmax);
    }
    double mgs = o->p.gs_m0 * exp(-1.0 * o->p.gs_hact / kT);
    double mhagb = o->p.hagb_m0 * exp(-1.0 * o->p.hagb_hact / kT);
    double m_max = (o->Pmax * mgs) + ((1.0 - o->Pmax) * mhagb);
    if (o->Pmax < (0.0 - O_DEFAULT_SMALL_NUMBER)) m_max = o->p.lagb_m0 * exp(-1.0 * o->p.lagb_hact / kT);
    double v_max = m_max * ((o->Gbhalfsq * o->rhomax) - o->pzmin);
    return ((o->p.maxfill * o->cellsize) / v_max);
}

/* get_dtmax_minimum (SCORE_Kernel.cpp:5687-5723); the nucleation rule returns
 * INFINITE for site saturation. */
static double o_dt_minimum(const oracle *o) {
    double mn = O_INFINITE;
    double d = o_dt_temp(o);
    if (d < mn) mn = d;
    double dr = o_dt_rho(o);
    if (dr < mn) mn = dr;
    double dz = o_dt_zener(o);
    if (dz < mn) mn = dz;
    double dm = o_dt_cells(o);
    if (dm < mn) mn = dm;
    return mn;
}

/* update_mydgoptimize_* + update_myrhomax (SCORE_RecoveryModels.cpp:128-165):
 * the largest rho among deformed grains that still own cells, floor 1e8 */
static void o_update_rhomax(oracle *o) {
    double rmax = O_RHOMAX_WELLANNEALED;
    double best = -1.0;
    int found = 0;
    for (uint32_t d = 0; d < o->n_defg; d++)
        if (o->defcnt[d] > 0 && (!found || o->rho[d] > best)) { best = o->rho[d]; found = 1; }
    if (found && best >= rmax) rmax = best;
    o->rhomax = rmax;
}

/* ---------------- front lists ---------------- */

static int o_list_init(o_list *L, uint32_t nslots, uint32_t rec_cap, uint32_t full_cap, int growable) {
    memset(L, 0, sizeof(*L));
    L->ntotal = nslots > 0 ? nslots : 1;
    L->c = (o_cell *)calloc(L->ntotal, sizeof(o_cell));
    L->rec_cap = rec_cap;
    L->rec = (uint32_t *)malloc(sizeof(uint32_t) * (rec_cap > 0 ? rec_cap : 1));
    L->full_cap = full_cap > 0 ? full_cap : 1;
    L->full = (uint32_t *)malloc(sizeof(uint32_t) * L->full_cap);
    L->growable = growable;
    if (!L->c || !L->rec || !L->full) return -1;
    for (uint32_t i = 0; i < L->ntotal; i++) {
        L->c[i].act = 0; L->c[i].inf = 26; L->c[i].x = L->c[i].y = L->c[i].z = -1;
        L->c[i].frac = 0.0; L->c[i].P = -1.0; L->c[i].defg = 4096000001u; L->c[i].rxg = 4096000001u;
    }
    return 0;
}

static void o_list_free(o_list *L) {
    free(L->c); free(L->rec); free(L->full);
    memset(L, 0, sizeof(*L));
}

/* omp_getNextFreeSlotInRXFrontInside / Border with CELLRECYCLE (SCORE_Kernel.cpp:3757-3874);
 * CELLAPPEND skips the recycling branch. Inside grows; Border returns INVALID_ADDRESS when full. */
static uint32_t o_next_free(o_list *L, int recycle) {
    if (recycle && L->rec_next < L->rec_first_not) return L->rec[L->rec_next++];
    if (L->nnever < L->ntotal) return L->nnever++;
    if (!L->growable) return O_INVALID_ADDRESS;
    uint32_t nsz = L->ntotal * 2u;
    o_cell *nc = (o_cell *)realloc(L->c, sizeof(o_cell) * nsz);
    if (!nc) return O_INVALID_ADDRESS;
    L->c = nc;
    for (uint32_t i = L->ntotal; i < nsz; i++) {
        L->c[i].act = 0; L->c[i].inf = 26; L->c[i].x = L->c[i].y = L->c[i].z = -1;
        L->c[i].frac = 0.0; L->c[i].P = -1.0; L->c[i].defg = 4096000001u; L->c[i].rxg = 4096000001u;
    }
    L->ntotal = nsz;
    return L->nnever++;
}

/* append_to_fullrx_*_list (SCORE_Kernel.cpp:3994-4030, :4133-4170): grows */
static int o_append_full(o_list *L, uint32_t c) {
    if (L->full_n >= L->full_cap) {
        uint32_t nsz = L->full_cap * 2u;
        uint32_t *nf = (uint32_t *)realloc(L->full, sizeof(uint32_t) * nsz);
        if (!nf) return -1;
        L->full = nf;
        L->full_cap = nsz;
    }
    L->full[L->full_n++] = c;
    return 0;
}

/* append_to_recycling_*_list (SCORE_Kernel.cpp:4034-4042, :4173-4180): silently capped */
static void o_append_rec(o_list *L, uint32_t c) {
    if (L->rec_first_not < L->rec_cap) L->rec[L->rec_first_not++] = c;
}

/* ompshar_sim_myCA_calcGrowthStepInside / Border (SCORE_Kernel.cpp:4046-4130, :4183-4249) */
static int o_sweep(oracle *o, o_list *L) {
    uint32_t active = 0;
    L->full_n = 0;
    L->rec_next = 0;
    L->rec_first_not = 0;
    for (uint32_t c = 0; c < L->nnever; c++) {
        o_cell *e = &L->c[c];
        if (e->act) {
            active++;
            double m;
            if (o->p.mobility_model == SCORE_GM_MOB_ROLLETTHOLM) m = e->P * o->mRHHAGB;
            else if (e->P < 0.0) m = o->mLAGB;
            else m = (e->P * o->mGS) + ((1.0 - e->P) * o->mHAGB);
            double p = (o->Gbhalfsq * (o->rho[e->defg] - O_RHO_RX)) - o_get_zener(o);
            if (p > O_DOUBLE_ACCURACY) {
                double v = m * p;
                v *= o_shape[e->inf];
                e->frac = e->frac + (v * o->dt * o->inv_dx);
                if (e->frac >= 1.0)
                    if (o_append_full(L, c)) return SCORE_GM_ERR_OOM;
            }
            continue;
        }
        o_append_rec(L, c);
    }
    L->sv = active;
    return 0;
}

static void o_write_cell(o_cell *d, const o_cell *parent, uint8_t code, int x, int y, int z, double carry,
                         uint32_t defg_t, double Pnew) {
    d->act = 1;
    d->inf = code;
    d->x = (int16_t)x; d->y = (int16_t)y; d->z = (int16_t)z;
    d->frac = carry;
    d->P = Pnew;
    d->rxg = parent->rxg;
    d->defg = defg_t;
}

static double o_pair_weight(oracle *o, uint32_t rxg, uint32_t defg) {
    return oracle_mobility_weight(&o->p, &o->rxq[4 * (size_t)rxg], &o->defq[4 * (size_t)defg]);
}

/* neighbour directions in the order of mynborregions (ompshar_init_neighbortopology,
 * SCORE_Kernel.cpp:2028-2080): myself, six faces, the x-z, x-y and y-z edges, the
 * corners; o_nb_opp[nb] is the opposite direction (the halo candidate that
 * ompshar_link_haloregions looks up, SCORE_GridPartitioning.cpp:237-276) */
static const int8_t o_nb_dir[27][3] = {
    {0, 0, 0},
    {-1, 0, 0}, {1, 0, 0}, {0, -1, 0}, {0, 1, 0}, {0, 0, -1}, {0, 0, 1},
    {-1, 0, -1}, {1, 0, -1}, {-1, 0, 1}, {1, 0, 1},
    {-1, -1, 0}, {-1, 1, 0}, {1, -1, 0}, {1, 1, 0},
    {0, -1, -1}, {0, -1, 1}, {0, 1, -1}, {0, 1, 1},
    {-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {1, 1, -1}, {-1, -1, 1}, {1, -1, 1}, {-1, 1, 1}, {1, 1, 1}};
static const uint8_t o_nb_opp[27] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15,
                                     26, 25, 24, 23, 22, 21, 20, 19};

static int o_in_box(const uint32_t *lo, const uint32_t *hi, int x, int y, int z) {
    return x >= (int)lo[0] && x <= (int)hi[0] && y >= (int)lo[1] && y <= (int)hi[1] && z >= (int)lo[2] && z <= (int)hi[2];
}

static int o_strictly_inside(const o_region *R, int x, int y, int z) {
    return x > (int)R->lo[0] && x < (int)R->hi[0] && y > (int)R->lo[1] && y < (int)R->hi[1] && z > (int)R->lo[2] &&
           z < (int)R->hi[2];
}

/* a new front record in list dst for the deformed cell (x,y,z) of grain st */
static int o_claim_into(oracle *o, o_list *dst, const o_cell *par, uint8_t code, int x, int y, int z, double carry,
                        uint32_t st, double Pn) {
    uint64_t cxyz = (uint64_t)x + (uint64_t)y * o->nx + (uint64_t)z * o->nxy;
    o->grid[cxyz] = O_CURRENTLY_INFECTED;
    uint32_t slot = o_next_free(dst, 1);
    if (slot == O_INVALID_ADDRESS) return SCORE_GM_ERR_CAPACITY;
    if (Pn >= o->reg_Pmax) o->reg_Pmax = Pn;
    o_write_cell(&dst->c[slot], par, code, x, y, z, carry, st, Pn);
    o->last_infected++;
    return 0;
}

/* an infection of region R directed at (x,y,z) in region nbortid: parked in R's
 * halo towards it unless that halo cell was used before (SCORE_Kernel.cpp:4494-4555) */
static void o_halo_attempt(o_region *R, uint32_t nbortid, int x, int y, int z, uint8_t code, double carry, uint32_t rxg) {
    for (uint32_t h = 1; h < 27; h++) {
        o_halo *H = &R->halo[h];
        if (H->inregion != nbortid || !H->status) continue;
        /* the reference's bounds reach one cell beyond the halo (:4506-4511) */
        if (x < (int)H->org[0] || x > (int)(H->org[0] + H->ext[0]) || y < (int)H->org[1] ||
            y > (int)(H->org[1] + H->ext[1]) || z < (int)H->org[2] || z > (int)(H->org[2] + H->ext[2]))
            continue;
        const uint32_t hx = (uint32_t)x - H->org[0], hy = (uint32_t)y - H->org[1], hz = (uint32_t)z - H->org[2];
        const uint32_t hid = hx + hy * H->ext[0] + hz * H->ext[0] * H->ext[1];
        if (H->status[hid]) return;
        H->status[hid] = 1;
        H->inf[hid] = code;
        H->xyz[3 * (size_t)hid] = (int16_t)x; H->xyz[3 * (size_t)hid + 1] = (int16_t)y; H->xyz[3 * (size_t)hid + 2] = (int16_t)z;
        H->carry[hid] = carry;
        H->rxg[hid] = rxg;
        H->refs[H->nref++] = hid;
        return;
    }
}

/* one attempt; omp_infect_OutOfRXFrontInside (SCORE_Kernel.cpp:4252-4344) when
 * from_border == 0 and omp_infect_OutOfRXFrontBorder (:4347-4556) when 1, for
 * region R */
static int o_infect(oracle *o, o_region *R, o_list *src, uint32_t pslot, int from_border, int intoborder, const o_dir *d,
                    double carry) {
    const o_cell par = src->c[pslot];
    int x = par.x + d->dx, y = par.y + d->dy, z = par.z + d->dz;
    o_list *dst;
    if (from_border) {
        if (x < 0) x += (int)o->nx;
        if (x >= (int)o->nx) x -= (int)o->nx;
        if (y < 0) y += (int)o->ny;
        if (y >= (int)o->ny) y -= (int)o->ny;
        if (z < 0) return 0;
        if (z >= (int)o->nz) return 0;
        if (o_strictly_inside(R, x, y, z)) {
            dst = &R->in;
        } else {
            /* the region holding the target, neighbours in mynborregions order (:4438-4447) */
            uint32_t nbortid = O_NO_HALO;
            for (uint32_t nb = 0; nb < 27; nb++) {
                const o_region *Q = &o->reg[R->nbor[nb]];
                if (o_in_box(Q->lo, Q->hi, x, y, z)) { nbortid = R->nbor[nb]; break; }
            }
            if (nbortid != R->nbor[0]) {
                o_halo_attempt(R, nbortid, x, y, z, (uint8_t)d->code, carry, par.rxg);
                return 0;
            }
            dst = &R->bd;
        }
    } else {
        dst = intoborder ? &R->bd : &R->in;
    }
    uint64_t cxyz = (uint64_t)x + (uint64_t)y * o->nx + (uint64_t)z * o->nxy;
    uint32_t st = o->grid[cxyz];
    if (st >= o->n_defg || st == O_CURRENTLY_INFECTED) return 0;
    double Pn = par.P;
    if (par.defg != st) Pn = o_pair_weight(o, par.rxg, st);
    return o_claim_into(o, dst, &par, (uint8_t)d->code, x, y, z, carry, st, Pn);
}

/* ompshar_sim_myCA_updateFullRXInside / Border (SCORE_Kernel.cpp:4559-4664, :4678-4774) */
static int o_update_fullrx(oracle *o, o_region *R, o_list *L, int from_border) {
    const double overFac1 = 0.0682, overFac2 = 0.0542, overFac3 = 0.0484;
    uint32_t nupd = 0;
    for (uint32_t k = 0; k < L->full_n; k++) {
        uint32_t c = L->full[k];
        uint32_t rgpid = L->c[c].rxg;
        double overshoot = L->c[c].frac - 1.0;
        double carry[4];
        carry[1] = overshoot * overFac1;
        carry[2] = overshoot * overFac2;
        carry[3] = overshoot * overFac3;
        int intoborder = 1;
        if (!from_border) {
            int x = L->c[c].x, y = L->c[c].y, z = L->c[c].z;
            if (x > (int)(R->lo[0] + 1) && x < (int)(R->hi[0] - 1) && y > (int)(R->lo[1] + 1) && y < (int)(R->hi[1] - 1) &&
                z > (int)(R->lo[2] + 1) && z < (int)(R->hi[2] - 1))
                intoborder = 0;
        }
        for (int dd = 0; dd < 26; dd++) {
            int rc = o_infect(o, R, L, c, from_border, intoborder, &o_dirs[dd], carry[o_dirs[dd].carry]);
            if (rc) return rc;
        }
        uint64_t cxyz = (uint64_t)L->c[c].x + (uint64_t)L->c[c].y * o->nx + (uint64_t)L->c[c].z * o->nxy;
        o->grid[cxyz] = o->n_defg + rgpid;
        o->rxcnt[L->c[c].rxg] += 1;
        o->defcnt[L->c[c].defg] -= 1;
        L->c[c].act = 0;
        L->c[c].frac = 0.0;
        nupd++;
    }
    L->nupdated = nupd;
    return 0;
}

/* ompshar_synchronize_haloregions_rx (SCORE_Kernel.cpp:4777-4872): region R applies
 * the infections its neighbours parked in their halos towards it, neighbour by
 * neighbour in its mynborregions order, each halo in the order of its refs, into
 * its Border list, with the weight recomputed; a cell no longer deformed is skipped */
static int o_sync_halos(oracle *o, uint32_t r) {
    o_region *R = &o->reg[r];
    for (uint32_t nb = 0; nb < 27; nb++) {
        o_region *A = &o->reg[R->nbor[nb]];
        const uint32_t hal = A->nborhalo[nb];
        if (hal == O_NO_HALO) continue;
        const o_halo *H = &A->halo[hal];
        for (uint32_t s2 = 0; s2 < H->nref; s2++) {
            const uint32_t hid = H->refs[s2];
            const int x = H->xyz[3 * (size_t)hid], y = H->xyz[3 * (size_t)hid + 1], z = H->xyz[3 * (size_t)hid + 2];
            uint64_t cxyz = (uint64_t)x + (uint64_t)y * o->nx + (uint64_t)z * o->nxy;
            uint32_t st = o->grid[cxyz];
            if (st >= o->n_defg || st == O_CURRENTLY_INFECTED) continue;
            o_cell par;
            memset(&par, 0, sizeof(par));
            par.rxg = H->rxg[hid];
            const double Pn = o_pair_weight(o, H->rxg[hid], st);
            int rc = o_claim_into(o, &R->bd, &par, H->inf[hid], x, y, z, H->carry[hid], st, Pn);
            if (rc) return rc;
        }
    }
    return 0;
}

/* ompshar_defragmentation (SCORE_Kernel.cpp:3877-3934): stable compaction of ACTIVE slots.
 * Slots past the compacted block keep their previous contents, as in the reference. */
static void o_defrag_list(o_list *L) {
    uint32_t first = 0;
    while (first < L->ntotal && L->c[first].act) first++;
    uint32_t contiguous = first;
    for (uint32_t c = first + 1; c < L->nnever; c++)
        if (L->c[c].act) L->c[contiguous++] = L->c[c];
    L->nnever = contiguous;
}

/* ---------------- set-up ---------------- */

static int o_validate(const score_gm_domain *d, const score_gm_params *p) {
    if (!d || !p) return SCORE_GM_ERR_ARG;
    if (d->nx < 3 || d->ny < 3 || d->nz < 3) return SCORE_GM_ERR_ARG;
    if (d->nx > 1600 || d->ny > 1600 || d->nz > 1600) return SCORE_GM_ERR_ARG;
    if (!d->cellgrid || !d->ori_quat || !d->defg_ori || !d->defg_rho || d->n_defg == 0) return SCORE_GM_ERR_ARG;
    if (d->n_rxg && (!d->rxg_ori || !d->rxg_site)) return SCORE_GM_ERR_ARG;
    if (d->cellsize <= 0.0) return SCORE_GM_ERR_ARG;
    if (p->mobility_model != SCORE_GM_MOB_ROLLETTHOLM && p->mobility_model != SCORE_GM_MOB_SEBALDGOTTSTEIN)
        return SCORE_GM_ERR_ARG;
    if (p->zener_mode < 0 || p->zener_mode > 2) return SCORE_GM_ERR_ARG;
    if (p->zener_mode != 0) {
        if (p->n_disp < (p->zener_mode == 2 ? 2u : 1u) || !p->disp_time || !p->disp_fr) return SCORE_GM_ERR_ARG;
        if (p->zener_mode == 2) {
            if (p->disp_time[0] != 0.0) return SCORE_GM_ERR_ARG;
            for (uint32_t i = 1; i < p->n_disp; i++)
                if (!(p->disp_time[i] > p->disp_time[i - 1])) return SCORE_GM_ERR_ARG;
        }
    }
    if (p->recovery_mode < 0 || p->recovery_mode > 4) return SCORE_GM_ERR_ARG;
    if (p->n_regions < 0 || p->n_regions > 1024) return SCORE_GM_ERR_ARG;
    if (p->n_regions > 1) {
        /* every region more than MINIMUM_REGION_SIZE cells thick (SCORE_GridPartitioning.cpp:87-93) */
        const uint32_t npy = oracle_smallest_y_factor((uint32_t)p->n_regions);
        const uint32_t npz = (uint32_t)p->n_regions / npy;
        if (d->ny / npy <= O_MINIMUM_REGION_SIZE || d->nz / npz <= O_MINIMUM_REGION_SIZE) return SCORE_GM_ERR_ARG;
    }
    if (p->recovery_mode == 4 && !(p->rec_tau > 0.0)) return SCORE_GM_ERR_ARG;
    if (p->n_grainlog && !p->grainlog_at) return SCORE_GM_ERR_ARG;
    if (p->n_sched < 2 || !p->sched_time || !p->sched_temp) return SCORE_GM_ERR_ARG;
    /* the range SCORE accepts for NMAX (SCORE_Kernel.cpp:274) */
    if (p->nmax < 1 || p->nmax > O_DEFAULT_NMAX) return SCORE_GM_ERR_ARG;
    if (p->n_defrag && !p->defrag_at) return SCORE_GM_ERR_ARG;
    uint64_t N = (uint64_t)d->nx * d->ny * d->nz;
    for (uint32_t i = 0; i < d->n_defg; i++)
        if (d->defg_ori[i] >= d->n_ori) return SCORE_GM_ERR_ARG;
    for (uint32_t i = 0; i < d->n_rxg; i++)
        if (d->rxg_ori[i] >= d->n_ori || d->rxg_site[i] >= N) return SCORE_GM_ERR_ARG;
    for (uint64_t i = 0; i < N; i++)
        if (d->cellgrid[i] >= d->n_defg) return SCORE_GM_ERR_ARG;
    return 0;
}

/* mathMethods::smallestYFactor (SCORE_Math.cpp:97-115): the y factor of the
 * region grid, the last of the listed primes that divides n before the next
 * prime reaches the cofactor */
uint32_t oracle_smallest_y_factor(uint32_t n) {
    static const uint32_t primes[32] = {1, 2, 3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47,
                                        53, 59, 61, 67, 71, 73, 79, 83, 89, 97, 101, 103, 107, 109, 113, 127};
    uint32_t z = n, y = 1;
    for (uint32_t c = 0; c < 31; c++) {
        const uint32_t q = n / primes[c];
        if (q * primes[c] == n) { z = q; y = primes[c]; }
        if (primes[c + 1] >= z) break;
    }
    return y;
}

static void o_regions_free(oracle *o) {
    if (!o->reg) return;
    for (uint32_t r = 0; r < o->nreg; r++) {
        o_region *R = &o->reg[r];
        o_list_free(&R->in);
        o_list_free(&R->bd);
        for (uint32_t h = 0; h < 27; h++) {
            o_halo *H = &R->halo[h];
            free(H->status); free(H->inf); free(H->xyz); free(H->carry); free(H->rxg); free(H->refs);
        }
    }
    free(o->reg);
    o->reg = NULL;
}

/* the region grid of ompshar_ca2regions (SCORE_GridPartitioning.cpp:69-166):
 * npy = smallestYFactor(T) slabs along y, npz = T / npy along z, x not split;
 * region z*npy + y spans [y*nry, y*nry + nry - 1] (the last one up to ny - 1)
 * and likewise in z.  Then its neighbour table (ompshar_init_neighbortopology,
 * SCORE_Kernel.cpp:2028-2080; get_nbortid, :2002-2026, periodic on the region
 * grid), its halos (ompshar_init_haloregions, :2223-2457; determine_origins,
 * :2134-2220) and the halo each neighbour keeps towards it
 * (ompshar_link_haloregions, SCORE_GridPartitioning.cpp:222-276). */
static int o_regions_init(oracle *o, uint32_t nreg) {
    o->nreg = nreg;
    o->npy = oracle_smallest_y_factor(nreg);
    o->npz = nreg / o->npy;
    o->reg = (o_region *)calloc(nreg, sizeof(o_region));
    if (!o->reg) return SCORE_GM_ERR_OOM;
    const uint32_t npy = o->npy, npz = o->npz;
    const uint32_t nry = o->ny / npy, nrz = o->nz / npz;
    for (uint32_t iz = 0; iz < npz; iz++)
        for (uint32_t iy = 0; iy < npy; iy++) {
            o_region *R = &o->reg[iz * npy + iy];
            R->lo[0] = 0; R->hi[0] = o->nx - 1;
            R->lo[1] = iy * nry; R->hi[1] = iy * nry + nry - 1;
            R->lo[2] = iz * nrz; R->hi[2] = iz * nrz + nrz - 1;
            if (iy == npy - 1) R->hi[1] = o->ny - 1;
            if (iz == npz - 1) R->hi[2] = o->nz - 1;
        }
    for (uint32_t r = 0; r < nreg; r++) {
        o_region *R = &o->reg[r];
        const int iy = (int)(r % npy), iz = (int)(r / npy);
        uint32_t ext[3];
        for (int a = 0; a < 3; a++) ext[a] = R->hi[a] - R->lo[a] + 1;
        for (uint32_t nb = 0; nb < 27; nb++) {
            int jy = iy + o_nb_dir[nb][1], jz = iz + o_nb_dir[nb][2];
            if (jy < 0) jy += (int)npy;
            if (jy >= (int)npy) jy -= (int)npy;
            if (jz < 0) jz += (int)npz;
            if (jz >= (int)npz) jz -= (int)npz;
            R->nbor[nb] = (uint32_t)jz * npy + (uint32_t)jy;
        }
        for (uint32_t h = 0; h < 27; h++) {
            o_halo *H = &R->halo[h];
            H->inregion = R->nbor[h];
            if (h == 0 || R->nbor[h] == r) continue; /* a dummy: extent 0 */
            const uint32_t dims[3] = {o->nx, o->ny, o->nz};
            for (int a = 0; a < 3; a++) {
                const int d = o_nb_dir[h][a];
                int org;
                if (d == -1) { org = (int)R->lo[a] - 1; if (org < 0) org += (int)dims[a]; }
                else if (d == 0) org = (int)R->lo[a];
                else { org = (int)R->hi[a] + 1; if (org >= (int)dims[a]) org -= (int)dims[a]; }
                H->org[a] = (uint32_t)org;
                H->ext[a] = d == 0 ? ext[a] : 1;
            }
            const size_t n = (size_t)H->ext[0] * H->ext[1] * H->ext[2];
            H->status = (uint8_t *)calloc(n, 1);
            H->inf = (uint8_t *)malloc(n);
            H->xyz = (int16_t *)malloc(3 * n * sizeof(int16_t));
            H->carry = (double *)malloc(n * sizeof(double));
            H->rxg = (uint32_t *)malloc(n * sizeof(uint32_t));
            H->refs = (uint32_t *)malloc(n * sizeof(uint32_t));
            if (!H->status || !H->inf || !H->xyz || !H->carry || !H->rxg || !H->refs) return SCORE_GM_ERR_OOM;
        }
    }
    /* region R + d keeps its halo opp(d) towards R whenever d moves off R on the
     * region grid (exists_nborhalo, SCORE_GridPartitioning.cpp:187-219) */
    for (uint32_t r = 0; r < nreg; r++) {
        o_region *R = &o->reg[r];
        R->nborhalo[0] = O_NO_HALO;
        for (uint32_t nb = 1; nb < 27; nb++) {
            const o_region *Q = &o->reg[R->nbor[nb]];
            R->nborhalo[nb] = Q->halo[o_nb_opp[nb]].status ? o_nb_opp[nb] : O_NO_HALO;
        }
    }
    return 0;
}

void oracle_destroy(oracle *o) {
    if (!o) return;
    free(o->grid); free(o->defq); free(o->rxq); free(o->rho); free(o->rho0); free(o->site);
    free(o->defcnt); free(o->rxcnt); free(o->sched_t); free(o->sched_T); free(o->defrag_at);
    free(o->disp_t); free(o->disp_fr);
    o_regions_free(o);
    free(o->traj);
    free(o->arr_ids); free(o->arr_off);
    free(o->glog_at); free(o->glog); free(o->glog_counts);
    free(o);
}

/* infect_JuvenileNucleation (SCORE_GrowthMachine.cpp:124-189): the nucleus takes
 * its site unless the cell is no longer deformed; its record (rxFrac 1.0,
 * infector 26) is appended to the Inside list for a strictly interior site,
 * else to the Border list */
static int o_juvenile_nucleation(oracle *o, uint32_t nuc, int *placed) {
    *placed = 0;
    uint32_t w = o->site[nuc];
    uint32_t z = w / o->nxy;
    uint32_t rem = w - (o->nxy * z);
    uint32_t y = rem / o->nx;
    uint32_t x = rem - (o->nx * y);
    /* the region whose box holds the site (ompshar_sim_myCA_sitesaturatedNucleation,
     * SCORE_GrowthMachine.cpp:75-100) */
    const uint32_t ry = o->ny / o->npy, rz = o->nz / o->npz;
    uint32_t iy = y / ry, iz = z / rz;
    if (iy >= o->npy) iy = o->npy - 1;
    if (iz >= o->npz) iz = o->npz - 1;
    o_region *R = &o->reg[iz * o->npy + iy];
    int inside = (x > R->lo[0] && x < R->hi[0] && y > R->lo[1] && y < R->hi[1] && z > R->lo[2] && z < R->hi[2]);
    uint32_t st = o->grid[w];
    if (st >= o->n_defg || st == O_CURRENTLY_INFECTED) return 0;
    o->grid[w] = O_CURRENTLY_INFECTED;
    double PP = o_pair_weight(o, nuc, st);
    o_list *L = inside ? &R->in : &R->bd;
    uint32_t slot = o_next_free(L, 0);
    if (slot == O_INVALID_ADDRESS) return SCORE_GM_ERR_CAPACITY;
    o_cell *e = &L->c[slot];
    e->act = 1; e->inf = 26; e->x = (int16_t)x; e->y = (int16_t)y; e->z = (int16_t)z;
    e->frac = 1.0; e->defg = st; e->rxg = nuc; e->P = PP;
    if (PP >= o->reg_Pmax) o->reg_Pmax = PP;
    o->nuclei++;
    *placed = 1;
    return 0;
}

/* solve_RXGROWTH set-up (SCORE_Kernel.cpp:6497-6612) including site-saturated
 * nucleation (SCORE_GrowthMachine.cpp:50-189) */
int oracle_create(const score_gm_domain *d, const score_gm_params *p, oracle **out) {
    int rc = o_validate(d, p);
    if (rc) return rc;
    oracle *o = (oracle *)calloc(1, sizeof(oracle));
    if (!o) return SCORE_GM_ERR_OOM;
    o->nx = d->nx; o->ny = d->ny; o->nz = d->nz; o->nxy = d->nx * d->ny;
    o->N = (uint64_t)d->nx * d->ny * d->nz;
    o->cellsize = d->cellsize;
    o->inv_dx = (1.0 / d->cellsize);
    o->n_defg = d->n_defg; o->n_rxg = d->n_rxg;
    o->p = *p;
    o->grid = (uint32_t *)malloc(sizeof(uint32_t) * o->N);
    o->defq = (double *)malloc(sizeof(double) * 4 * (size_t)o->n_defg);
    o->rxq = (double *)malloc(sizeof(double) * 4 * (size_t)(o->n_rxg + 1));
    o->rho = (double *)malloc(sizeof(double) * o->n_defg);
    o->rho0 = (double *)malloc(sizeof(double) * o->n_defg);
    o->site = (uint32_t *)malloc(sizeof(uint32_t) * (o->n_rxg + 1));
    o->defcnt = (uint32_t *)calloc(o->n_defg, sizeof(uint32_t));
    o->rxcnt = (uint32_t *)calloc(o->n_rxg + 1, sizeof(uint32_t));
    o->sched_t = (double *)malloc(sizeof(double) * p->n_sched);
    o->sched_T = (double *)malloc(sizeof(double) * p->n_sched);
    o->defrag_at = (double *)malloc(sizeof(double) * (p->n_defrag + 1));
    o->disp_t = (double *)malloc(sizeof(double) * (p->n_disp + 1));
    o->disp_fr = (double *)malloc(sizeof(double) * (p->n_disp + 1));
    if (!o->grid || !o->defq || !o->rxq || !o->rho || !o->rho0 || !o->site || !o->defcnt || !o->rxcnt || !o->sched_t ||
        !o->sched_T || !o->defrag_at || !o->disp_t || !o->disp_fr) {
        oracle_destroy(o);
        return SCORE_GM_ERR_OOM;
    }
    memcpy(o->grid, d->cellgrid, sizeof(uint32_t) * o->N);
    for (uint32_t i = 0; i < o->n_defg; i++) {
        memcpy(&o->defq[4 * (size_t)i], &d->ori_quat[4 * (size_t)d->defg_ori[i]], 4 * sizeof(double));
        o->rho[i] = d->defg_rho[i];
        o->rho0[i] = d->defg_rho[i];
    }
    for (uint32_t i = 0; i < o->n_rxg; i++) {
        memcpy(&o->rxq[4 * (size_t)i], &d->ori_quat[4 * (size_t)d->rxg_ori[i]], 4 * sizeof(double));
        o->site[i] = d->rxg_site[i];
    }
    memcpy(o->sched_t, p->sched_time, sizeof(double) * p->n_sched);
    memcpy(o->sched_T, p->sched_temp, sizeof(double) * p->n_sched);
    if (p->n_defrag) memcpy(o->defrag_at, p->defrag_at, sizeof(double) * p->n_defrag);
    o->p.sched_time = o->sched_t;
    o->p.sched_temp = o->sched_T;
    o->p.defrag_at = o->defrag_at;
    if (p->zener_mode != 0) {
        memcpy(o->disp_t, p->disp_time, sizeof(double) * p->n_disp);
        memcpy(o->disp_fr, p->disp_fr, sizeof(double) * p->n_disp);
    }
    o->p.disp_time = o->disp_t;
    o->p.disp_fr = o->disp_fr;
    o->fr = 0.0;
    o->glog_at = (double *)malloc(sizeof(double) * (p->n_grainlog + 1));
    if (!o->glog_at) { oracle_destroy(o); return SCORE_GM_ERR_OOM; }
    if (p->n_grainlog) memcpy(o->glog_at, p->grainlog_at, sizeof(double) * p->n_grainlog);
    o->p.grainlog_at = o->glog_at;
    /* nuclei arriving at a later step, grouped by step in pool order */
    o->arr_last = 0;
    if (d->rxg_step)
        for (uint32_t i = 0; i < o->n_rxg; i++)
            if (d->rxg_step[i] > o->arr_last) o->arr_last = d->rxg_step[i];
    if (o->arr_last) {
        o->arr_off = (uint32_t *)calloc((size_t)o->arr_last + 2, sizeof(uint32_t));
        o->arr_ids = (uint32_t *)malloc(sizeof(uint32_t) * (o->n_rxg + 1));
        if (!o->arr_off || !o->arr_ids) { oracle_destroy(o); return SCORE_GM_ERR_OOM; }
        for (uint32_t i = 0; i < o->n_rxg; i++) o->arr_off[d->rxg_step[i] + 1]++;
        for (uint32_t k = 0; k <= o->arr_last; k++) o->arr_off[k + 1] += o->arr_off[k];
        uint32_t *fillp = (uint32_t *)malloc(sizeof(uint32_t) * ((size_t)o->arr_last + 1));
        if (!fillp) { oracle_destroy(o); return SCORE_GM_ERR_OOM; }
        memcpy(fillp, o->arr_off, sizeof(uint32_t) * ((size_t)o->arr_last + 1));
        for (uint32_t i = 0; i < o->n_rxg; i++) o->arr_ids[fillp[d->rxg_step[i]]++] = i;
        free(fillp);
    }
    /* init_processing (SCORE_Kernel.cpp:5888-5901) */
    o->tprocessingend = 0.0;
    for (uint32_t i = 0; i < p->n_sched; i++)
        if (o->sched_t[i] >= o->tprocessingend) o->tprocessingend = o->sched_t[i];
    /* counts after solve_REPLACE_CA_STRUCTURE (SCORE_Kernel.cpp:6390-6420) */
    for (uint64_t i = 0; i < o->N; i++) o->defcnt[o->grid[i]]++;

    /* solve_RXGROWTH prologue (SCORE_Kernel.cpp:6501-6530) */
    o->t = 0.0; o->X = 0.0; o->Xcells = 0.0; o->step = 0;
    o->Pmax = O_DEFAULT_PMAX;
    o->reg_Pmax = O_DEFAULT_PMAX;
    o->pzmin = 0.0;
    o_update_rhomax(o);
    o_update_temperature(o);
    o_update_atomistic(o);
    o_update_mobilities(o);
    o->dt = o_dt_minimum(o);

    if ((rc = o_regions_init(o, p->n_regions > 1 ? (uint32_t)p->n_regions : 1u))) { oracle_destroy(o); return rc; }
    /* ompshar_init_myRXFrontInside / Border (SCORE_Kernel.cpp:3587-3710), per region
     * from its own volume and faces */
    for (uint32_t r = 0; r < o->nreg; r++) {
        o_region *R = &o->reg[r];
        const uint32_t ex = R->hi[0] - R->lo[0] + 1, ey = R->hi[1] - R->lo[1] + 1, ez = R->hi[2] - R->lo[2] + 1;
        const double vol = (double)(ex * ey * ez);
        uint32_t nin = (uint32_t)(p->rel_cell_caching * vol);
        uint32_t nin_rec = (uint32_t)(p->rel_cell_caching * vol * p->maxfill * O_FULLRECYCLING_ATT_FACTOR);
        uint32_t nin_full = (uint32_t)(p->rel_cell_caching * vol * p->maxfill * O_FULLRXCACHING_ATT_FACTOR);
        uint32_t nbd = 2 * ((ex * ey) + (ex * ez) + (ey * ez));
        uint32_t nbd_rec = (uint32_t)((double)nbd * p->maxfill * O_FULLRECYCLING_ATT_FACTOR);
        uint32_t nbd_full = (uint32_t)((double)nbd * p->maxfill * O_FULLRXCACHING_ATT_FACTOR);
        if (o_list_init(&R->in, nin, nin_rec, nin_full, 1) || o_list_init(&R->bd, nbd, nbd_rec, nbd_full, 0)) {
            oracle_destroy(o);
            return SCORE_GM_ERR_OOM;
        }
    }

    /* ompshar_sim_myCA_sitesaturatedNucleation + infect_JuvenileNucleation (SCORE_GrowthMachine.cpp:50-189):
     * the nuclei that arrive at step 0 */
    for (uint32_t nuc = 0; nuc < o->n_rxg; nuc++) {
        if (d->rxg_step && d->rxg_step[nuc] != 0) continue;
        int placed = 0;
        if ((rc = o_juvenile_nucleation(o, nuc, &placed))) { oracle_destroy(o); return rc; }
    }
    /* master reduction of myMobilityWeightMax (SCORE_Kernel.cpp:6598-6605) */
    if (o->reg_Pmax >= o->Pmax) o->Pmax = o->reg_Pmax;
    *out = o;
    return 0;
}

static int o_push_traj(oracle *o, const score_gm_traj_row *r) {
    if (o->ntraj == o->captraj) {
        size_t nc = o->captraj ? 2 * o->captraj : 256;
        score_gm_traj_row *nt = (score_gm_traj_row *)realloc(o->traj, nc * sizeof(*nt));
        if (!nt) return SCORE_GM_ERR_OOM;
        o->traj = nt;
        o->captraj = nc;
    }
    o->traj[o->ntraj++] = *r;
    return 0;
}

/* log_grainevo (SCORE_OnTheFlyCharacterization.cpp:478-530): time (before t
 * advances), X, the log counter, step and Sv, and the cell count of every grain,
 * deformed grains first */
static int o_log_grainevo(oracle *o, uint32_t sv) {
    const size_t ng = (size_t)o->n_defg + o->n_rxg;
    if (o->nglog == o->capglog) {
        size_t nc = o->capglog ? 2 * o->capglog : 16;
        score_gm_grainlog_row *nr = (score_gm_grainlog_row *)realloc(o->glog, nc * sizeof(*nr));
        if (!nr) return SCORE_GM_ERR_OOM;
        o->glog = nr;
        uint32_t *nc2 = (uint32_t *)realloc(o->glog_counts, nc * (ng ? ng : 1) * sizeof(uint32_t));
        if (!nc2) return SCORE_GM_ERR_OOM;
        o->glog_counts = nc2;
        o->capglog = nc;
    }
    score_gm_grainlog_row *r = &o->glog[o->nglog];
    memset(r, 0, sizeof(*r));
    r->t = o->t;
    r->X = o->X;
    r->logcnt = o->glog_cnt;
    r->step = o->step;
    r->sv = sv;
    r->ngrains = (uint32_t)ng;
    uint32_t *c = o->glog_counts + o->nglog * ng;
    memcpy(c, o->defcnt, sizeof(uint32_t) * o->n_defg);
    memcpy(c + o->n_defg, o->rxcnt, sizeof(uint32_t) * o->n_rxg);
    o->nglog++;
    o->glog_cnt++;
    return 0;
}

/* termination test (SCORE_Kernel.cpp:6619); the loop itself runs at most
 * DEFAULT_NMAX iterations (:6615) */
static int o_done(const oracle *o) {
    return (o->Xcells >= (double)o->N) || (o->X > o->p.xmax) || ((int64_t)o->step > o->p.nmax) ||
           (o->t > o->tprocessingend) || (o->t > o->p.tmax) || ((int64_t)o->step >= O_DEFAULT_NMAX);
}

/* one iteration of the main loop (SCORE_Kernel.cpp:6615-6747); returns 1 when
 * the termination test fired, 0 after a step, <0 on error */
int oracle_step(oracle *o) {
    if (o->err) return o->err;
    if (o->completed) return 1;
    if (o_done(o)) { o->completed = 1; return 1; }
    if (o->defrag_cnt < o->p.n_defrag && o->Xcells >= o->defrag_at[o->defrag_cnt]) {
        for (uint32_t q = 0; q < o->nreg; q++) {
            o_defrag_list(&o->reg[q].in);
            o_defrag_list(&o->reg[q].bd);
        }
        o->defrag_cnt++;
    }
    int rc;
    /* nuclei arriving at this step, in pool order, at the point where the
     * reference marks the injection of time-dependent nucleation: after the
     * defragmentation, before the growth step (SCORE_Kernel.cpp:6632; the sweep
     * must see them, :4068) */
    uint32_t nplaced = 0;
    if (o->step >= 1 && o->step <= o->arr_last) {
        for (uint32_t i = o->arr_off[o->step]; i < o->arr_off[o->step + 1]; i++) {
            int placed = 0;
            if ((rc = o_juvenile_nucleation(o, o->arr_ids[i], &placed))) { o->err = rc; return rc; }
            nplaced += (uint32_t)placed;
        }
    }
    /* the regions run their growth step side by side and touch only their own
     * cells, lists and halos until the barrier (SCORE_Kernel.cpp:6634-6648), so
     * taking them one after another gives the same result */
    for (uint32_t q = 0; q < o->nreg; q++)
        if ((rc = o_sweep(o, &o->reg[q].in)) || (rc = o_sweep(o, &o->reg[q].bd))) { o->err = rc; return rc; }
    o->last_infected = 0;
    for (uint32_t q = 0; q < o->nreg; q++) {
        o_region *R = &o->reg[q];
        if ((rc = o_update_fullrx(o, R, &R->in, 0))) { o->err = rc; return rc; }
        for (uint32_t h = 0; h < 27; h++) R->halo[h].nref = 0; /* ompshar_clear_halo_bookkeeping (:4667-4675) */
        if ((rc = o_update_fullrx(o, R, &R->bd, 1))) { o->err = rc; return rc; }
    }
    /* after the barrier every region takes up what its neighbours parked for it (:6654) */
    if (o->nreg > 1)
        for (uint32_t q = 0; q < o->nreg; q++)
            if ((rc = o_sync_halos(o, q))) { o->err = rc; return rc; }
    uint32_t sv = 0, nfi = 0, nfb = 0, nsi = 0, nsb = 0;
    for (uint32_t q = 0; q < o->nreg; q++) {
        const o_region *R = &o->reg[q];
        sv += R->in.sv + R->bd.sv;
        nfi += R->in.nupdated; nfb += R->bd.nupdated;
        nsi += R->in.nnever; nsb += R->bd.nnever;
    }
    score_gm_traj_row r;
    memset(&r, 0, sizeof(r));
    r.step = o->step;
    r.t = o->t;
    r.dt = o->dt;
    r.temperature = o->T;
    /* master section (SCORE_Kernel.cpp:6688-6742) */
    double mymobmax = O_DEFAULT_PMAX;
    if (o->reg_Pmax >= mymobmax) mymobmax = o->reg_Pmax;
    o->Pmax = mymobmax;
    for (uint32_t q = 0; q < o->nreg; q++)
        o->Xcells = o->Xcells + ((double)(o->reg[q].in.nupdated + o->reg[q].bd.nupdated));
    o->X = (o->Xcells / ((double)o->N));
    o->sum_sv += sv;
    o_update_rhomax(o);
    r.sv = sv;
    r.nfull_inside = nfi;
    r.nfull_border = nfb;
    r.nslots_inside = nsi;
    r.nslots_border = nsb;
    r.ninfected = o->last_infected;
    r.X = o->X;
    r.xcells = o->Xcells;
    r.rhomax = o->rhomax;
    r.pmax = o->Pmax;
    r.nplaced = nplaced;
    /* log_OUTPUT_Diehl (SCORE_OnTheFlyCharacterization.cpp:646-651): at most one
     * grain-evolution log per step, once Xcells reaches the next threshold */
    if (o->glog_cnt < o->p.n_grainlog && o->Xcells >= o->glog_at[o->glog_cnt]) {
        if ((rc = o_log_grainevo(o, sv))) { o->err = rc; return rc; }
    }
    o->t += o->dt;
    o_update_temperature(o);
    o_update_atomistic(o);
    o_update_microchemistry(o);
    o_update_mobilities(o);
    o_update_deformedsubstructure(o);
    o->dt = o_dt_minimum(o);
    o->step++;
    if ((rc = o_push_traj(o, &r))) { o->err = rc; return rc; }
    return 0;
}

int oracle_run(oracle *o, uint32_t max_steps) {
    for (uint32_t k = 0; k < max_steps; k++) {
        int rc = oracle_step(o);
        if (rc != 0) return rc;
    }
    return 0;
}

/* ---------------- accessors ---------------- */

int oracle_get_grainmap(const oracle *o, uint32_t *out, int encoding) {
    for (uint64_t i = 0; i < o->N; i++) {
        uint32_t v = o->grid[i];
        if (encoding == SCORE_GM_MAP_RAW2GID) v = (v == O_CURRENTLY_INFECTED) ? 1u : 2u + v;
        out[i] = v;
    }
    return 0;
}

int oracle_get_counts(const oracle *o, uint32_t *defg, uint32_t *rxg) {
    if (defg) memcpy(defg, o->defcnt, sizeof(uint32_t) * o->n_defg);
    if (rxg) memcpy(rxg, o->rxcnt, sizeof(uint32_t) * o->n_rxg);
    return 0;
}

size_t oracle_get_trajectory(const oracle *o, score_gm_traj_row *rows, size_t cap) {
    size_t n = o->ntraj < cap ? o->ntraj : cap;
    if (rows && n) memcpy(rows, o->traj, n * sizeof(*rows));
    return o->ntraj;
}

int oracle_get_result(const oracle *o, score_gm_result *r) {
    memset(r, 0, sizeof(*r));
    r->t = o->t; r->X = o->X; r->xcells = o->Xcells; r->step = o->step; r->completed = o->completed;
    r->sum_sv = o->sum_sv; r->nuclei_placed = o->nuclei; r->pmax = o->Pmax;
    return o->err;
}

double oracle_get_dt(const oracle *o) { return o->dt; }

size_t oracle_get_grainlog(const oracle *o, score_gm_grainlog_row *rows, uint32_t *counts, size_t cap) {
    const size_t n = o->nglog < cap ? o->nglog : cap;
    const size_t ng = (size_t)o->n_defg + o->n_rxg;
    if (rows && n) memcpy(rows, o->glog, n * sizeof(*rows));
    if (counts && n) memcpy(counts, o->glog_counts, n * ng * sizeof(uint32_t));
    return o->nglog;
}

int oracle_get_front(const oracle *o, int list, uint32_t cap, uint32_t *n, uint8_t *act, uint8_t *inf, uint32_t *cell,
                     double *frac, double *P, uint32_t *defg, uint32_t *rxg) {
    /* list 2r + SCORE_GM_LIST_* is the Inside or Border list of region r */
    if (list < 0 || (uint32_t)list >= 2 * o->nreg) return SCORE_GM_ERR_ARG;
    const o_region *R = &o->reg[list / 2];
    const o_list *L = (list & 1) == SCORE_GM_LIST_BORDER ? &R->bd : &R->in;
    *n = L->nnever;
    uint32_t m = L->nnever < cap ? L->nnever : cap;
    for (uint32_t i = 0; i < m; i++) {
        const o_cell *e = &L->c[i];
        if (act) act[i] = e->act;
        if (inf) inf[i] = e->inf;
        if (cell) cell[i] = (uint32_t)((int64_t)e->x + (int64_t)e->y * o->nx + (int64_t)e->z * o->nxy);
        if (frac) frac[i] = e->frac;
        if (P) P[i] = e->P;
        if (defg) defg[i] = e->defg;
        if (rxg) rxg[i] = e->rxg;
    }
    return 0;
}

/* ompcrit_clarify_status_infectedcells (SCORE_Io_Characterization.cpp:925-997):
 * region by region, the ACTIVE records of its Inside then its Border list, in
 * slot order; returns the count */
uint32_t oracle_get_status(const oracle *o, double thr, uint32_t cap, uint32_t *ixyz, uint32_t *gid) {
    uint32_t n = 0;
    for (uint32_t l = 0; l < 2 * o->nreg; l++) {
        const o_list *L = (l & 1) ? &o->reg[l / 2].bd : &o->reg[l / 2].in;
        for (uint32_t c = 0; c < L->nnever; c++) {
            const o_cell *e = &L->c[c];
            if (!e->act) continue;
            if (n < cap) {
                ixyz[n] = (uint32_t)((int64_t)e->x + (int64_t)e->y * o->nx + (int64_t)e->z * o->nxy);
                gid[n] = e->frac < thr ? e->defg : o->n_defg + e->rxg;
            }
            n++;
        }
    }
    return n;
}

/* ompshar_defragmentation of both lists outside the step loop (a test hook) */
void oracle_defragment(oracle *o) {
    for (uint32_t q = 0; q < o->nreg; q++) {
        o_defrag_list(&o->reg[q].in);
        o_defrag_list(&o->reg[q].bd);
    }
}

/* list capacities the reference derives (SCORE_Kernel.cpp:3589,3615,3631,3679,3682,3697)
 * for region r: inside, inside recycle cap, border, border recycle cap */
void oracle_capacities(const oracle *o, uint32_t r, uint32_t *out4) {
    const o_region *R = &o->reg[r < o->nreg ? r : 0];
    out4[0] = R->in.ntotal; out4[1] = R->in.rec_cap; out4[2] = R->bd.ntotal; out4[3] = R->bd.rec_cap;
}

/* the region grid: npy, npz, then per region its inclusive box lo x,y,z, hi x,y,z */
uint32_t oracle_regions(const oracle *o, uint32_t *npy_npz, uint32_t *boxes) {
    if (npy_npz) { npy_npz[0] = o->npy; npy_npz[1] = o->npz; }
    if (boxes)
        for (uint32_t r = 0; r < o->nreg; r++)
            for (int a = 0; a < 3; a++) { boxes[6 * r + a] = o->reg[r].lo[a]; boxes[6 * r + 3 + a] = o->reg[r].hi[a]; }
    return o->nreg;
}
