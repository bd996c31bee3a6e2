/*
 * score_setup.h -- C ABI of SCORE's input deck and per-domain setup, on the host.
 *
 * What SCORE does before solve_RXGROWTH on one solitary-unit domain, so that a
 * stock .uds input gives the growth machine (score_gm.h) the same domain:
 *
 *   score_uds_*      the UDS file format (io::readDataBlock, geTReal, geTInt,
 *                    getReal; SCORE_Io.cpp:265-434): named blocks of key/value
 *                    pairs (read as double / long) and of typed columns ('f'
 *                    columns are read as float, as SCORE does);
 *   score_mt_*       randomClass's Mersenne Twister (SCORE_Random.cpp:49-76,
 *                    :400-458): the 1998 seeding by 69069 multiplication, 500000
 *                    warm-up draws, values in [0, 1 - DBL_EPSILON];
 *   score_setup_ensemble_*  ensembleHdl::init_parameter_* (SCORE_Kernel.cpp:262-940):
 *                    the settings of the growth path and the pools of deformed
 *                    and recrystallized grains, orientations merged within 0.5
 *                    degree (ens_check_disjunctness_core, SCORE_OrientationHdl.cpp:80-155);
 *   score_setup_domain_*    one caHdl: its random streams (init_cahdlprng,
 *                    SCORE_Kernel.cpp:5729-5745), the cuboid deformed structure
 *                    (determine_polycrystalline_geometry, pick_deformedgrains,
 *                    ompshar_voxelize_defms; SCORE_StructureSynthesis.cpp:355-529),
 *                    site-saturated CSR nucleation with Shoemake orientations
 *                    (solve_nucmodeling_csr_enforced, SCORE_NucleationModels.cpp:207-282)
 *                    and the grain-id replacement (solve_REPLACE_CA_STRUCTURE,
 *                    SCORE_Kernel.cpp:6353-6445), as a score_gm_domain and
 *                    score_gm_params.
 *
 * Supported: DefStructureSynthesis 1 (cuboid grains), CSRNucleation 1
 * (csr_enforced), no grain-boundary or clustered nucleation, site saturation,
 * the three recovery models, Zener drag constant or linear in time, no EBSD input; other settings return
 * SCORE_SETUP_ERR_UNSUPPORTED.  All functions return 0 or a negative
 * SCORE_SETUP_ERR_* code; score_setup_last_error() describes the last failure
 * of the calling thread.  Objects are immutable once created and may be shared
 * between threads.
 */
#ifndef SCORE_SETUP_H
#define SCORE_SETUP_H

#include <stdint.h>

#include "score_gm.h"

#ifdef __cplusplus
extern "C" {
#endif

#define SCORE_SETUP_ABI_VERSION 3

#define SCORE_SETUP_OK 0
#define SCORE_SETUP_ERR_ARG -1          /* invalid argument */
#define SCORE_SETUP_ERR_FILE -2         /* the file cannot be read */
#define SCORE_SETUP_ERR_FORMAT -3       /* a block, key or value is missing or unreadable (SCORE's exitus) */
#define SCORE_SETUP_ERR_RANGE -4        /* a value SCORE rejects (its status = false) */
#define SCORE_SETUP_ERR_UNSUPPORTED -5  /* a setting this setup does not implement */
#define SCORE_SETUP_ERR_OOM -6

/* DEFAULT_SEED (SCORE_Random.h:97): the ensemble seed unless one is passed
 * explicitly (ensembleHdl::init_ensprng(int), SCORE_Kernel.cpp:208-222) */
#define SCORE_SETUP_DEFAULT_SEED (-46356)

int score_setup_abi_version(void);
const char *score_setup_last_error(void);

/* ---- randomClass's Mersenne Twister ---------------------------------------- */

typedef struct score_mt {
    uint32_t mt[624];
    uint32_t mti;
    uint32_t reserved;
} score_mt;

/* randomClass::init(seed) for the Mersenne Twister: seedMT = (uint32_t)seed,
 * sgenrand(), warmupMT(500000) (SCORE_Random.cpp:49-76, SCORE_Random.h:126-139) */
int score_mt_init(score_mt *g, int64_t seed);
/* randomClass::MersenneTwister() (SCORE_Random.cpp:409-458) */
double score_mt_next(score_mt *g);
/* mathMethods::rnd_quat_shoemake (SCORE_Math.cpp:660-686) drawing from g */
void score_mt_shoemake(score_mt *g, double q[4]);

/* ---- UDS files ----------------------------------------------------------------- */

typedef struct score_uds score_uds;

/* reads the whole file; blocks are looked up by name later (the first block of a name wins) */
int score_uds_open(const char *path, score_uds **out);
void score_uds_close(score_uds *u);
/* geTReal (SCORE_Io.cpp:369-411): the value of `key` in key/value block `block` as a double */
int score_uds_real(const score_uds *u, const char *block, const char *key, double *out);
/* geTInt (SCORE_Io.cpp:413-425) */
int score_uds_int(const score_uds *u, const char *block, const char *key, int64_t *out);
/* rows and columns of a block */
int score_uds_shape(const score_uds *u, const char *block, uint32_t *rows, uint32_t *cols);
/* getReal (SCORE_Io.cpp:428-432): row (0-based) and column (0-based) of a block
 * whose column is typed 'f', read as float like SCORE, widened to double */
int score_uds_table_real(const score_uds *u, const char *block, uint32_t row, uint32_t col, double *out);

/* ---- ensemble ------------------------------------------------------------------ */

typedef struct score_setup_ensemble score_setup_ensemble;

typedef struct score_setup_ensemble_info {
    uint32_t n_domains;           /* CAEnsembleSize */
    uint32_t nx, ny, nz;          /* 3DCAEdgeLengthInCells RD, TD, ND */
    double cellsize;              /* m */
    double defg_mean[3];          /* MeanDefGrainSizein RD, TD, ND, m */
    uint32_t n_nuclei;            /* NucleationDensityLocalCSR: nuclei per domain */
    uint32_t n_world_ori;         /* orientations after merging */
    uint32_t n_world_defg;        /* DeformedGrainsPool rows */
    uint32_t n_world_rxg;         /* RXGrainsPool rows */
    uint32_t n_ideal;             /* IdealComponents rows (texture classes 1..n_ideal; 0 is RANDOM) */
} score_setup_ensemble_info;

/* all settings of the growth path from a UDS file (ensembleHdl::init_parameter_*) */
int score_setup_ensemble_create(const score_uds *u, score_setup_ensemble **out);
int score_setup_ensemble_get_info(const score_setup_ensemble *e, score_setup_ensemble_info *out);
/* quaternion of world orientation i (after merging) and the orientation of world deformed grain g */
int score_setup_world_quat(const score_setup_ensemble *e, uint32_t i, double q[4]);
int score_setup_world_defg(const score_setup_ensemble *e, uint32_t g, uint32_t *ori, double *rho);
void score_setup_ensemble_free(score_setup_ensemble *e);
/* ideal texture component i (init_parameter_idealtexture, SCORE_Kernel.cpp:668-712):
 * Bunge angles and scatter in radians, its quaternion */
int score_setup_ideal(const score_setup_ensemble *e, uint32_t i, double bunge[3], double q[4], double *scatter);
/* the texture class of an orientation (ens_get_closest_standardlage,
 * SCORE_OrientationHdl.cpp:51-77): 1 + the closest ideal component within its
 * scatter (the last of equally close ones), or 0 (RANDOM) */
int score_setup_closest_ideal(const score_setup_ensemble *e, const double q[4], uint32_t *cls);

/* ---- one domain ----------------------------------------------------------------- */

typedef struct score_setup_domain score_setup_domain;

typedef struct score_setup_domain_info {
    int64_t local_seed;           /* the seed of localprng and localmath.r */
    double offset[3];             /* u_xrd, v_ytd, w_znd: origin of the cuboid grid, m (<= 0) */
    uint32_t ngr[3];              /* cuboid grains along RD, TD, ND */
    uint32_t n_defg;              /* distinct deformed grains of the domain (mydefgpool) */
    uint32_t nuclei_desired;
    uint32_t nuclei_placed;
    uint32_t collisions;          /* site draws that hit an occupied site (maxattempts) */
    uint32_t reserved;
} score_setup_domain_info;

/* Domain `jobid` of the ensemble.  ensemble_seed is SCORE_SETUP_DEFAULT_SEED for
 * SCORE's default, where domain j draws from seed -(j+1)-10, or an explicit
 * seed that every domain then shares (init_cahdlprng, SCORE_Kernel.cpp:5729-5745). */
int score_setup_domain_create(const score_setup_ensemble *e, uint32_t jobid, int64_t ensemble_seed,
                              score_setup_domain **out);
/* views of the domain's arrays and settings, valid while both objects live */
int score_setup_domain_view(const score_setup_domain *s, score_gm_domain *dom, score_gm_params *prm);
int score_setup_domain_get_info(const score_setup_domain *s, score_setup_domain_info *out);
void score_setup_domain_free(score_setup_domain *s);

#ifdef __cplusplus
}
#endif

#endif
