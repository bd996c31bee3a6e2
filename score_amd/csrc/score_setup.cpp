/*
 * score_setup.cpp -- SCORE's input deck and per-domain setup on the host
 * (include/score_setup.h).  Host C++ only; compiled with -ffp-contract=off and
 * without -march so that every double is rounded as in the reference's GCC build.
 */
#include "../../include/score_setup.h"

#include <float.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <string>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "score_orient.h"

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &what) {
    g_err = what;
    return code;
}

/* ---- constants of the setup path (SCORE_Defs_Physics.h, SCORE_Defs_Functionality.h) ---- */
constexpr double kPiDefs = 3.1415926535897932384626433832795;             /* pi, SCORE_Defs_Physics.h:55 */
constexpr double kMicron2Meter = (1.0e-6);                                   /* :59 */
constexpr double kToffset = (273.15);                                        /* :52 */
constexpr double kEcharge = (1.602176565e-19);                               /* :53 */
constexpr double kRhoRecrystallized = (1.0e11);                              /* :66 */
constexpr double kResolutionSO3Grid = ((0.5) / (180.0) * (kPiDefs));        /* :102 */
constexpr double kBungeMin = ((0.0) / (180.0) * (kPiDefs));                 /* :106 */
constexpr double kBungeMax = ((360.0) / (180.0) * (kPiDefs));               /* :107 */
constexpr double kMaxFcc = (1.099);                                          /* MAX_FCC, :101 */
constexpr double kDefaultMaxFill = (0.1), kMinFillIn = (0.01), kMaxFillIn = (0.2);      /* :123-125 */
constexpr double kDefaultRelCell = (0.10), kDefaultTransRelCell = (1.0);                /* :127-128 */
constexpr double kDefaultTmax = (360000.0);                                  /* :132 */
constexpr long kDefaultNmax = 1000000;                                       /* :133 */
constexpr long kDefaultNucDensityCsr = 1;                                    /* :134 */
constexpr double kCellsizeMin = (1e-7), kCellsizeMax = (1e-5);              /* :135-136 */
constexpr double kMinRelCell = (0.05), kMaxRelCell = (0.25);                 /* :137-138 */
constexpr double kMinTransRelCell = (0.5), kMaxTransRelCell = (2.0);         /* :139-140 */
constexpr double kMinGrainDiscr = 10;                                        /* :141 */
constexpr uint32_t kMaxAttemptsNucSite = 100000;                             /* :143 */
constexpr uint32_t kCaDimensionsMinimum = 50;                                /* SCORE_Defs_Functionality.h:84 */
constexpr uint64_t kCaCellsMaximum = 1600ull * 1600ull * 1600ull;            /* CUBE(CA_DIMENSIONS_MAXIMUM), :83 */
constexpr long kIntegerRangeMax = 2147483647;                                /* :135 */
constexpr long kMaxNcas = 1048576;                                           /* :161 */
constexpr long kMaxDeformedGrains = 65536;                                   /* :162 */
constexpr long kMaxRxGrains = 1048576;                                       /* :164 */
constexpr size_t kMaxInputDataLen = 4096;                                    /* DEFAULT_MAX_INPUTDATALEN, :165 */
constexpr int kBufsize = 1024;                                               /* BUFSIZE, SCORE_Io.h:62 */
constexpr int64_t kDefaultSeed = SCORE_SETUP_DEFAULT_SEED;

inline double deg2rad(double deg) { return ((deg) / (180.0) * (score_orient::kPi)); } /* DEG2RAD, SCORE_Math.h:57 */

/* ---- Mersenne Twister (SCORE_Random.cpp:400-458, SCORE_Random.h:82-92) ---- */
constexpr uint32_t kNMT = 624, kMMT = 397;
constexpr uint32_t kMatrixA = 0x9908b0dfu, kUpperMask = 0x80000000u, kLowerMask = 0x7fffffffu;
constexpr uint32_t kTemperB = 0x9d2c5680u, kTemperC = 0xefc60000u;
constexpr double kRandMax = (1.0 - DBL_EPSILON);

void mt_sgenrand(score_mt *g, uint32_t seed) {
    g->mt[0] = seed & 0xffffffffu;
    for (g->mti = 1; g->mti < kNMT; g->mti++) g->mt[g->mti] = (69069u * g->mt[g->mti - 1]) & 0xffffffffu;
}

double mt_next(score_mt *g) {
    static const uint32_t mag01[2] = {0x0u, kMatrixA};
    uint32_t y;
    if (g->mti >= kNMT) {
        /* init() always seeds, so the reseed-from-time branch of the reference never runs */
        uint32_t kk;
        for (kk = 0; kk < kNMT - kMMT; kk++) {
            y = (g->mt[kk] & kUpperMask) | (g->mt[kk + 1] & kLowerMask);
            g->mt[kk] = g->mt[kk + kMMT] ^ (y >> 1) ^ mag01[y & 0x1u];
        }
        for (; kk < kNMT - 1; kk++) {
            y = (g->mt[kk] & kUpperMask) | (g->mt[kk + 1] & kLowerMask);
            g->mt[kk] = g->mt[kk + kMMT - kNMT] ^ (y >> 1) ^ mag01[y & 0x1u];
        }
        y = (g->mt[kNMT - 1] & kUpperMask) | (g->mt[0] & kLowerMask);
        g->mt[kNMT - 1] = g->mt[kMMT - 1] ^ (y >> 1) ^ mag01[y & 0x1u];
        g->mti = 0;
    }
    y = g->mt[g->mti++];
    y ^= (y >> 11);
    y ^= (y << 7) & kTemperB;
    y ^= (y << 15) & kTemperC;
    y ^= (y >> 18);
    const double randr = ((double)y / (uint32_t)0xffffffffu);
    if (randr > kRandMax) return kRandMax;
    return randr;
}

void mt_init(score_mt *g, int64_t seed) {
    /* randomClass::init(long sd): seedMT = (uint32_t) sd; mti = NMT+1; sgenrand(); warmupMT(500000) */
    mt_sgenrand(g, (uint32_t)seed);
    g->reserved = 0;
    for (unsigned r = 0; r < 500000u; r++) (void)mt_next(g);
}

void mt_shoemake(score_mt *g, double *q) {
    const double X0 = mt_next(g);
    const double X1 = mt_next(g);
    const double X2 = mt_next(g);
    score_orient::shoemake(X0, X1, X2, q);
}

}  // namespace

/* ---------------------------------------------------------------- UDS files */

struct UdsCell {
    char type = 's';
    float f = 0.0f;
    long i = 0;
    std::string s;
};

struct UdsBlock {
    std::string name;
    std::string types;       /* column types, e.g. "ss" or "ff" */
    bool defective = false;  /* header without "*(...)" */
    std::string error;       /* a data line SCORE rejects (exitus), reported when the block is read */
    std::vector<std::vector<UdsCell>> lines;
};

struct score_uds {
    std::vector<UdsBlock> blocks;
};

namespace {

const UdsBlock *find_block(const score_uds *u, const char *name) {
    for (const UdsBlock &b : u->blocks)
        if (b.name == name) return &b;
    return nullptr;
}

/* the data lines of a block as readDataBlock takes them (SCORE_Io.cpp:299-356):
 * per column skip blanks and tabs, take one whitespace-delimited token; a line
 * whose first token is missing is skipped, one whose first token starts with
 * "||" ends the block, a line with fewer tokens than columns is an error */
void parse_block_lines(UdsBlock &b, const std::vector<std::string> &text, size_t first) {
    const size_t ncol = b.types.size();
    for (size_t ln = first; ln < text.size(); ln++) {
        const std::string &buf1 = text[ln];
        std::vector<UdsCell> line(ncol);
        size_t weiter = 0;
        bool have = true;
        for (size_t i = 0; i < ncol; i++) {
            while (weiter < buf1.size() && (buf1[weiter] == ' ' || buf1[weiter] == '\t')) weiter++;
            char buf2[kBufsize];
            buf2[0] = 0;
            if (weiter < buf1.size()) sscanf(buf1.c_str() + weiter, "%1023s", buf2);
            if ((buf2[0] == '|' && buf2[1] == '|') || strlen(buf2) == 0) {
                if (i > 0) {
                    if (b.error.empty()) b.error = "incomplete line in Block -->" + b.name + "<--";
                } else if (buf2[0] == '|' && buf2[1] == '|') {
                    return;
                }
                have = false;
                break;
            }
            weiter += strlen(buf2) + 1;
            UdsCell &c = line[i];
            c.type = b.types[i];
            int cnt = 1;
            if (c.type == 's') c.s = buf2;
            else if (c.type == 'f') cnt = sscanf(buf2, "%f", &c.f);
            else if (c.type == 'i') {
                int v = 0;
                cnt = sscanf(buf2, "%i", &v);
                c.i = v;
            } else {
                cnt = 0;
            }
            if (cnt != 1 && b.error.empty()) b.error = "file error 6 in Block -->" + b.name + "<--";
        }
        if (have) b.lines.push_back(std::move(line));
    }
}

/* geTString (SCORE_Io.cpp:369-395): the value of key `s`, one leading and one
 * trailing double quote removed from name and value */
int get_string(const UdsBlock &b, const char *s, std::string *out) {
    if (b.types.size() < 2 || b.types[0] != 's' || b.types[1] != 's')
        return fail(SCORE_SETUP_ERR_FORMAT, "block -->" + b.name + "<-- is not a (ss) parameter block");
    for (const auto &line : b.lines) {
        std::string name = line[0].s;
        if (!name.empty() && name[0] == '"') name.erase(0, 1);
        if (!name.empty() && name.back() == '"') name.pop_back();
        if (name != s) continue;
        std::string v = line[1].s;
        if (!v.empty() && v[0] == '"') v.erase(0, 1);
        if (!v.empty() && v.back() == '"') v.pop_back();
        *out = v;
        return 0;
    }
    return fail(SCORE_SETUP_ERR_FORMAT, std::string("-->") + s + "<-- gibts nicht im Block -->" + b.name + "<--");
}

int block_or_fail(const score_uds *u, const char *block, const UdsBlock **out) {
    const UdsBlock *b = find_block(u, block);
    if (!b) return fail(SCORE_SETUP_ERR_FORMAT, std::string("Block -->") + block + "<-- does not exist in file");
    if (b->defective) return fail(SCORE_SETUP_ERR_FORMAT, std::string("block header defective: ") + block);
    if (!b->error.empty()) return fail(SCORE_SETUP_ERR_FORMAT, b->error);
    *out = b;
    return 0;
}

int uds_real(const score_uds *u, const char *block, const char *key, double *out) {
    const UdsBlock *b;
    int rc = block_or_fail(u, block, &b);
    if (rc) return rc;
    std::string v;
    if ((rc = get_string(*b, key, &v))) return rc;
    double x;
    char dummy;
    const int cnt = sscanf(v.c_str(), "%lf%c", &x, &dummy);
    if (cnt < 1) return fail(SCORE_SETUP_ERR_FORMAT, std::string("can't read a double: ") + key);
    if (cnt > 1) return fail(SCORE_SETUP_ERR_FORMAT, std::string("something's behind the double: ") + key);
    *out = x;
    return 0;
}

int uds_int(const score_uds *u, const char *block, const char *key, long *out) {
    const UdsBlock *b;
    int rc = block_or_fail(u, block, &b);
    if (rc) return rc;
    std::string v;
    if ((rc = get_string(*b, key, &v))) return rc;
    long x;
    char dummy;
    const int cnt = sscanf(v.c_str(), "%ld%c", &x, &dummy);
    if (cnt < 1) return fail(SCORE_SETUP_ERR_FORMAT, std::string("can't read a long: ") + key);
    if (cnt > 1) return fail(SCORE_SETUP_ERR_FORMAT, std::string("something's behind the long: ") + key);
    *out = x;
    return 0;
}

/* getReal (SCORE_Io.cpp:428-432): a float column, widened */
int table_real(const UdsBlock &b, size_t row, size_t col, double *out) {
    if (row >= b.lines.size() || col >= b.types.size())
        return fail(SCORE_SETUP_ERR_ARG, "row or column outside block -->" + b.name + "<--");
    if (b.types[col] != 'f') return fail(SCORE_SETUP_ERR_FORMAT, "column of block -->" + b.name + "<-- is not of type f");
    *out = b.lines[row][col].f;
    return 0;
}

}  // namespace

/* ---------------------------------------------------------------- ensemble */

struct WorldOri {
    double q[4];
};
struct WorldDefg {
    uint32_t ori;
    double rho0, rho;
};
struct Ideal {
    double bunge[3];  /* rad */
    double q[4];
    double scatter;   /* rad */
};

struct score_setup_ensemble {
    /* RuntimeControl */
    double xmax = 0, tmax = 0, maxfill = 0, relcell = 0, recache = 0;
    long nmax = 0;
    bool defrag_on = false;
    /* EnsembleDefinition */
    uint32_t n_domains = 0, nx = 0, ny = 0, nz = 0;
    uint64_t nvol = 0;
    double cellsize = 0, boxedge[3] = {0, 0, 0}, defgmean[3] = {0, 0, 0};
    /* NucleationModeling */
    uint32_t nuc_density = 0;
    /* MaterialProperties, BoundaryMobilityModel */
    double G0 = 0, b0 = 0;
    int32_t model = 0;
    double lagb_m0 = 0, lagb_h = 0, hagb_m0 = 0, hagb_h = 0, gs_m0 = 0, gs_h = 0;
    double rh_m0 = 0, rh_h = 0, rh_cut = 0, rh_trans = 0, rh_exp = 0;
    /* RecoveryParameter (enthalpies in J) */
    int32_t recovery_mode = 0;
    double rec_vd0 = 0, rec_vh = 0, rec_sd0 = 0, rec_sh = 0, rec_ls = 0, rec_conc = 0;
    double rec_alpha3 = 0, rec_kappa2 = 0, rec_c3 = 0, rec_c5 = 0;
    /* DragParameter, EvoMicrochemistry */
    int32_t zener_mode = 0;
    double zener_alpha = 0, zener_gamma = 0;
    std::vector<double> disp_t, disp_fr;
    /* IdealComponents */
    std::vector<Ideal> ideals;
    /* AnnealingSchedule */
    std::vector<double> ttime, ttemp;
    /* HeuristicRXFrontListDefragmentation: X, ascending */
    std::vector<double> defrag_x;
    std::vector<double> log_x;   /* SingleGrainVolOverTime (UserDefLogPoint_X_Output) */
    /* pools */
    std::vector<WorldOri> oris;
    std::vector<WorldDefg> defg;
    uint32_t n_rxg = 0;
};

namespace {

/* ens_check_disjunctness_core (SCORE_OrientationHdl.cpp:80-155): an orientation
 * within RESOLUTION_SO3GRID of one already known (the last of the closest) is
 * that one; else it is added */
uint32_t ens_check_disjunctness(score_setup_ensemble *e, const double *q) {
    double closest = kResolutionSO3Grid;
    uint32_t closestid = UINT32_MAX;
    for (uint32_t cand = 0; cand < e->oris.size(); cand++) {
        const double disori = score_orient::disori_angle(q, e->oris[cand].q);
        if (disori <= closest) {
            closest = disori;
            closestid = cand;
        }
    }
    if (closestid == UINT32_MAX) {
        WorldOri o;
        memcpy(o.q, q, sizeof(o.q));
        e->oris.push_back(o);
        return (uint32_t)(e->oris.size() - 1);
    }
    return closestid;
}

int read_runtime(const score_uds *u, score_setup_ensemble *e) {
    /* init_parameter_runtime (SCORE_Kernel.cpp:262-317); values out of the
     * recommended ranges are reset as on SCORE's master rank */
    const char *B = "RuntimeControl";
    int rc;
    long v;
    if ((rc = uds_real(u, B, "XMAX", &e->xmax))) return rc;
    if (e->xmax < 0.0 || e->xmax > 1.0) return fail(SCORE_SETUP_ERR_RANGE, "Invalid argument in XMAX");
    if ((rc = uds_real(u, B, "TIMEMAX", &e->tmax))) return rc;
    if (e->tmax < 0.0 || e->tmax > kDefaultTmax) return fail(SCORE_SETUP_ERR_RANGE, "Invalid argument in TMAX");
    if ((rc = uds_int(u, B, "NMAX", &e->nmax))) return rc;
    if (e->nmax < 1 || e->nmax > kDefaultNmax || e->nmax >= kIntegerRangeMax)
        return fail(SCORE_SETUP_ERR_RANGE, "Invalid argument for NMAX");
    if ((rc = uds_real(u, B, "MAXFILLPERSTEP", &e->maxfill))) return rc;
    if (e->maxfill < kMinFillIn || e->maxfill > kMaxFillIn) e->maxfill = kDefaultMaxFill;
    if ((rc = uds_real(u, B, "InitialCacheSizeRXCells", &e->relcell))) return rc;
    if (e->relcell < kMinRelCell || e->relcell > kMaxRelCell) e->relcell = kDefaultRelCell;
    if ((rc = uds_real(u, B, "ReCacheSizeRXCells", &e->recache))) return rc;
    if (e->recache < kMinTransRelCell || e->recache > kMaxTransRelCell) e->recache = kDefaultTransRelCell;
    if ((rc = uds_int(u, B, "OnTheFlyDefragmentation", &v))) return rc;
    e->defrag_on = v == 1;
    /* read but not on the growth path */
    if ((rc = uds_int(u, B, "PERCOLATIONANALYSIS", &v)) || (rc = uds_int(u, B, "RediscretizationSteps", &v)) ||
        (rc = uds_int(u, B, "SolitaryUnitPostProcessing", &v)))
        return rc;
    if ((rc = uds_int(u, B, "ReadEBSDMapping", &v))) return rc;
    if (v == 1) return fail(SCORE_SETUP_ERR_UNSUPPORTED, "ReadEBSDMapping 1 (EBSD input) is not supported");
    double step;
    if ((rc = uds_real(u, B, "ReadEBSDStepsize", &step))) return rc;
    if (step <= 0.05e-6) return fail(SCORE_SETUP_ERR_RANGE, "Invalid argument in ReadEBSDStepsize"); /* DEFAULT_EBSDSTEPSIZE_MINIMUM */
    return 0;
}

int read_ensemble(const score_uds *u, score_setup_ensemble *e) {
    /* init_parameter_ensemble (SCORE_Kernel.cpp:393-466) */
    const char *B = "EnsembleDefinition";
    int rc;
    long v;
    double x;
    if ((rc = uds_int(u, B, "CAEnsembleSize", &v))) return rc;
    if (v < 1 || v >= kIntegerRangeMax || v > kMaxNcas) return fail(SCORE_SETUP_ERR_RANGE, "Invalid argument for the number of domains");
    e->n_domains = (uint32_t)v;
    if ((rc = uds_real(u, B, "CellSize", &x))) return rc;
    e->cellsize = x * kMicron2Meter;
    if (e->cellsize < kCellsizeMin || e->cellsize > kCellsizeMax) return fail(SCORE_SETUP_ERR_RANGE, "Invalid argument for the cellsize");
    const char *edges[3] = {"3DCAEdgeLengthInCellsRD", "3DCAEdgeLengthInCellsTD", "3DCAEdgeLengthInCellsND"};
    uint32_t n[3];
    for (int a = 0; a < 3; a++) {
        if ((rc = uds_int(u, B, edges[a], &v))) return rc;
        if (v < (long)kCaDimensionsMinimum || v > 0xFFFFFFFFl)
            return fail(SCORE_SETUP_ERR_RANGE, std::string("Invalid argument for automaton size ") + edges[a]);
        n[a] = (uint32_t)v;
    }
    e->nx = n[0];
    e->ny = n[1];
    e->nz = n[2];
    e->nvol = (uint64_t)e->nx * e->ny * e->nz;
    if (e->nvol > kCaCellsMaximum || e->nvol > 0xFFFFFFFFull)
        return fail(SCORE_SETUP_ERR_RANGE, "the automaton holds more cells than SCORE handles");
    e->boxedge[0] = e->cellsize * e->nx;
    e->boxedge[1] = e->cellsize * e->ny;
    e->boxedge[2] = e->cellsize * e->nz;
    const char *means[3] = {"MeanDefGrainSizeinRD", "MeanDefGrainSizeinTD", "MeanDefGrainSizeinND"};
    for (int a = 0; a < 3; a++) {
        if ((rc = uds_real(u, B, means[a], &x))) return rc;
        e->defgmean[a] = x * kMicron2Meter;
    }
    const double mindiscretization = kMinGrainDiscr * e->cellsize;
    if (e->defgmean[0] < mindiscretization || e->defgmean[1] < mindiscretization || e->defgmean[2] < mindiscretization)
        return fail(SCORE_SETUP_ERR_RANGE, "Invalid argument for the average size of the cuboid deformed grain");
    if ((rc = uds_int(u, B, "DefStructureSynthesis", &v))) return rc;
    if (v != 1 && v >= 1 && v <= 6)
        return fail(SCORE_SETUP_ERR_UNSUPPORTED, "DefStructureSynthesis other than 1 (cuboid grains) is not supported");
    /* any other value leaves SCORE's default, cuboid grains */
    if ((rc = uds_real(u, B, "MeanDefSizePoissonDiameter", &x))) return rc;
    if (x * kMicron2Meter < mindiscretization)
        return fail(SCORE_SETUP_ERR_RANGE, "Invalid argument for the average size of deformed grain for method poisson");
    return 0;
}

int read_nucleation(const score_uds *u, score_setup_ensemble *e) {
    /* init_parameter_nucleation (SCORE_Kernel.cpp:467-535) */
    const char *B = "NucleationModeling";
    int rc;
    long v;
    if ((rc = uds_int(u, B, "NucleationDensityLocalCSR", &v))) return rc;
    if (v >= kIntegerRangeMax) return fail(SCORE_SETUP_ERR_RANGE, "Too high NucleationDensityLocalCSR");
    if (v < kDefaultNucDensityCsr) return fail(SCORE_SETUP_ERR_RANGE, "Too low NucleationDensityLocalCSR");
    e->nuc_density = (uint32_t)v;
    long csr, clust, gb, tinc;
    if ((rc = uds_int(u, B, "CSRNucleation", &csr)) || (rc = uds_int(u, B, "ClusteredNucleation", &clust)) ||
        (rc = uds_int(u, B, "GBNucleation", &gb)) || (rc = uds_int(u, B, "IncubationTimeModel", &tinc)))
        return rc;
    const bool csr_any = csr >= 1 && csr <= 6, gb_any = gb >= 1 && gb <= 3, clust_any = clust == 1;
    if (!csr_any && !gb_any && !clust_any) return fail(SCORE_SETUP_ERR_RANGE, "Choose a valid nucleation model");
    /* the remaining keys of the block; the grain-boundary ones must not be negative */
    const char *ikeys[] = {"ClusteredNucleationNCluster", "ClusteredNucleationExtendA", "ClusteredNucleationExtendB",
                           "ClusteredNucleationExtendC"};
    for (const char *k : ikeys)
        if ((rc = uds_int(u, B, k, &v))) return rc;
    double lam, scal, d2n, r2d, scat, tis;
    if ((rc = uds_real(u, B, "ClusteredNucleationLCluster", &lam)) || (rc = uds_real(u, B, "ClusteredNucleationScaling", &scal)) ||
        (rc = uds_real(u, B, "GBNucDensity2Number", &d2n)) || (rc = uds_real(u, B, "GBNucRhoDiff2Density", &r2d)) ||
        (rc = uds_real(u, B, "GBNucMaximumScatter", &scat)) || (rc = uds_real(u, B, "IncubationTimeScatter", &tis)))
        return rc;
    if (d2n < 0.0 || r2d < 0.0 || deg2rad(scat) < 0.0) return fail(SCORE_SETUP_ERR_RANGE, "Invalid grain boundary nucleation properties");
    if (tinc != 1) return fail(SCORE_SETUP_ERR_RANGE, "Only site-saturated nucleation implemented in the MPI/OpenMP model");
    if (gb_any) return fail(SCORE_SETUP_ERR_UNSUPPORTED, "GBNucleation is not supported");
    if (clust_any) return fail(SCORE_SETUP_ERR_UNSUPPORTED, "ClusteredNucleation is not supported");
    if (csr != 1) return fail(SCORE_SETUP_ERR_UNSUPPORTED, "CSRNucleation other than 1 (csr_enforced) is not supported");
    return 0;
}

int read_physics(const score_uds *u, score_setup_ensemble *e) {
    int rc;
    long v;
    double x;
    /* init_parameter_material (SCORE_Kernel.cpp:536-553) */
    const char *M = "MaterialProperties";
    double tmelt, dgdt, c0, c1, c2;
    if ((rc = uds_real(u, M, "ZeroKelvinShearModulusG0", &e->G0)) || (rc = uds_real(u, M, "FirstOrderdG0dT", &dgdt)) ||
        (rc = uds_real(u, M, "ZeroCelsiusBurgersVector", &e->b0)) ||
        (rc = uds_real(u, M, "AlloyConstantThermalExpCoeff", &c0)) || (rc = uds_real(u, M, "FirstOrderThermalExpCoeff", &c1)) ||
        (rc = uds_real(u, M, "SecondOrderThermalExpCoeff", &c2)) || (rc = uds_real(u, M, "MeltingTemperature", &tmelt)))
        return rc;
    tmelt += kToffset;
    if (e->G0 <= 0.0 || e->b0 <= 0.0 || tmelt <= 0.0) return fail(SCORE_SETUP_ERR_RANGE, "Inrealistic material parameter");
    /* init_parameter_mobilities (SCORE_Kernel.cpp:555-594): activation enthalpies in eV times echarge */
    const char *B = "BoundaryMobilityModel";
    if ((rc = uds_int(u, B, "MobilityModel", &v))) return rc;
    e->model = v == 2 ? SCORE_GM_MOB_ROLLETTHOLM : SCORE_GM_MOB_SEBALDGOTTSTEIN;
    if ((rc = uds_real(u, B, "LAGBm0", &e->lagb_m0)) || (rc = uds_real(u, B, "LAGBHact", &x))) return rc;
    e->lagb_h = x * kEcharge;
    if ((rc = uds_real(u, B, "HAGBm0", &e->hagb_m0)) || (rc = uds_real(u, B, "HAGBHact", &x))) return rc;
    e->hagb_h = x * kEcharge;
    if ((rc = uds_real(u, B, "GSm0", &e->gs_m0)) || (rc = uds_real(u, B, "GSHact", &x))) return rc;
    e->gs_h = x * kEcharge;
    if ((rc = uds_real(u, B, "RHModelHAGBm0", &e->rh_m0)) || (rc = uds_real(u, B, "RHModelHAGBHact", &x))) return rc;
    e->rh_h = x * kEcharge;
    if ((rc = uds_real(u, B, "RHModelLAGBHAGBCut", &e->rh_cut)) || (rc = uds_real(u, B, "RHModelLAGBHAGBTrans", &e->rh_trans)) ||
        (rc = uds_real(u, B, "RHModelLAGBHAGBExponent", &e->rh_exp)))
        return rc;
    if (e->model == SCORE_GM_MOB_SEBALDGOTTSTEIN &&
        (e->lagb_m0 <= 0.0 || e->lagb_h <= 0.0 || e->hagb_m0 <= 0.0 || e->hagb_h <= 0.0 || e->gs_m0 <= 0.0 || e->gs_h <= 0))
        return fail(SCORE_SETUP_ERR_RANGE, "Unrealistic mobility model parameterization");
    if (e->model == SCORE_GM_MOB_ROLLETTHOLM &&
        (e->rh_m0 <= 0.0 || e->rh_h <= 0.0 || e->rh_cut <= 0.0 || e->rh_trans <= 0.0 || e->rh_exp <= 0.0))
        return fail(SCORE_SETUP_ERR_RANGE, "Unrealistic mobility model parameterization");
    /* recovery and Zener drag (init_parameter_recovery, _drag, SCORE_Kernel.cpp:595-665):
     * RecoveryConsider 1-3 selects a recovery model, any other value none; drag is
     * constant (1) or linear in time (2) */
    const char *R = "RecoveryParameter";
    if ((rc = uds_int(u, R, "RecoveryConsider", &v))) return rc;
    e->recovery_mode = (v >= 1 && v <= 3) ? (int32_t)v : 0;
    double vgeo, c4;
    if ((rc = uds_real(u, R, "RecoveryVacancyDiffGeometry", &vgeo)) ||
        (rc = uds_real(u, R, "RecoveryVacancyDiffPreexp", &e->rec_vd0)) ||
        (rc = uds_real(u, R, "RecoveryVacancyDiffHact", &x)))
        return rc;
    e->rec_vh = x * kEcharge;
    if ((rc = uds_real(u, R, "RecoverySoluteDiffPreexp", &e->rec_sd0)) ||
        (rc = uds_real(u, R, "RecoverySoluteDiffHact", &x)))
        return rc;
    e->rec_sh = x * kEcharge;
    if ((rc = uds_real(u, R, "RecoverySoluteLsPropFactor", &e->rec_ls)) ||
        (rc = uds_real(u, R, "RecoverySoluteConcentration", &e->rec_conc)) ||
        (rc = uds_real(u, R, "RecoveryParameterAlpha3", &e->rec_alpha3)) ||
        (rc = uds_real(u, R, "RecoveryParameterKappa2", &e->rec_kappa2)) ||
        (rc = uds_real(u, R, "RecoveryParameterC3", &e->rec_c3)) || (rc = uds_real(u, R, "RecoveryParameterC4", &c4)) ||
        (rc = uds_real(u, R, "RecoveryParameterC5", &e->rec_c5)))
        return rc;
    const char *D = "DragParameter";
    if ((rc = uds_int(u, D, "ZenerConsider", &v)) || (rc = uds_real(u, D, "ZenerAlpha", &e->zener_alpha)) ||
        (rc = uds_real(u, D, "ZenerGamma", &e->zener_gamma)))
        return rc;
    e->zener_mode = (v == 1 || v == 2) ? (int32_t)v : 0;
    const UdsBlock *b;
    if ((rc = block_or_fail(u, "EvoMicrochemistry", &b))) return rc;
    for (size_t t = 0; t < b->lines.size(); t++) {
        double ti, fri;
        if ((rc = table_real(*b, t, 0, &ti)) || (rc = table_real(*b, t, 1, &fri))) return rc;
        e->disp_t.push_back(ti);
        e->disp_fr.push_back(fri);
    }
    if (e->zener_mode != 0 && e->disp_t.empty()) return fail(SCORE_SETUP_ERR_RANGE, "EvoMicrochemistry is empty");
    if (e->zener_mode == 2) {
        /* SCORE reads before the table's first entry for times up to its first time */
        bool ok = e->disp_t.size() >= 2 && e->disp_t[0] == 0.0;
        for (size_t t = 1; ok && t < e->disp_t.size(); t++) ok = e->disp_t[t] > e->disp_t[t - 1];
        if (!ok)
            return fail(SCORE_SETUP_ERR_UNSUPPORTED,
                        "time-dependent Zener drag needs an EvoMicrochemistry table from t = 0 with increasing times");
    }
    return 0;
}

int read_ideals(const score_uds *u, score_setup_ensemble *e) {
    /* init_parameter_idealtexture (SCORE_Kernel.cpp:668-712): name, Bunge angles
     * and scatter in degrees (float columns 2-5) */
    const UdsBlock *b;
    int rc = block_or_fail(u, "IdealComponents", &b);
    if (rc) return rc;
    if (b->lines.size() > kMaxInputDataLen) return fail(SCORE_SETUP_ERR_RANGE, "too many ideal components");
    for (size_t ii = 0; ii < b->lines.size(); ii++) {
        Ideal c;
        double raw;
        for (int a = 0; a < 3; a++) {
            if ((rc = table_real(*b, ii, 1 + a, &raw))) return rc;
            c.bunge[a] = deg2rad(raw);
        }
        if ((rc = table_real(*b, ii, 4, &raw))) return rc;
        c.scatter = deg2rad(raw);
        if (c.bunge[0] < kBungeMin || c.bunge[0] > kBungeMax || c.bunge[1] < kBungeMin || c.bunge[1] > kBungeMax ||
            c.bunge[2] < kBungeMin || c.bunge[2] > kBungeMax)
            return fail(SCORE_SETUP_ERR_RANGE, "Invalid Euler angles in ideal orientation defintion");
        score_orient::euler2quat(c.bunge, c.q);
        e->ideals.push_back(c);
    }
    return 0;
}

/* ens_get_closest_standardlage (SCORE_OrientationHdl.cpp:51-77): the ideal
 * component within its scatter that is closest (the last of equally close
 * ones), numbered from 1; 0 is RANDOM */
uint32_t closest_ideal(const score_setup_ensemble *e, const double *q) {
    double closest = kMaxFcc;
    uint32_t best = UINT32_MAX;
    for (uint32_t cand = 0; cand < e->ideals.size(); cand++) {
        const double disori = score_orient::disori_angle(q, e->ideals[cand].q);
        if (disori <= e->ideals[cand].scatter && disori <= closest) {
            closest = disori;
            best = cand;
        }
    }
    return best == UINT32_MAX ? 0u : best + 1u;
}

int read_schedule(const score_uds *u, score_setup_ensemble *e) {
    /* init_parameter_processing (SCORE_Kernel.cpp:716-754): time in s and
     * temperature in degrees Celsius, float columns, strictly increasing time */
    const UdsBlock *b;
    int rc = block_or_fail(u, "AnnealingSchedule", &b);
    if (rc) return rc;
    if (b->lines.size() < 2) return fail(SCORE_SETUP_ERR_RANGE, "An insufficient number (<2) of Time-Temperature data");
    double ti_old = -1.0;
    for (size_t ts = 0; ts < b->lines.size(); ts++) {
        double ti_now, Ti;
        if ((rc = table_real(*b, ts, 0, &ti_now)) || (rc = table_real(*b, ts, 1, &Ti))) return rc;
        Ti += kToffset;
        if (ti_now < 0.0 || Ti < 0.0 || ti_old >= ti_now)
            return fail(SCORE_SETUP_ERR_RANGE, "Invalid annealing schedule data (<0.0 or not consecutive in time)");
        ti_old = ti_now;
        e->ttime.push_back(ti_now);
        e->ttemp.push_back(Ti);
    }
    return 0;
}

int read_pools(const score_uds *u, score_setup_ensemble *e) {
    /* init_parameter_defgpool (SCORE_Kernel.cpp:756-801): Bunge angles in degrees
     * and rho0, float columns; orientations merged */
    const UdsBlock *b;
    int rc = block_or_fail(u, "DeformedGrainsPool", &b);
    if (rc) return rc;
    long n = (long)b->lines.size();
    if (n > kMaxDeformedGrains) n = kMaxDeformedGrains;
    for (long dg = 0; dg < n; dg++) {
        double bunge[3], raw, rho0;
        for (int a = 0; a < 3; a++) {
            if ((rc = table_real(*b, dg, a, &raw))) return rc;
            bunge[a] = deg2rad(raw);
        }
        if (bunge[0] < kBungeMin || bunge[0] > kBungeMax || bunge[1] < kBungeMin || bunge[1] > kBungeMax ||
            bunge[2] < kBungeMin || bunge[2] > kBungeMax)
            return fail(SCORE_SETUP_ERR_RANGE, "Invalid Euler angles in deformed grain orientation defintion");
        double q[4];
        score_orient::euler2quat(bunge, q); /* ens_check_disjunctness_io, SCORE_OrientationHdl.cpp:157-169 */
        WorldDefg g;
        g.ori = ens_check_disjunctness(e, q);
        if ((rc = table_real(*b, dg, 3, &rho0))) return rc;
        g.rho0 = rho0;
        g.rho = g.rho0;
        if (g.rho <= kRhoRecrystallized) return fail(SCORE_SETUP_ERR_RANGE, "Unphysical low dislocation density for deformed material");
        e->defg.push_back(g);
    }
    if (e->defg.empty()) return fail(SCORE_SETUP_ERR_RANGE, "DeformedGrainsPool is empty");
    /* init_parameter_rxgpool (SCORE_Kernel.cpp:803-845): csr_enforced nucleation
     * draws one pool index per nucleus but orients nuclei by Shoemake, so only the
     * pool's size matters; the rows are checked as SCORE checks them (the merge
     * of their orientations changes no quaternion that reaches a domain) */
    if ((rc = block_or_fail(u, "RXGrainsPool", &b))) return rc;
    n = (long)b->lines.size();
    if (n > kMaxRxGrains) n = kMaxRxGrains;
    for (long rg = 0; rg < n; rg++) {
        double bunge[3], raw, tincub;
        for (int a = 0; a < 3; a++) {
            if ((rc = table_real(*b, rg, a, &raw))) return rc;
            bunge[a] = deg2rad(raw);
        }
        if (bunge[0] < kBungeMin || bunge[0] > kBungeMax || bunge[1] < kBungeMin || bunge[1] > kBungeMax ||
            bunge[2] < kBungeMin || bunge[2] > kBungeMax)
            return fail(SCORE_SETUP_ERR_RANGE, "Invalid Euler angles in recrystallized orientation defintion");
        if ((rc = table_real(*b, rg, 3, &tincub))) return rc;
        if (tincub < 0.0) return fail(SCORE_SETUP_ERR_RANGE, "Invalid incubation time");
    }
    e->n_rxg = (uint32_t)n;
    return 0;
}

int read_defrag(const score_uds *u, score_setup_ensemble *e) {
    /* init_parameter_additional (SCORE_Kernel.cpp:918-936): X in column 2, sorted ascending */
    const UdsBlock *b;
    int rc = block_or_fail(u, "HeuristicRXFrontListDefragmentation", &b);
    if (rc) return rc;
    for (size_t d = 0; d < b->lines.size(); d++) {
        double X;
        if ((rc = table_real(*b, d, 1, &X))) return rc;
        if (X < 0.0 || X > 1.0) return fail(SCORE_SETUP_ERR_RANGE, "Invalid recrystallized volume fraction X at which to defragment list");
        e->defrag_x.push_back(X);
    }
    std::sort(e->defrag_x.begin(), e->defrag_x.end());
    return 0;
}

int read_logpoints(const score_uds *u, score_setup_ensemble *e) {
    /* init_parameter_additional (SCORE_Kernel.cpp:938-957): X in column 2, sorted ascending */
    const UdsBlock *b;
    int rc = block_or_fail(u, "SingleGrainVolOverTime", &b);
    if (rc) return rc;
    for (size_t d = 0; d < b->lines.size(); d++) {
        double X;
        if ((rc = table_real(*b, d, 1, &X))) return rc;
        if (X < 0.0 || X > 1.0)
            return fail(SCORE_SETUP_ERR_RANGE, "Invalid recrystallized volume fraction X at which to log locally all grains");
        e->log_x.push_back(X);
    }
    std::sort(e->log_x.begin(), e->log_x.end());
    return 0;
}

}  // namespace

/* ---------------------------------------------------------------- domain */

struct score_setup_domain {
    score_setup_domain_info info{};
    uint32_t nx = 0, ny = 0, nz = 0;
    double cellsize = 0;
    std::vector<uint32_t> cellgrid;
    std::vector<double> ori_quat;
    std::vector<uint32_t> defg_ori, rxg_ori, rxg_site;
    std::vector<double> defg_rho;
    score_gm_params prm{};
    std::vector<double> sched_t, sched_T, defrag_at, disp_t, disp_fr, grainlog_at;
};

extern "C" {

int score_setup_abi_version(void) { return SCORE_SETUP_ABI_VERSION; }

const char *score_setup_last_error(void) { return g_err.c_str(); }

int score_mt_init(score_mt *g, int64_t seed) {
    if (!g) return fail(SCORE_SETUP_ERR_ARG, "null generator");
    mt_init(g, seed);
    return 0;
}

double score_mt_next(score_mt *g) { return mt_next(g); }

void score_mt_shoemake(score_mt *g, double q[4]) { mt_shoemake(g, q); }

int score_uds_open(const char *path, score_uds **out) {
    if (!path || !out) return fail(SCORE_SETUP_ERR_ARG, "null argument");
    *out = nullptr;
    FILE *fh = fopen(path, "r");
    if (!fh) return fail(SCORE_SETUP_ERR_FILE, std::string("cannot open ") + path);
    /* lines as fgets with BUFSIZE gives them: a longer line continues as the next one */
    std::vector<std::string> text;
    char buf[kBufsize];
    while (fgets(buf, kBufsize, fh)) text.push_back(buf);
    fclose(fh);
    score_uds *u = new (std::nothrow) score_uds;
    if (!u) return fail(SCORE_SETUP_ERR_OOM, "out of memory");
    /* block headers: "||<name> <anything> *(<types>)" (readDataBlock, SCORE_Io.cpp:276-297) */
    std::vector<size_t> heads;
    for (size_t ln = 0; ln < text.size(); ln++) {
        const std::string &l = text[ln];
        if (l.size() < 2 || l[0] != '|' || l[1] != '|') continue;
        char name[kBufsize], skip[kBufsize], types[kBufsize];
        name[0] = types[0] = 0;
        const int cnt = sscanf(l.c_str() + 2, "%1023s%1023s%1023s", name, skip, types);
        if (cnt < 1) continue;
        UdsBlock b;
        b.name = name;
        const size_t tl = strlen(types);
        if (cnt < 3 || tl < 3 || types[0] != '*' || types[1] != '(') {
            b.defective = true;
        } else {
            size_t i = 0;
            while (2 + i < tl && types[2 + i] != ')') i++;
            if (2 + i >= tl) b.defective = true;
            else b.types.assign(types + 2, i);
        }
        u->blocks.push_back(std::move(b));
        heads.push_back(ln);
    }
    for (size_t k = 0; k < heads.size(); k++) {
        UdsBlock &b = u->blocks[k];
        if (b.defective) continue;
        parse_block_lines(b, text, heads[k] + 1);
    }
    *out = u;
    return 0;
}

void score_uds_close(score_uds *u) { delete u; }

int score_uds_real(const score_uds *u, const char *block, const char *key, double *out) {
    if (!u || !block || !key || !out) return fail(SCORE_SETUP_ERR_ARG, "null argument");
    return uds_real(u, block, key, out);
}

int score_uds_int(const score_uds *u, const char *block, const char *key, int64_t *out) {
    if (!u || !block || !key || !out) return fail(SCORE_SETUP_ERR_ARG, "null argument");
    long v;
    const int rc = uds_int(u, block, key, &v);
    if (!rc) *out = v;
    return rc;
}

int score_uds_shape(const score_uds *u, const char *block, uint32_t *rows, uint32_t *cols) {
    if (!u || !block || !rows || !cols) return fail(SCORE_SETUP_ERR_ARG, "null argument");
    const UdsBlock *b;
    const int rc = block_or_fail(u, block, &b);
    if (rc) return rc;
    *rows = (uint32_t)b->lines.size();
    *cols = (uint32_t)b->types.size();
    return 0;
}

int score_uds_table_real(const score_uds *u, const char *block, uint32_t row, uint32_t col, double *out) {
    if (!u || !block || !out) return fail(SCORE_SETUP_ERR_ARG, "null argument");
    const UdsBlock *b;
    const int rc = block_or_fail(u, block, &b);
    if (rc) return rc;
    return table_real(*b, row, col, out);
}

int score_setup_ensemble_create(const score_uds *u, score_setup_ensemble **out) {
    if (!u || !out) return fail(SCORE_SETUP_ERR_ARG, "null argument");
    *out = nullptr;
    score_setup_ensemble *e = new (std::nothrow) score_setup_ensemble;
    if (!e) return fail(SCORE_SETUP_ERR_OOM, "out of memory");
    int rc;
    /* the order of SCORE's init_parameter_* calls (SCORE_Kernel.cpp:1550-1612) */
    if ((rc = read_runtime(u, e)) || (rc = read_ensemble(u, e)) || (rc = read_nucleation(u, e)) ||
        (rc = read_physics(u, e)) || (rc = read_ideals(u, e)) || (rc = read_schedule(u, e)) || (rc = read_pools(u, e)) ||
        (rc = read_defrag(u, e)) || (rc = read_logpoints(u, e))) {
        delete e;
        return rc;
    }
    *out = e;
    return 0;
}

int score_setup_ensemble_get_info(const score_setup_ensemble *e, score_setup_ensemble_info *out) {
    if (!e || !out) return fail(SCORE_SETUP_ERR_ARG, "null argument");
    memset(out, 0, sizeof(*out));
    out->n_domains = e->n_domains;
    out->nx = e->nx;
    out->ny = e->ny;
    out->nz = e->nz;
    out->cellsize = e->cellsize;
    for (int a = 0; a < 3; a++) out->defg_mean[a] = e->defgmean[a];
    out->n_nuclei = e->nuc_density;
    out->n_world_ori = (uint32_t)e->oris.size();
    out->n_world_defg = (uint32_t)e->defg.size();
    out->n_world_rxg = e->n_rxg;
    out->n_ideal = (uint32_t)e->ideals.size();
    return 0;
}

int score_setup_ideal(const score_setup_ensemble *e, uint32_t i, double bunge[3], double q[4], double *scatter) {
    if (!e || !bunge || !q || !scatter || i >= e->ideals.size())
        return fail(SCORE_SETUP_ERR_ARG, "ideal component index out of range");
    memcpy(bunge, e->ideals[i].bunge, 3 * sizeof(double));
    memcpy(q, e->ideals[i].q, 4 * sizeof(double));
    *scatter = e->ideals[i].scatter;
    return 0;
}

int score_setup_closest_ideal(const score_setup_ensemble *e, const double q[4], uint32_t *cls) {
    if (!e || !q || !cls) return fail(SCORE_SETUP_ERR_ARG, "null argument");
    *cls = closest_ideal(e, q);
    return 0;
}

int score_setup_world_quat(const score_setup_ensemble *e, uint32_t i, double q[4]) {
    if (!e || !q || i >= e->oris.size()) return fail(SCORE_SETUP_ERR_ARG, "orientation index out of range");
    memcpy(q, e->oris[i].q, 4 * sizeof(double));
    return 0;
}

int score_setup_world_defg(const score_setup_ensemble *e, uint32_t g, uint32_t *ori, double *rho) {
    if (!e || !ori || !rho || g >= e->defg.size()) return fail(SCORE_SETUP_ERR_ARG, "grain index out of range");
    *ori = e->defg[g].ori;
    *rho = e->defg[g].rho;
    return 0;
}

void score_setup_ensemble_free(score_setup_ensemble *e) { delete e; }

int score_setup_domain_create(const score_setup_ensemble *e, uint32_t jobid, int64_t ensemble_seed,
                              score_setup_domain **out) {
    if (!e || !out) return fail(SCORE_SETUP_ERR_ARG, "null argument");
    *out = nullptr;
    score_setup_domain *s = new (std::nothrow) score_setup_domain;
    if (!s) return fail(SCORE_SETUP_ERR_OOM, "out of memory");
    try {
        /* init_cahdlprng (SCORE_Kernel.cpp:5729-5745): localprng and localmath.r share the seed */
        int64_t localseed = (int64_t)(-1 * ((int64_t)jobid + 1));
        if (ensemble_seed == kDefaultSeed) localseed = localseed - 10;
        else localseed = ensemble_seed;
        score_mt prng, mathr;
        mt_init(&prng, localseed);
        mt_init(&mathr, localseed);
        s->info.local_seed = localseed;

        /* determine_polycrystalline_geometry (SCORE_StructureSynthesis.cpp:355-392) */
        const double u = -1.0 * mt_next(&prng) * e->defgmean[0];
        const double v = -1.0 * mt_next(&prng) * e->defgmean[1];
        const double w = -1.0 * mt_next(&prng) * e->defgmean[2];
        uint32_t ngx = (uint32_t)((e->boxedge[0] - u) / e->defgmean[0]);
        ngx++;
        uint32_t ngy = (uint32_t)((e->boxedge[1] - v) / e->defgmean[1]);
        ngy++;
        uint32_t ngz = (uint32_t)((e->boxedge[2] - w) / e->defgmean[2]);
        ngz++;
        const uint64_t nseeds = (uint64_t)ngx * ngy * ngz;
        if (nseeds >= 4096000000ull) { /* CA_ALLOCATION_MAXIMUM */
            delete s;
            return fail(SCORE_SETUP_ERR_RANGE, "too many cuboid grains");
        }
        s->info.offset[0] = u;
        s->info.offset[1] = v;
        s->info.offset[2] = w;
        s->info.ngr[0] = ngx;
        s->info.ngr[1] = ngy;
        s->info.ngr[2] = ngz;

        /* pick_deformedgrains (SCORE_StructureSynthesis.cpp:394-463): one draw per
         * cuboid; a world grain picked again maps to its first local copy */
        const uint32_t nens = (uint32_t)e->defg.size();
        std::vector<uint32_t> seed_local((size_t)nseeds);
        std::unordered_map<uint32_t, uint32_t> local_of;
        local_of.reserve((size_t)std::min<uint64_t>(nseeds, nens) * 2 + 1);
        std::vector<uint32_t> world_of_local;
        for (uint64_t dg = 0; dg < nseeds; dg++) {
            const uint32_t lucky = (uint32_t)(mt_next(&prng) * nens);
            auto it = local_of.find(lucky);
            if (it == local_of.end()) {
                const uint32_t id = (uint32_t)world_of_local.size();
                local_of.emplace(lucky, id);
                world_of_local.push_back(lucky);
                seed_local[dg] = id;
            } else {
                seed_local[dg] = it->second;
            }
        }
        const uint32_t ndefg = (uint32_t)world_of_local.size();
        s->info.n_defg = ndefg;
        /* mydefgpool and its orientations (ca_check_disjunctness_core appends,
         * SCORE_OrientationHdl.cpp:204-281) */
        s->defg_ori.resize(ndefg);
        s->defg_rho.resize(ndefg);
        for (uint32_t i = 0; i < ndefg; i++) {
            const WorldDefg &g = e->defg[world_of_local[i]];
            const WorldOri &o = e->oris[g.ori];
            s->ori_quat.insert(s->ori_quat.end(), o.q, o.q + 4);
            s->defg_ori[i] = i;
            s->defg_rho[i] = g.rho;
        }

        /* ompshar_voxelize_defms (SCORE_StructureSynthesis.cpp:466-529) with one region,
         * then solve_REPLACE_CA_STRUCTURE (SCORE_Kernel.cpp:6353-6445): cuboid -> local grain */
        s->nx = e->nx;
        s->ny = e->ny;
        s->nz = e->nz;
        s->cellsize = e->cellsize;
        s->cellgrid.resize((size_t)e->nvol);
        const double cs = e->cellsize, cs2 = 0.5 * cs;
        const uint32_t ngrxy = ngx * ngy;
        size_t c = 0;
        for (uint32_t iz = 0; iz < e->nz; ++iz) {
            const uint32_t gz = (uint32_t)((((double)iz * cs) + cs2 - w) / e->defgmean[2]);
            const uint32_t gz_ngrxy = gz * ngrxy;
            for (uint32_t iy = 0; iy < e->ny; ++iy) {
                const uint32_t gy = (uint32_t)((((double)iy * cs) + cs2 - v) / e->defgmean[1]);
                const uint32_t gy_ngrx = gy * ngx;
                for (uint32_t ix = 0; ix < e->nx; ++ix) {
                    const uint32_t gx = (uint32_t)((((double)ix * cs) + cs2 - u) / e->defgmean[0]);
                    const uint32_t gxyz = gx + gy_ngrx + gz_ngrxy;
                    if (gxyz >= nseeds) {
                        delete s;
                        return fail(SCORE_SETUP_ERR_RANGE, "cell outside the cuboid grain grid");
                    }
                    s->cellgrid[c++] = seed_local[gxyz];
                }
            }
        }

        /* solve_nucmodeling_csr_enforced (SCORE_NucleationModels.cpp:207-282) */
        const uint32_t ncells = (uint32_t)e->nvol;
        uint32_t nuc = 0, maxattempts = 0;
        std::unordered_set<uint32_t> taken;
        taken.reserve((size_t)e->nuc_density * 2 + 1);
        do {
            const uint32_t lucky_site = (uint32_t)(mt_next(&prng) * ncells);
            if (taken.find(lucky_site) == taken.end()) {
                (void)(uint32_t)(mt_next(&prng) * e->n_rxg); /* luckyNucleus: drawn, not used */
                double q[4] = {1.0, 0.0, 0.0, 0.0};
                mt_shoemake(&mathr, q);
                s->ori_quat.insert(s->ori_quat.end(), q, q + 4);
                s->rxg_ori.push_back(ndefg + nuc);
                s->rxg_site.push_back(lucky_site);
                taken.insert(lucky_site);
                nuc++;
                continue;
            }
            maxattempts++;
        } while (nuc < e->nuc_density && maxattempts < kMaxAttemptsNucSite);
        s->info.nuclei_desired = e->nuc_density;
        s->info.nuclei_placed = nuc;
        s->info.collisions = maxattempts;
    } catch (const std::bad_alloc &) {
        delete s;
        return fail(SCORE_SETUP_ERR_OOM, "out of memory");
    }

    /* settings of caHdl::init_parameter (SCORE_Kernel.cpp:5747-5880) as score_gm_params */
    score_gm_params &p = s->prm;
    memset(&p, 0, sizeof(p));
    p.mobility_model = e->model;
    /* caHdl::init_parameter copies the recovery model (SCORE_Kernel.cpp:5835-5847) */
    p.recovery_mode = e->recovery_mode;
    p.rec_vacancy_d0 = e->rec_vd0;
    p.rec_vacancy_hact = e->rec_vh;
    p.rec_solute_d0 = e->rec_sd0;
    p.rec_solute_hact = e->rec_sh;
    p.rec_solute_ls = e->rec_ls;
    p.rec_solute_conc = e->rec_conc;
    p.rec_alpha3 = e->rec_alpha3;
    p.rec_kappa2 = e->rec_kappa2;
    p.rec_c3 = e->rec_c3;
    p.rec_c5 = e->rec_c5;
    p.rh_hagb_m0 = e->rh_m0;
    p.rh_hagb_hact = e->rh_h;
    p.rh_cut = e->rh_cut;
    p.rh_trans = e->rh_trans;
    p.rh_exponent = e->rh_exp;
    p.lagb_m0 = e->lagb_m0;
    p.lagb_hact = e->lagb_h;
    p.hagb_m0 = e->hagb_m0;
    p.hagb_hact = e->hagb_h;
    p.gs_m0 = e->gs_m0;
    p.gs_hact = e->gs_h;
    p.G0 = e->G0;
    p.b0 = e->b0;
    s->sched_t = e->ttime;
    s->sched_T = e->ttemp;
    p.n_sched = (uint32_t)s->sched_t.size();
    p.sched_time = s->sched_t.data();
    p.sched_temp = s->sched_T.data();
    p.maxfill = e->maxfill;
    p.xmax = e->xmax;
    p.tmax = e->tmax;
    p.nmax = e->nmax;
    p.rel_cell_caching = e->relcell;
    p.rel_cell_recaching = e->recache;
    /* log_initialization (SCORE_OnTheFlyCharacterization.cpp:51-58): X times the box volume */
    if (e->defrag_on) {
        const double boxvolume = (double)e->nvol;
        for (double X : e->defrag_x) s->defrag_at.push_back(X * boxvolume);
    }
    p.n_defrag = (uint32_t)s->defrag_at.size();
    p.defrag_at = s->defrag_at.empty() ? nullptr : s->defrag_at.data();
    {
        const double boxvolume = (double)e->nvol;
        for (double X : e->log_x) s->grainlog_at.push_back(X * boxvolume);
    }
    p.n_grainlog = (uint32_t)s->grainlog_at.size();
    p.grainlog_at = s->grainlog_at.empty() ? nullptr : s->grainlog_at.data();
    /* caHdl::init_zenerdrag (SCORE_Kernel.cpp:5904-5911) */
    p.zener_mode = e->zener_mode;
    p.zener_alpha = e->zener_alpha;
    p.zener_gamma = e->zener_gamma;
    s->disp_t = e->disp_t;
    s->disp_fr = e->disp_fr;
    p.n_disp = e->zener_mode ? (uint32_t)s->disp_t.size() : 0u;
    p.disp_time = s->disp_t.empty() ? nullptr : s->disp_t.data();
    p.disp_fr = s->disp_fr.empty() ? nullptr : s->disp_fr.data();
    *out = s;
    return 0;
}

int score_setup_domain_view(const score_setup_domain *s, score_gm_domain *dom, score_gm_params *prm) {
    if (!s || !dom || !prm) return fail(SCORE_SETUP_ERR_ARG, "null argument");
    memset(dom, 0, sizeof(*dom));
    dom->nx = s->nx;
    dom->ny = s->ny;
    dom->nz = s->nz;
    dom->cellsize = s->cellsize;
    dom->cellgrid = s->cellgrid.data();
    dom->n_ori = (uint32_t)(s->ori_quat.size() / 4);
    dom->ori_quat = s->ori_quat.data();
    dom->n_defg = (uint32_t)s->defg_ori.size();
    dom->defg_ori = s->defg_ori.data();
    dom->defg_rho = s->defg_rho.data();
    dom->n_rxg = (uint32_t)s->rxg_ori.size();
    dom->rxg_ori = s->rxg_ori.data();
    dom->rxg_site = s->rxg_site.data();
    *prm = s->prm;
    return 0;
}

int score_setup_domain_get_info(const score_setup_domain *s, score_setup_domain_info *out) {
    if (!s || !out) return fail(SCORE_SETUP_ERR_ARG, "null argument");
    *out = s->info;
    return 0;
}

void score_setup_domain_free(score_setup_domain *s) { delete s; }

}  // extern "C"
