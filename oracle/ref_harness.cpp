/*
 * ref_harness.cpp -- TEST INFRASTRUCTURE ONLY.
 *
 * C entry points into the reference's own random-number and orientation code,
 * compiled together with SCORE_Random.cpp and SCORE_Math.cpp where they lie in
 * the reference tree (oracle/Makefile, target `ref`, output oracle/_ref/).  The
 * tests use it to pin the oracle and the setup library against the reference:
 *
 *   - randomClass::init + MersenneTwister (SCORE_Random.cpp:49-76, :400-458):
 *     the stream every caHdl draws its cuboid offsets, grain picks and nucleus
 *     sites from (caHdl::init_cahdlprng, SCORE_Kernel.cpp:5729-5745);
 *   - mathMethods::rnd_quat_shoemake (SCORE_Math.cpp:660-686): nucleus
 *     orientations of csr_enforced nucleation (SCORE_NucleationModels.cpp:255);
 *   - mathMethods::disori_angle_oriquat_cubic and disori_fromoriquat_cubic_imm
 *     (SCORE_Math.cpp:491-622, :1067-1131): the disorientation that decides the
 *     mobility weight (SCORE_Kernel.cpp:2460-2527);
 *   - mathMethods::euler2quat (SCORE_Math.cpp:286-296): Bunge angles of the
 *     orientation pools to quaternions.
 *
 * Nothing here restates the reference: every value comes from its classes.
 */
#include <stdint.h>

#include "SCORE_Math.h"
#include "SCORE_Random.h"

extern "C" {

/* n values of MersenneTwister() after init(seed) */
int ref_mt_stream(long seed, uint64_t n, double *out) {
    randomClass r;
    r.init(seed);
    for (uint64_t i = 0; i < n; i++) out[i] = r.MersenneTwister();
    return 0;
}

/* n Shoemake quaternions from a math object whose generator was init(seed),
 * as caHdl::init_cahdlprng seeds localmath.r */
int ref_shoemake(long seed, uint64_t n, double *q) {
    mathMethods m;
    m.r.init(seed);
    for (uint64_t i = 0; i < n; i++) m.rnd_quat_shoemake(q + 4 * i);
    return 0;
}

/* disorientation angle (radians) of n quaternion pairs p[4i..], q[4i..] */
int ref_disori_angle(uint64_t n, const double *p, const double *q, double *out) {
    mathMethods m;
    for (uint64_t i = 0; i < n; i++) {
        double a[4] = {p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]};
        double b[4] = {q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
        out[i] = m.disori_angle_oriquat_cubic(a, b);
    }
    return 0;
}

/* disorientation quaternion of n pairs */
int ref_disori_quat(uint64_t n, const double *p, const double *q, double *out) {
    mathMethods m;
    for (uint64_t i = 0; i < n; i++) {
        double a[4] = {p[4 * i], p[4 * i + 1], p[4 * i + 2], p[4 * i + 3]};
        double b[4] = {q[4 * i], q[4 * i + 1], q[4 * i + 2], q[4 * i + 3]};
        m.disori_fromoriquat_cubic_imm(a, b, out + 4 * i);
    }
    return 0;
}

/* quaternions of n Bunge triplets (radians) */
int ref_euler2quat(uint64_t n, const double *bunge, double *q) {
    mathMethods m;
    for (uint64_t i = 0; i < n; i++) {
        double e[3] = {bunge[3 * i], bunge[3 * i + 1], bunge[3 * i + 2]};
        m.euler2quat(e, q + 4 * i);
    }
    return 0;
}

}  // extern "C"
