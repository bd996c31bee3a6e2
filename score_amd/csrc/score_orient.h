/*
 * score_orient.h -- cubic orientation arithmetic shared by the growth machine's
 * host code and the domain setup (header only, host code).
 *
 * Each function follows the reference line for line in operation order, so the
 * doubles it returns are the reference's: compile with -ffp-contract=off and
 * without -march (no FMA is formed), as the reference's GCC build.  The
 * disorientation and euler2quat are checked bit for bit against the reference's
 * own SCORE_Math.cpp (tests/test_ref_pins.py, tests/test_setup.py).
 */
#ifndef SCORE_ORIENT_H
#define SCORE_ORIENT_H

#include <math.h>

namespace score_orient {

constexpr double kPi = 3.1415926535897932384626433832795;   /* _PI_, SCORE_Math.h:54 */
constexpr double kDoubleAccuracy = 5.0e-16;                  /* DOUBLE_ACCURACY */

/* q^-1 p with the Hamilton product of multiply_quat_degraef (SCORE_Math.cpp:230-239, :564-572) */
inline void misorientation(const double *p, const double *q, double *r) {
    const double m0 = q[0], m1 = -1.0 * q[1], m2 = -1.0 * q[2], m3 = -1.0 * q[3];
    r[0] = m0 * p[0] - m1 * p[1] - m2 * p[2] - m3 * p[3];
    r[1] = p[0] * m1 + m0 * p[1] + (m2 * p[3] - m3 * p[2]);
    r[2] = p[0] * m2 + m0 * p[2] + (-(m1 * p[3]) + m3 * p[1]);
    r[3] = p[0] * m3 + m0 * p[3] + (m1 * p[2] - m2 * p[1]);
}

/* six Grimmer candidates (SCORE_Math.cpp:592-602) */
inline void grimmer(const double *r, double (*c)[4]) {
    const double a = r[0], b = r[1], cc = r[2], d = r[3];
    const double fac = 1.0 / sqrt(2.0);
    c[0][0] = (a + b) * fac; c[0][1] = (a - b) * fac; c[0][2] = (cc + d) * fac; c[0][3] = (cc - d) * fac;
    c[1][0] = (a + cc) * fac; c[1][1] = (a - cc) * fac; c[1][2] = (b + d) * fac; c[1][3] = (b - d) * fac;
    c[2][0] = (a + d) * fac; c[2][1] = (a - d) * fac; c[2][2] = (b + cc) * fac; c[2][3] = (b - cc) * fac;
    c[3][0] = (a + b + cc + d) * 0.5; c[3][1] = (a + b - cc - d) * 0.5; c[3][2] = (a - b + cc - d) * 0.5; c[3][3] = (a - b - cc + d) * 0.5;
    c[4][0] = (a + b + cc - d) * 0.5; c[4][1] = (a + b - cc + d) * 0.5; c[4][2] = (a - b + cc + d) * 0.5; c[4][3] = (a - b - cc - d) * 0.5;
    c[5][0] = a; c[5][1] = b; c[5][2] = cc; c[5][3] = d;
}

/* disori_angle_oriquat_cubic_imm (SCORE_Math.cpp:558-622), radians */
inline double disori_angle(const double *p, const double *q) {
    double r[4], c[6][4];
    misorientation(p, q, r);
    grimmer(r, c);
    double omega = 0.0;
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 4; j++)
            if (fabs(c[i][j]) > omega) omega = fabs(c[i][j]);
    if (omega < (1.0 - kDoubleAccuracy)) return 2.0 * acos(omega);
    return 0.0;
}

/* disori_fromoriquat_cubic_imm (SCORE_Math.cpp:1067-1131): the candidate with the
 * largest |component|, its absolute components in descending order */
inline void disori_quat(const double *p, const double *q, double *out) {
    double r[4], c[6][4];
    misorientation(p, q, r);
    grimmer(r, c);
    int mi = 0;
    double mx = 0.0;
    for (int i = 0; i < 6; i++)
        for (int j = 0; j < 4; j++)
            if (fabs(c[i][j]) > mx) { mx = fabs(c[i][j]); mi = i; }
    double v[4] = {fabs(c[mi][0]), fabs(c[mi][1]), fabs(c[mi][2]), fabs(c[mi][3])};
    for (int i = 0; i < 3; i++) /* ascending sort of four values */
        for (int j = 0; j < 3 - i; j++)
            if (v[j] > v[j + 1]) { const double t = v[j]; v[j] = v[j + 1]; v[j + 1] = t; }
    out[0] = v[3]; out[1] = v[2]; out[2] = v[1]; out[3] = v[0];
}

/* euler2quat_imm (SCORE_Math.cpp:334-357): Bunge (phi1, Phi, phi2) in radians */
inline void euler2quat(const double *bunge, double *q) {
    const double sigma = (bunge[0] + bunge[2]) / 2;
    const double delta = (bunge[0] - bunge[2]) / 2;
    const double c = cos(bunge[1] / 2);
    const double s = sin(bunge[1] / 2);
    q[0] = c * cos(sigma);
    q[1] = s * cos(delta);
    q[2] = s * sin(delta);
    q[3] = c * sin(sigma);
}

/* rnd_quat_shoemake (SCORE_Math.cpp:660-686) from its three uniform draws */
inline void shoemake(double X0, double X1, double X2, double *q) {
    const double r1 = sqrt(1 - X0);
    const double r2 = sqrt(X0);
    const double theta1 = 2 * kPi * X1;
    const double theta2 = 2 * kPi * X2;
    q[0] = r1 * sin(theta1);
    q[1] = r1 * cos(theta1);
    q[2] = r2 * sin(theta2);
    q[3] = r2 * cos(theta2);
    const double qnorm = 1.0 / sqrt((q[0] * q[0]) + (q[1] * q[1]) + (q[2] * q[2]) + (q[3] * q[3]));
    q[0] *= qnorm;
    q[1] *= qnorm;
    q[2] *= qnorm;
    q[3] *= qnorm;
}

}  // namespace score_orient

#endif
