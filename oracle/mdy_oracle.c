/*
 * mdy_oracle.c -- CPU restatement of MEDYAN's mechanical-equilibration path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP implementation. Loaded by
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by
 * the product library.
 *
 * Written from the reference's mathematics and behaviour. File:line citations
 * are relative to the MEDYAN source tree (src/). Operation order inside each
 * expression follows the cited lines so that the IEEE results are reproducible
 * when built with -ffp-contract=off (the parity build).
 *
 * Parity status per function is listed in mdy_oracle.h; the Simpson
 * quadrature fallbacks of the excluded-volume force field are "parity
 * unpinned" (only a run of the reference could pin their exact outputs).
 */
#include "mdy_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../medyan_amd/csrc/mdy_xmath.h"

/* Numerics mode. Reference mode (default): the C library's atan/exp and
 * sequential left-to-right sums, as the reference computes them. Exact mode:
 * the shared mdy_atan/mdy_exp and order-independent exact sums (mdy_xsum),
 * the numerics the HIP kernels use, so results can be compared bit for bit.
 * The two modes differ by rounding only (checked in tests/test_oracle_numerics.py). */
#ifndef M_PI
#define M_PI 3.14159265358979323846   /* the C library's value (not in strict C11) */
#endif
static int g_exact = 0;
void orc_set_exact(int on) { g_exact = on; }
int orc_get_exact(void) { return g_exact; }
static inline double o_atan(double x) { return g_exact ? mdy_atan(x) : atan(x); }
static inline double o_exp(double x) { return g_exact ? mdy_exp(x) : exp(x); }
static inline double o_acos(double x) { return g_exact ? mdy_acos(x) : acos(x); }
/* sin(acos(c)) for c in [-1, 1]: the C library's sin of the angle in
 * reference mode (as the reference), sqrt((1 - c)(1 + c)) in exact mode */
static inline double o_sin_of(double c, double theta) { return g_exact ? sqrt((1 - c) * (1 + c)) : sin(theta); }

/* sum helper: sequential in reference mode, exact in exact mode. In exact
 * mode every interaction energy is also added to g_total, so the manager's
 * total is the exact sum of all interaction energies rounded once. */
typedef struct osum { double seq; mdy_xsum xs; } osum;
static mdy_xsum *g_total = NULL;
static inline void os_zero(osum *o) { o->seq = 0.0; mdy_xs_zero(&o->xs); }
static inline void os_add(osum *o, double v) {
    if (g_exact) {
        mdy_xs_add(&o->xs, v);
        if (g_total) mdy_xs_add(g_total, v);
    } else {
        o->seq += v;
    }
}
static inline double os_val(const osum *o) { return g_exact ? mdy_xs_value(&o->xs) : o->seq; }

/* ------------------------------------------------------------------------- */
/* scalar helpers: src/MathFunctions.h, src/utility.h                        */
/* ------------------------------------------------------------------------- */
/* std::max / std::min semantics (first argument kept unless the comparison
 * holds), which decide what happens to NaN. */
static inline double smax(double a, double b) { return (a < b) ? b : a; }
static inline double smin(double a, double b) { return (b < a) ? b : a; }

/* areEqual, utility.h:75-77 (ZERO_PREC = 1e-6 for double, utility.h:57-66) */
static inline int are_equal(double a, double b) { return fabs(a - b) < 1e-6; }
/* areEqualLT, utility.h:80-83 */
static inline int are_equal_lt(double a, double b) { return fabs(a - b) < 1e-1; }

/* scalarProduct(v1,v2,v3,v4) = (v2-v1).(v4-v3), MathFunctions.h:469-476 */
static inline double sprod(const double *v1, const double *v2, const double *v3, const double *v4) {
    return ((v2[0] - v1[0]) * (v4[0] - v3[0]) +
            (v2[1] - v1[1]) * (v4[1] - v3[1]) +
            (v2[2] - v1[2]) * (v4[2] - v3[2]));
}
/* twoPointDistance, MathFunctions.h:294-297 */
static inline double tpdist(const double *v1, const double *v2) {
    return sqrt((v2[0] - v1[0]) * (v2[0] - v1[0]) + (v2[1] - v1[1]) * (v2[1] - v1[1]) +
                (v2[2] - v1[2]) * (v2[2] - v1[2]));
}
/* sqmagnitude(c1,c2), MathFunctions.h:210-215 */
static inline double sqmag2(const double *c1, const double *c2) {
    double x = c1[0] - c2[0], y = c1[1] - c2[1], z = c1[2] - c2[2];
    return (x * x + y * y + z * z);
}
/* midPointCoordinate(v, v1, v2, alpha), MathFunctions.h:963-970 */
static inline void midpt(double *v, const double *v1, const double *v2, double alpha) {
    double beta = 1 - alpha;
    v[0] = (v1[0] * beta + alpha * v2[0]);
    v[1] = (v1[1] * beta + alpha * v2[1]);
    v[2] = (v1[2] * beta + alpha * v2[2]);
}
/* crossProduct, MathFunctions.h:910-916 */
static inline void cross3(double *cp, const double *v1, const double *v2) {
    cp[0] = v1[1] * v2[2] - v1[2] * v2[1];
    cp[1] = v1[2] * v2[0] - v1[0] * v2[2];
    cp[2] = v1[0] * v2[1] - v1[1] * v2[0];
}
/* normalizeVector, MathFunctions.h:178-190 */
static inline void normalize3(double *v) {
    double x2 = v[0] * v[0], y2 = v[1] * v[1], z2 = v[2] * v[2];
    double norm = sqrt(x2 + y2 + z2);
    v[0] = v[0] / norm; v[1] = v[1] / norm; v[2] = v[2] / norm;
}
static inline double dot3(const double *a, const double *b) {
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

/* areInPlane, MathFunctions.h:1101-1133 */
int orc_are_in_plane(const double *p1, const double *p2, const double *p3, const double *p4) {
    double v1[3], v2[3], v3[3], cp[3];
    for (int d = 0; d < 3; d++) { v1[d] = p2[d] - p1[d]; v2[d] = p3[d] - p1[d]; v3[d] = p4[d] - p1[d]; }
    cross3(cp, v1, v2);
    normalize3(cp);
    normalize3(v3);
    return are_equal_lt(dot3(v3, cp), 0.0);
}
/* movePointOutOfPlane (in-place array version), MathFunctions.h:1221-1270 */
static void move_out_of_plane(const double *p1, const double *p2, const double *p3, const double *p4,
                              double *newp, int which, double d) {
    double nrm[3], v1[3], v2[3];
    for (int k = 0; k < 3; k++) { v1[k] = p2[k] - p1[k]; v2[k] = p3[k] - p1[k]; }
    cross3(nrm, v1, v2);
    normalize3(nrm);
    const double *src = which == 1 ? p1 : which == 2 ? p2 : which == 3 ? p3 : p4;
    for (int k = 0; k < 3; k++) newp[k] = (src[k] + nrm[k] * d);
}
static inline int finite3(const double *v) { return isfinite(v[0]) && isfinite(v[1]) && isfinite(v[2]); }

/* ------------------------------------------------------------------------- */
/* a3: FilamentStretchingHarmonic  (Filament/FilamentStretchingHarmonic.cpp:185-207, 209-275) */
/* ------------------------------------------------------------------------- */
double orc_fs_energy(const double *x, int32_t n, const int32_t *bs, const double *k, const double *l0) {
    osum U;
    os_zero(&U);
    for (int32_t i = 0; i < n; i++) {
        const double *c1 = x + bs[2 * i], *c2 = x + bs[2 * i + 1];
        double dist = tpdist(c1, c2) - l0[i];
        double Ui = 0.5 * k[i] * dist * dist;
        os_add(&U, Ui);
    }
    return os_val(&U);
}
void orc_fs_forces(const double *x, double *f, int32_t n, const int32_t *bs, const double *k, const double *l0) {
    for (int32_t i = 0; i < n; i++) {
        const double *c1 = x + bs[2 * i], *c2 = x + bs[2 * i + 1];
        double dist = tpdist(c1, c2);
        double invL = 1 / dist;
        double f0 = k[i] * (dist - l0[i]) * invL;
        double *f1 = f + bs[2 * i], *f2 = f + bs[2 * i + 1];
        for (int d = 0; d < 3; d++) {
            double r = c2[d] - c1[d];
            double F = -f0 * (r);
            f2[d] += F;
            f1[d] += -F;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* a4 / a8: cosine bending (FilamentBendingCosine.cpp:199-243, 246-398;      */
/*          BranchingBendingCosine.cpp:196-241, 243-411)                      */
/* ------------------------------------------------------------------------- */
/* U = k(1-x) for theta0 == 0, else k(1 - cos(theta - theta0)) with x = cos(theta) */
static double cosine_bend_energy(double l1l2, double L1, double L2, double k, double th0) {
    double L1L2 = L1 * L2;
    double xx = l1l2 / L1L2;
    if (xx < -1.0) xx = -1.0;
    else if (xx > 1.0) xx = 1.0;
    if (are_equal(th0, 0.0)) return k * (1 - xx);
    double cosA = xx;
    double sinA = smax(sqrt(1 - cosA * cosA), 0.0);
    double cosAmB = cosA * cos(th0) + sinA * sin(th0);
    return k * (1 - cosAmB);
}
/* effective force constant and A, B, C of the cosine bending force */
static void cosine_bend_coef(double L1sq, double L2sq, double l1l2, double kb, double th0,
                             double *A, double *B, double *C, double *keff) {
    double a = 1 / sqrt(L1sq * L2sq);
    double xx = l1l2 * a;
    *A = a;
    *B = xx / L1sq;
    *C = xx / L2sq;
    if (are_equal(th0, 0.0)) { *keff = kb; return; }
    if (fabs(fabs(xx) - 1.0) < 0.001) xx = 0.999 * xx;
    if (xx < -1.0) xx = -1.0;
    else if (xx > 1.0) xx = 1.0;
    double cosp = xx;
    double sinp = sqrt(smax((1 - cosp * cosp), 0.0));
    double sinpmq = sinp * cos(th0) - cosp * sin(th0);
    *keff = kb * sinpmq / sinp;
}
double orc_fb_energy(const double *x, int32_t n, const int32_t *bs, const double *k, const double *th0) {
    osum U;
    os_zero(&U);
    for (int32_t i = 0; i < n; i++) {
        const double *c1 = x + bs[3 * i], *c2 = x + bs[3 * i + 1], *c3 = x + bs[3 * i + 2];
        double L1 = sqrt(sprod(c1, c2, c1, c2));
        double L2 = sqrt(sprod(c2, c3, c2, c3));
        double l1l2 = sprod(c1, c2, c2, c3);
        os_add(&U, cosine_bend_energy(l1l2, L1, L2, k[i], th0[i]));
    }
    return os_val(&U);
}
void orc_fb_forces(const double *x, double *f, int32_t n, const int32_t *bs, const double *k, const double *th0) {
    for (int32_t i = 0; i < n; i++) {
        const double *c1 = x + bs[3 * i], *c2 = x + bs[3 * i + 1], *c3 = x + bs[3 * i + 2];
        double *f1 = f + bs[3 * i], *f2 = f + bs[3 * i + 1], *f3 = f + bs[3 * i + 2];
        double L1sq = sprod(c1, c2, c1, c2), L2sq = sprod(c2, c3, c2, c3), l1l2 = sprod(c1, c2, c2, c3);
        double A, B, C, ke;
        cosine_bend_coef(L1sq, L2sq, l1l2, k[i], th0[i], &A, &B, &C, &ke);
        for (int d = 0; d < 3; d++) {
            double r1 = c2[d] - c1[d], r2 = c3[d] - c2[d];
            double Fr1 = ke * (r2 * A - r1 * B);
            double Fr2 = ke * (r1 * A - r2 * C);
            f1[d] += -Fr1;
            f2[d] += Fr1 - Fr2;
            f3[d] += Fr2;
        }
    }
}
double orc_bb_energy(const double *x, int32_t n, const int32_t *bs, const double *k, const double *th0) {
    osum U;
    os_zero(&U);
    for (int32_t i = 0; i < n; i++) {
        const double *c1 = x + bs[4 * i], *c2 = x + bs[4 * i + 1], *c3 = x + bs[4 * i + 2], *c4 = x + bs[4 * i + 3];
        double L1 = sqrt(sprod(c1, c2, c1, c2));
        double L2 = sqrt(sprod(c3, c4, c3, c4));
        double l1l2 = sprod(c1, c2, c3, c4);
        os_add(&U, cosine_bend_energy(l1l2, L1, L2, k[i], th0[i]));
    }
    return os_val(&U);
}
void orc_bb_forces(const double *x, double *f, int32_t n, const int32_t *bs, const double *k, const double *th0,
                   double *stretch3) {
    for (int32_t i = 0; i < n; i++) {
        const double *c1 = x + bs[4 * i], *c2 = x + bs[4 * i + 1], *c3 = x + bs[4 * i + 2], *c4 = x + bs[4 * i + 3];
        double *f1 = f + bs[4 * i], *f2 = f + bs[4 * i + 1], *f3 = f + bs[4 * i + 2], *f4 = f + bs[4 * i + 3];
        double L1sq = sprod(c1, c2, c1, c2), L2sq = sprod(c3, c4, c3, c4), l1l2 = sprod(c1, c2, c3, c4);
        double A, B, C, ke;
        cosine_bend_coef(L1sq, L2sq, l1l2, k[i], th0[i], &A, &B, &C, &ke);
        for (int d = 0; d < 3; d++) {
            double r1 = c2[d] - c1[d], r2 = c4[d] - c3[d];
            double Fr1 = ke * (r2 * A - r1 * B);
            double Fr2 = ke * (r1 * A - r2 * C);
            f1[d] += -Fr1;
            f2[d] += Fr1;
            double f3d = -Fr2;
            f3[d] += f3d;
            if (stretch3) stretch3[3 * i + d] = f3d;
            f4[d] += Fr2;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* a5 / a6: linker and motor harmonic springs between cylinder points         */
/* (LinkerStretchingHarmonic.cpp:191-222, 225-336; MotorGhostStretchingHarmonic.cpp:234-389) */
/* ------------------------------------------------------------------------- */
double orc_ls_energy(const double *x, int32_t n, const int32_t *bs, const double *k, const double *l0,
                     const double *p1, const double *p2) {
    osum U;
    os_zero(&U);
    for (int32_t i = 0; i < n; i++) {
        double v1[3], v2[3];
        midpt(v1, x + bs[4 * i], x + bs[4 * i + 1], p1[i]);
        midpt(v2, x + bs[4 * i + 2], x + bs[4 * i + 3], p2[i]);
        double dist = tpdist(v1, v2) - l0[i];
        double Ui = 0.5 * k[i] * dist * dist;
        os_add(&U, Ui);
    }
    return os_val(&U);
}
void orc_ls_forces(const double *x, double *f, int32_t n, const int32_t *bs, const double *k, const double *l0,
                   const double *p1, const double *p2, double *stretch) {
    for (int32_t i = 0; i < n; i++) {
        double v1[3], v2[3];
        midpt(v1, x + bs[4 * i], x + bs[4 * i + 1], p1[i]);
        midpt(v2, x + bs[4 * i + 2], x + bs[4 * i + 3], p2[i]);
        double dist = tpdist(v1, v2);
        double invL = 1 / dist;
        double f0 = k[i] * (dist - l0[i]) * invL;
        double *f1 = f + bs[4 * i], *f2 = f + bs[4 * i + 1], *f3 = f + bs[4 * i + 2], *f4 = f + bs[4 * i + 3];
        for (int d = 0; d < 3; d++) {
            double l = v2[d] - v1[d];
            double Fl = -f0 * (l);
            f1[d] += -Fl * (1 - p1[i]);
            f2[d] += -Fl * (p1[i]);
            f3[d] += Fl * (1 - p2[i]);
            f4[d] += Fl * (p2[i]);
        }
        if (stretch) stretch[i] = f0 / invL;
    }
}

/* ------------------------------------------------------------------------- */
/* a7: BranchingStretchingHarmonic (BranchingStretchingHarmonic.cpp:195-221, 224-318) */
/* ------------------------------------------------------------------------- */
double orc_bs_energy(const double *x, int32_t n, const int32_t *bs, const double *k, const double *l0, const double *p) {
    osum U;
    os_zero(&U);
    for (int32_t i = 0; i < n; i++) {
        double v1[3];
        midpt(v1, x + bs[3 * i], x + bs[3 * i + 1], p[i]);
        double dist = tpdist(v1, x + bs[3 * i + 2]) - l0[i];
        double Ui = 0.5 * k[i] * dist * dist;
        os_add(&U, Ui);
    }
    return os_val(&U);
}
void orc_bs_forces(const double *x, double *f, int32_t n, const int32_t *bs, const double *k, const double *l0,
                   const double *p, double *stretch3) {
    for (int32_t i = 0; i < n; i++) {
        const double *c3 = x + bs[3 * i + 2];
        double *f1 = f + bs[3 * i], *f2 = f + bs[3 * i + 1], *f3 = f + bs[3 * i + 2];
        double v1[3];
        midpt(v1, x + bs[3 * i], x + bs[3 * i + 1], p[i]);
        double dist = tpdist(v1, c3);
        double invL = 1 / dist;
        double f0 = k[i] * (dist - l0[i]) * invL;
        for (int d = 0; d < 3; d++) {
            double r = c3[d] - v1[d];
            f1[d] += -f0 * (r) * (p[i] - 1);
            f2[d] += f0 * (r) * p[i];
            f3[d] += -f0 * (r);
            if (stretch3) stretch3[3 * i + d] = -f0 * (r);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Branching position / dihedral, cosine forms                               */
/* BranchingPositionCosine.cpp:190-248 (energy), 306-531 (forces);           */
/* BranchingDihedralCosine.cpp:211-269 (energy), 271-661 (forces).           */
/* A branch point at the mother's plus end (areEqual(pos, 1)) sits at 0.5 of */
/* a virtual cylinder with plus end c2' = (1/0.5)(c2 - 0.5 c1).               */
/* ------------------------------------------------------------------------- */
static inline int are_equal1(double a) { return fabs(a - 1.0) < 1e-6; }
static void extend_plus(double *c2p, const double *c1, const double *c2, double position) {
    for (int d = 0; d < 3; d++) c2p[d] = (1 / position) * (c2[d] - (1 - position) * c1[d]);
}
double orc_bp_energy(const double *x, int32_t n, const int32_t *bs, const double *kpos, const double *pos) {
    osum U;
    os_zero(&U);
    const double posheta = 0.5 * M_PI;
    for (int32_t i = 0; i < n; i++) {
        const double *c1 = x + bs[3 * i], *c2 = x + bs[3 * i + 1], *c3 = x + bs[3 * i + 2];
        double c2p[3], mp[3];
        double position = pos[i];
        if (are_equal1(position)) {
            position = 0.5;
            extend_plus(c2p, c1, c2, position);
            c2 = c2p;
        }
        midpt(mp, c1, c2, position);
        double X = sqrt(sprod(mp, c2, mp, c2));
        double D = sqrt(sprod(mp, c3, mp, c3));
        double XD = X * D;
        double xd = sprod(mp, c2, mp, c3);
        double xx = xd / XD;
        if (fabs(fabs(xx) - 1.0) < 0.001) {
            xd = 0.999 * XD;
            xx = xd / XD;
        }
        if (xx < -1.0) xx = -1.0;
        else if (xx > 1.0) xx = 1.0;
        double cosp = xx;
        double sinp = sqrt(smax(1 - cosp * cosp, 0.0));
        double cospminusq = cosp * cos(posheta) + sinp * sin(posheta);
        os_add(&U, kpos[i] * (1 - cospminusq));
    }
    return os_val(&U);
}
void orc_bp_forces(const double *x, double *f, int32_t n, const int32_t *bs, const double *kpos, const double *pos,
                   double *stretch3) {
    const double posheta = 0.5 * M_PI;
    for (int32_t i = 0; i < n; i++) {
        const double *c1 = x + bs[3 * i], *c2 = x + bs[3 * i + 1], *c3 = x + bs[3 * i + 2];
        double *f1 = f + bs[3 * i], *f2 = f + bs[3 * i + 1], *f3 = f + bs[3 * i + 2];
        double c2p[3], mp[3];
        double position = pos[i];
        if (are_equal1(position)) {
            position = 0.5;
            extend_plus(c2p, c1, c2, position);
            c2 = c2p;
        }
        midpt(mp, c1, c2, position);
        double Xsq = sprod(mp, c2, mp, c2);
        double Dsq = sprod(mp, c3, mp, c3);
        double xd = sprod(mp, c2, mp, c3);
        double A = 1 / sqrt(Xsq * Dsq);
        double xx = xd * A;
        double B = xx / Xsq;
        double C = xx / Dsq;
        if (fabs(fabs(xx) - 1.0) < 0.001) xx = 0.999 * xx;
        if (xx < -1.0) xx = -1.0;
        else if (xx > 1.0) xx = 1.0;
        double cosp = xx;
        double sinp = sqrt(smax(1 - cosp * cosp, 0.0));
        double sinpminusq = sinp * cos(posheta) - cosp * sin(posheta);
        double k = kpos[i] * sinpminusq / sinp;
        double f1t[3], f2t[3], f3t[3];
        for (int d = 0; d < 3; d++) {
            double r1 = c2[d] - mp[d], r2 = c3[d] - mp[d];
            double Fr1 = k * (r2 * A - r1 * B);
            double Fr2 = k * (r1 * A - r2 * C);
            double fmp = -Fr1 - Fr2;
            f1t[d] = (1 - position) * fmp;
            f2t[d] = Fr1 + position * fmp;
            f3t[d] = Fr2;
        }
        for (int d = 0; d < 3; d++) {
            f3[d] += f3t[d];
            if (stretch3) stretch3[3 * i + d] = f3t[d];
        }
        if (are_equal1(pos[i])) {
            double factor = (position - 1) / position;
            for (int d = 0; d < 3; d++) {
                f1[d] += f1t[d] + f2t[d] * factor;
                f2[d] += f2t[d] * (1 / position);
            }
        } else {
            for (int d = 0; d < 3; d++) {
                f1[d] += f1t[d];
                f2[d] += f2t[d];
            }
        }
    }
}
double orc_bdh_energy(const double *x, int32_t n, const int32_t *bs, const double *kdih, const double *pos) {
    osum U;
    os_zero(&U);
    for (int32_t i = 0; i < n; i++) {
        const double *c1 = x + bs[4 * i], *c2 = x + bs[4 * i + 1], *c3 = x + bs[4 * i + 2], *c4 = x + bs[4 * i + 3];
        double c2p[3], mp[3], a[3], b[3], n1[3], n2[3];
        double position = pos[i];
        if (are_equal1(position)) {
            position = 0.5;
            extend_plus(c2p, c1, c2, position);
            c2 = c2p;
        }
        midpt(mp, c1, c2, position);
        /* vectorProduct(v, v1, v2, v3, v4) = (v2 - v1) x (v4 - v3), MathFunctions.h:695-716 */
        for (int d = 0; d < 3; d++) { a[d] = c2[d] - mp[d]; b[d] = c3[d] - mp[d]; }
        cross3(n1, a, b);
        for (int d = 0; d < 3; d++) { a[d] = c4[d] - c3[d]; b[d] = c3[d] - mp[d]; }
        cross3(n2, a, b);
        normalize3(n1);
        normalize3(n2);
        double n1n2 = n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2];
        os_add(&U, kdih[i] * (1 - n1n2));
    }
    return os_val(&U);
}
void orc_bdh_forces(const double *x, double *f, int32_t n, const int32_t *bs, const double *kdih, const double *pos,
                    double *stretch3) {
    for (int32_t i = 0; i < n; i++) {
        double c1[3], c2[3], c3[3], c4[3], mp[3];
        for (int d = 0; d < 3; d++) {
            c1[d] = x[bs[4 * i] + d]; c2[d] = x[bs[4 * i + 1] + d];
            c3[d] = x[bs[4 * i + 2] + d]; c4[d] = x[bs[4 * i + 3] + d];
        }
        double position = pos[i];
        if (are_equal1(position)) {
            position = 0.5;
            for (int d = 0; d < 3; d++) c2[d] = (1 / position) * (c2[d] - (1 - position) * c1[d]);
        }
        midpt(mp, c1, c2, position);
        double *f1 = f + bs[4 * i], *f2 = f + bs[4 * i + 1], *f3 = f + bs[4 * i + 2], *f4 = f + bs[4 * i + 3];
        double vb1x = c2[0] - mp[0], vb1y = c2[1] - mp[1], vb1z = c2[2] - mp[2];
        double vb2x = c3[0] - mp[0], vb2y = c3[1] - mp[1], vb2z = c3[2] - mp[2];
        double vb2xm = -vb2x, vb2ym = -vb2y, vb2zm = -vb2z;
        double vb3x = c4[0] - c3[0], vb3y = c4[1] - c3[1], vb3z = c4[2] - c3[2];
        double ax = vb1y * vb2zm - vb1z * vb2ym, ay = vb1z * vb2xm - vb1x * vb2zm, az = vb1x * vb2ym - vb1y * vb2xm;
        double bx = vb3y * vb2zm - vb3z * vb2ym, by = vb3z * vb2xm - vb3x * vb2zm, bz = vb3x * vb2ym - vb3y * vb2xm;
        double rasq = ax * ax + ay * ay + az * az;
        double rbsq = bx * bx + by * by + bz * bz;
        double rgsq = vb2xm * vb2xm + vb2ym * vb2ym + vb2zm * vb2zm;
        double rg = sqrt(rgsq);
        double rginv = 0.0, ra2inv = 0.0, rb2inv = 0.0;
        if (rg > 0) rginv = 1.0 / rg;
        if (rasq > 0) ra2inv = 1.0 / rasq;
        if (rbsq > 0) rb2inv = 1.0 / rbsq;
        double rabinv = sqrt(ra2inv * rb2inv);
        double s = rg * rabinv * (ax * vb3x + ay * vb3y + az * vb3z);
        double fg = vb1x * vb2xm + vb1y * vb2ym + vb1z * vb2zm;
        double hg = vb3x * vb2xm + vb3y * vb2ym + vb3z * vb2zm;
        double fga = fg * ra2inv * rginv;
        double hgb = hg * rb2inv * rginv;
        double gaa = -ra2inv * rg;
        double gbb = rb2inv * rg;
        double dtf[3] = {gaa * ax, gaa * ay, gaa * az};
        double dtg[3] = {fga * ax - hgb * bx, fga * ay - hgb * by, fga * az - hgb * bz};
        double dth[3] = {gbb * bx, gbb * by, gbb * bz};
        double df = -kdih[i] * s;
        double L1[3], L2[3], L3[3], L4[3];
        for (int d = 0; d < 3; d++) {
            double s2 = df * dtg[d];
            L4[d] = df * dth[d];
            L3[d] = -s2 - L4[d];
            if (are_equal1(pos[i])) {
                double l2p = df * dtf[d];
                L2[d] = s2 - l2p;
                L1[d] = l2p * ((position - 1) / position);
                L2[d] += l2p * (1 / position);
            } else {
                L2[d] = df * dtf[d];
                double lmp = s2 - L2[d];
                L1[d] = (1 - position) * lmp;
                L2[d] += (position) * lmp;
            }
        }
        for (int d = 0; d < 3; d++) {
            f1[d] += L1[d]; f2[d] += L2[d]; f3[d] += L3[d]; f4[d] += L4[d];
            if (stretch3) stretch3[3 * i + d] = L3[d];
        }
    }
}

/* BranchingDihedralQuadratic (BranchingDihedralQuadratic.cpp:26-86 energy,
 * 88-398 forces): E = kdih theta^2. The energy takes cos(theta) from the
 * normals n1 = b1 x b2, n2 = b3 x b2 (b1 = c2 - mp, b2 = c3 - mp,
 * b3 = c4 - c3); the forces take it from the bond angles,
 * (cos<b1,b3> - cos<b1,b2> cos<b3,b2>) / (sin<b1,b2> sin<b3,b2>) with the
 * sines held at >= 1e-3, and use dE = eFac d(cos theta),
 * eFac = -2 kdih theta / max(sin theta, 1e-5). */
double orc_bdhq_energy(const double *x, int32_t n, const int32_t *bs, const double *kdih, const double *pos) {
    osum U;
    os_zero(&U);
    for (int32_t i = 0; i < n; i++) {
        const double *c1 = x + bs[4 * i], *c2 = x + bs[4 * i + 1], *c3 = x + bs[4 * i + 2], *c4 = x + bs[4 * i + 3];
        double c2p[3], mp[3], b1[3], b2[3], b3[3], n1[3], n2[3];
        double p = pos[i];
        if (are_equal1(p)) {
            p = 0.5;
            extend_plus(c2p, c1, c2, p);
            c2 = c2p;
        }
        midpt(mp, c1, c2, p);
        for (int d = 0; d < 3; d++) { b1[d] = c2[d] - mp[d]; b2[d] = c3[d] - mp[d]; b3[d] = c4[d] - c3[d]; }
        cross3(n1, b1, b2);
        cross3(n2, b3, b2);
        double n1mag = sqrt(n1[0] * n1[0] + n1[1] * n1[1] + n1[2] * n1[2]);
        double n2mag = sqrt(n2[0] * n2[0] + n2[1] * n2[1] + n2[2] * n2[2]);
        double ct = (n1[0] * n2[0] + n1[1] * n2[1] + n1[2] * n2[2]) / (n1mag * n2mag);
        if (ct < -1.0) ct = -1.0;
        if (ct > 1.0) ct = 1.0;
        double th = o_acos(ct);
        os_add(&U, kdih[i] * th * th);
    }
    return os_val(&U);
}
static inline double dmax(double a, double b) { return (a < b) ? b : a; }   /* std::max */
void orc_bdhq_forces(const double *x, double *f, int32_t n, const int32_t *bs, const double *kdih, const double *pos,
                     double *stretch3) {
    for (int32_t i = 0; i < n; i++) {
        double c1[3], c2[3], c3[3], c4[3], mp[3], b1[3], b2[3], b3[3];
        for (int d = 0; d < 3; d++) {
            c1[d] = x[bs[4 * i] + d]; c2[d] = x[bs[4 * i + 1] + d];
            c3[d] = x[bs[4 * i + 2] + d]; c4[d] = x[bs[4 * i + 3] + d];
        }
        double p = pos[i];
        if (are_equal1(p)) {
            p = 0.5;
            for (int d = 0; d < 3; d++) c2[d] = (1 / p) * (c2[d] - (1 - p) * c1[d]);
        }
        midpt(mp, c1, c2, p);
        for (int d = 0; d < 3; d++) { b1[d] = c2[d] - mp[d]; b2[d] = c3[d] - mp[d]; b3[d] = c4[d] - c3[d]; }
        const double l1sq = b1[0] * b1[0] + b1[1] * b1[1] + b1[2] * b1[2];
        const double l2sq = b2[0] * b2[0] + b2[1] * b2[1] + b2[2] * b2[2];
        const double l3sq = b3[0] * b3[0] + b3[1] * b3[1] + b3[2] * b3[2];
        const double l1 = sqrt(l1sq), l2 = sqrt(l2sq), l3 = sqrt(l3sq);
        const double i1sq = 1.0 / l1sq, i2sq = 1.0 / l2sq, i3sq = 1.0 / l3sq;
        const double i13 = 1.0 / (l1 * l3), i12 = 1.0 / (l1 * l2), i32 = 1.0 / (l3 * l2);
        const double d13 = b1[0] * b3[0] + b1[1] * b3[1] + b1[2] * b3[2];
        const double d12 = b1[0] * b2[0] + b1[1] * b2[1] + b1[2] * b2[2];
        const double d32 = b3[0] * b2[0] + b3[1] * b2[1] + b3[2] * b2[2];
        const double c12 = d12 * i12;
        const double s12 = dmax(sqrt(dmax(1 - c12 * c12, 0.0)), 0.001);
        const double is12 = 1.0 / s12, is12sq = is12 * is12;
        const double c32 = d32 * i32;
        const double s32 = dmax(sqrt(dmax(1 - c32 * c32, 0.0)), 0.001);
        const double is32 = 1.0 / s32, is32sq = is32 * is32;
        const double c13 = d13 * i13;
        const double num = c13 - c12 * c32, iden = is12 * is32;
        double ct = num * iden;
        if (ct < -1.0) ct = -1.0;
        if (ct > 1.0) ct = 1.0;
        const double th = o_acos(ct);
        const double st = dmax(o_sin_of(ct, th), 1e-5);
        const double eFac = -2 * kdih[i] * th * (1.0 / st);
        /* dE/db_i = sum_j e_ij b_j */
        const double e11 = -eFac * ct * is12sq * i1sq;
        const double e12 = eFac * i12 * (-c32 * iden + c12 * ct * is12sq);
        const double e13 = eFac * i13 * iden;
        const double e22 = eFac * i2sq * (2 * c13 * iden - ct * (is12sq + is32sq));
        const double e23 = eFac * i32 * (-c12 * iden + c32 * ct * is32sq);
        const double e33 = -eFac * ct * is32sq * i3sq;
        /* chain rule through b1 = c2 - mp, b2 = c3 - mp, b3 = c4 - c3, mp = (1-p) c1 + p c2 */
        const double w[4][3] = {
            {(1 - p) * (e11 + e12), (1 - p) * (e12 + e22), (1 - p) * (e13 + e23)},
            {(p - 1) * e11 + p * e12, (p - 1) * e12 + p * e22, (p - 1) * e13 + p * e23},
            {-e12 + e13, -e22 + e23, -e23 + e33},
            {-e13, -e23, -e33}};
        double L[4][3];
        for (int k = 0; k < 4; k++)
            for (int d = 0; d < 3; d++) L[k][d] = w[k][0] * b1[d] + w[k][1] * b2[d] + w[k][2] * b3[d];
        if (are_equal1(pos[i])) {   /* the energy is a function of (c1, c2', c3, c4) */
            const double factor1 = (p - 1) / p, factor2 = 1 / p;
            for (int d = 0; d < 3; d++) {
                L[0][d] = L[0][d] + L[1][d] * factor1;
                L[1][d] = L[1][d] * factor2;
            }
        }
        double *F[4] = {f + bs[4 * i], f + bs[4 * i + 1], f + bs[4 * i + 2], f + bs[4 * i + 3]};
        for (int d = 0; d < 3; d++) {
            for (int k = 0; k < 4; k++) F[k][d] += L[k][d];
            if (stretch3) stretch3[3 * i + d] = L[2][d];
        }
    }
}

/* ------------------------------------------------------------------------- */
/* a9: CylinderExclVolRepulsion (Volume/CylinderExclVolRepulsion.cpp)          */
/* ------------------------------------------------------------------------- */
/* Geometric invariants of a cylinder pair (c1,c2) x (c3,c4), :355-392 / :626-672. */
typedef struct ev_inv {
    double a, b, c, d, e, F, g, h, I;
    double AA, BB, CC, DD, EE, FF, GG, HH, JJ;
    double ATG1, ATG2, ATG3, ATG4;
} ev_inv;

static void ev_invariants(const double *c1, const double *c2, const double *c3, const double *c4, ev_inv *v) {
    v->a = sprod(c1, c2, c1, c2);
    v->b = sprod(c3, c4, c3, c4);
    v->c = sprod(c3, c1, c3, c1);
    v->d = sprod(c1, c2, c3, c4);
    v->e = sprod(c1, c2, c3, c1);
    v->F = sprod(c3, c4, c3, c1);
    v->g = sprod(c1, c2, c1, c4);
    v->h = sprod(c3, c4, c3, c2);
    v->I = sprod(c3, c4, c1, c4);
    double sqD = sqmag2(c1, c4), sqE = sqmag2(c3, c2);
    double vA[3], vB[3], vC[3], cp[3];
    for (int k = 0; k < 3; k++) { vA[k] = c1[k] - c2[k]; vB[k] = c3[k] - c4[k]; vC[k] = c3[k] - c1[k]; }
    double a = v->a, b = v->b, c = v->c, d = v->d, e = v->e, F = v->F, g = v->g, h = v->h, I = v->I;
    double ac = sqrt(a * c), bc = sqrt(b * c), aD = sqrt(a * sqD), bE = sqrt(b * sqE);
    v->AA = sqrt(smax(0.0, (ac + e) * (ac - e)));
    v->BB = sqrt(smax(0.0, (bc + F) * (bc - F)));
    v->CC = d * e - a * F;
    v->DD = b * e - d * F;
    v->EE = sqrt(smax(0.0, (aD + g) * (aD - g)));
    v->FF = sqrt(smax(0.0, (bE + h) * (bE - h)));
    v->GG = d * g - a * I;
    v->HH = v->CC + v->GG - v->DD;
    cross3(cp, vA, vB);
    double JJ = dot3(cp, vC);
    v->JJ = -JJ * JJ;
    double AA = v->AA, BB = v->BB, EE = v->EE, FF = v->FF;
    v->ATG1 = o_atan((a + e) / AA) - o_atan(e / AA);
    v->ATG2 = o_atan((a + e - d) / EE) - o_atan((e - d) / EE);
    v->ATG3 = o_atan((F) / BB) - o_atan((F - b) / BB);
    v->ATG4 = o_atan((d + F) / FF) - o_atan((d + F - b) / FF);
}

double orc_ev_pair_energy_analytic(const double *c1, const double *c2, const double *c3, const double *c4, double kr) {
    ev_inv v;
    ev_invariants(c1, c2, c3, c4, &v);
    /* U = 1/2 kR / JJ * (CC/AA ATG1 + GG/EE ATG2 + DD/BB ATG3 + HH/FF ATG4), :393-435 */
    return 0.5 * kr / v.JJ *
           (v.CC / v.AA * v.ATG1 + v.GG / v.EE * v.ATG2 + v.DD / v.BB * v.ATG3 + v.HH / v.FF * v.ATG4);
}

/* 2-D composite Simpson rule over s,t in [0,1] of 1/r^4 as coded in energyN
 * (:884-999): the running abscissae are advanced before their first use, so
 * interior nodes sit at (i+1)/N; the reference's sum structure is kept. */
static inline double r2_st(double a, double b, double c, double d, double e, double F, double s, double t) {
    return c + 2 * s * e + s * s * a - 2 * t * F - 2 * s * t * d + t * t * b;   /* CylinderExclVolRepulsion.h:53-58 */
}
static inline double eint(double a, double b, double c, double d, double e, double F, double s, double t) {
    double r2 = r2_st(a, b, c, d, e, F, s, t);
    return (1.0 / (r2 * r2));
}
double orc_ev_pair_energy_simpson(const double *c1, const double *c2, const double *c3, const double *c4,
                                  double kr, int N1, int N2) {
    double a = sprod(c1, c2, c1, c2), b = sprod(c3, c4, c3, c4), c = sprod(c3, c1, c3, c1);
    double d = sprod(c1, c2, c3, c4), e = sprod(c1, c2, c3, c1), F = sprod(c3, c4, c3, c1);
    double ds = 1.0 / ((double)N1), dt = 1.0 / ((double)N2);
    double T1 = eint(a, b, c, d, e, F, 0.0, 0.0) + eint(a, b, c, d, e, F, 0.0, 1.0) +
                eint(a, b, c, d, e, F, 1.0, 0.0) + eint(a, b, c, d, e, F, 1.0, 1.0);
    double T2 = 0.0, T3 = 0.0, T4 = 0.0, s = 0.0, t = 0.0;
    mdy_xsum X;   /* exact mode: one exact accumulator over all node terms */
    mdy_xs_zero(&X);
    mdy_xs_add(&X, T1);
    for (int i = 0; i <= N1; i++) {
        t = 0.0;
        s = g_exact ? (double)(i + 1) * ds : s + ds;
        double fs = (i % 2 == 0) ? 2.0 : 4.0;
        if (i > 0 && i < N1) {
            double v = fs * (eint(a, b, c, d, e, F, s, 0.0) + eint(a, b, c, d, e, F, s, 1.0));
            if (g_exact) mdy_xs_add(&X, v); else T2 += v;
        }
        for (int j = 0; j <= N2; j++) {
            t = g_exact ? (double)(j + 1) * dt : t + dt;
            double ft = (j % 2 == 0) ? 2.0 : 4.0;
            if (j > 0 && j < N2) {
                if (i == 0) {
                    double v = ft * (eint(a, b, c, d, e, F, 0.0, t) + eint(a, b, c, d, e, F, 1.0, t));
                    if (g_exact) mdy_xs_add(&X, v); else T3 += v;
                } else if (i < N1) {
                    double v = fs * ft * eint(a, b, c, d, e, F, s, t);
                    if (g_exact) mdy_xs_add(&X, v); else T4 += v;
                }
            }
        }
    }
    double T = g_exact ? mdy_xs_value(&X) : (T1 + T2 + T3 + T4);
    return kr * (ds * dt / 9.0) * T;
}

/* forceN (:1001-1254): Simpson integrals of {1, s, s^2, st, t, t^2}/r^6 and the
 * force assembly; f12 receives the 4x3 forces. */
static void fint(double a, double b, double c, double d, double e, double F, double s, double t, double *o) {
    double r2 = r2_st(a, b, c, d, e, F, s, t);
    double xx = (1.0 / (r2 * r2 * r2));   /* CylinderExclVolRepulsion.h:60-72 */
    o[0] = xx; o[1] = s * xx; o[2] = s * s * xx; o[3] = s * t * xx; o[4] = t * xx; o[5] = t * t * xx;
}
void orc_ev_pair_force_simpson(const double *c1, const double *c2, const double *c3, const double *c4,
                               double kr, int N1, int N2, double *f12) {
    double a = sprod(c1, c2, c1, c2), b = sprod(c3, c4, c3, c4), c = sprod(c3, c1, c3, c1);
    double d = sprod(c1, c2, c3, c4), e = sprod(c1, c2, c3, c1), F = sprod(c3, c4, c3, c1);
    double vA[3], vB[3], vC[3];
    for (int k = 0; k < 3; k++) { vA[k] = c1[k] - c2[k]; vB[k] = c3[k] - c4[k]; vC[k] = c3[k] - c1[k]; }
    double ds = 1.0 / ((double)N1), dt = 1.0 / ((double)N2);
    double T1[6] = {0}, T2[6] = {0}, T3[6] = {0}, T4[6] = {0}, o[6];
    mdy_xsum X[6];
    for (int q = 0; q < 6; q++) mdy_xs_zero(&X[q]);
    const double lim[2] = {0.0, 1.0};
    for (int i = 0; i < 2; i++)
        for (int j = 0; j < 2; j++) {
            fint(a, b, c, d, e, F, lim[i], lim[j], o);
            for (int q = 0; q < 6; q++) {
                if (g_exact) mdy_xs_add(&X[q], o[q]); else T1[q] += o[q];
            }
        }
    double s = 0.0, t = 0.0;
    for (int i = 0; i <= N1; i++) {
        t = 0.0;
        s = g_exact ? (double)(i + 1) * ds : s + ds;
        double fs = (i % 2 == 0) ? 2.0 : 4.0;
        if (i > 0 && i < N1)
            for (int l = 0; l < 2; l++) {
                fint(a, b, c, d, e, F, s, lim[l], o);
                for (int q = 0; q < 6; q++) {
                    if (g_exact) mdy_xs_add(&X[q], fs * o[q]); else T2[q] += fs * o[q];
                }
            }
        for (int j = 0; j <= N2; j++) {
            t = g_exact ? (double)(j + 1) * dt : t + dt;
            double ft = (j % 2 == 0) ? 2.0 : 4.0;
            if (j > 0 && j < N2) {
                if (i == 0) {
                    for (int l = 0; l < 2; l++) {
                        fint(a, b, c, d, e, F, lim[l], t, o);
                        for (int q = 0; q < 6; q++) {
                            if (g_exact) mdy_xs_add(&X[q], ft * o[q]); else T3[q] += ft * o[q];
                        }
                    }
                } else if (i < N1) {
                    fint(a, b, c, d, e, F, s, t, o);
                    for (int q = 0; q < 6; q++) {
                        if (g_exact) mdy_xs_add(&X[q], fs * ft * o[q]); else T4[q] += fs * ft * o[q];
                    }
                }
            }
        }
    }
    double in[6];
    for (int q = 0; q < 6; q++)
        in[q] = (ds * dt / 9.0) * (g_exact ? mdy_xs_value(&X[q]) : (T1[q] + T2[q] + T3[q] + T4[q]));
    for (int k = 0; k < 3; k++) {
        f12[0 + k] = 4.0 * kr * (-in[0] * vC[k] + in[1] * (vC[k] - vA[k]) + in[2] * vA[k] + in[4] * vB[k] - in[3] * vB[k]);
        f12[3 + k] = 4.0 * kr * (-in[1] * vC[k] - in[2] * vA[k] + in[3] * vB[k]);
        f12[6 + k] = 4.0 * kr * (in[0] * vC[k] + in[1] * vA[k] - in[4] * (vB[k] + vC[k]) - in[3] * vA[k] + in[5] * vB[k]);
        f12[9 + k] = 4.0 * kr * (in[4] * vC[k] + in[3] * vA[k] - in[5] * vB[k]);
    }
}

static inline int ev_energy_bad(double U) { return fabs(U) == INFINITY || U != U || U < -1.0; }

/* CylinderExclVolRepulsion::energy, :329-575. The per-call side table of pair
 * energies (cylEnergies, :558-565) has no arithmetic effect and is omitted. */
double orc_ev_energy(const double *x, int32_t n, const int32_t *bs, const double *krep, const double *eql) {
    osum U;
    os_zero(&U);
    for (int32_t i = 0; i < n; i++) {
        double kr = krep[i] * eql[2 * i] * eql[2 * i + 1];
        const double *c1 = x + bs[4 * i], *c2 = x + bs[4 * i + 1], *c3 = x + bs[4 * i + 2], *c4 = x + bs[4 * i + 3];
        double Ui = orc_ev_pair_energy_analytic(c1, c2, c3, c4, kr);
        if (ev_energy_bad(Ui)) {
            /* energyN without bead moves; its retry result is discarded by the reference (:989-995) */
            Ui = orc_ev_pair_energy_simpson(c1, c2, c3, c4, kr, 400, 400);
            if (ev_energy_bad(Ui)) return -1;
        }
        os_add(&U, Ui);
    }
    return os_val(&U);
}

/* Analytic force of one pair, :626-860. out[4][3] = forces on c1..c4, with the
 * direction vectors taken from the (possibly moved) c2. */
static void ev_pair_force_analytic(const double *c1, const double *c2, const double *c3, const double *c4,
                                   double kr, double out[12]) {
    ev_inv v;
    ev_invariants(c1, c2, c3, c4, &v);
    const double a = v.a, b = v.b, c = v.c, d = v.d, e = v.e, F = v.F;
    const double AA = v.AA, BB = v.BB, CC = v.CC, DD = v.DD, EE = v.EE, FF = v.FF, GG = v.GG, HH = v.HH;
    const double ATG1 = v.ATG1, ATG2 = v.ATG2, ATG3 = v.ATG3, ATG4 = v.ATG4;
    const double invJJ = 1 / v.JJ;
    const double sumBlock = CC / AA * ATG1 + GG / EE * ATG2 + DD / BB * ATG3 + HH / FF * ATG4;
    const double enFactor = kr * invJJ / 2;
    /* derivatives of the arctan differences */
    const double A1 = AA * AA / (AA * AA + (a + e) * (a + e));
    const double A2 = AA * AA / (AA * AA + e * e);
    const double E1 = EE * EE / (EE * EE + (a + e - d) * (a + e - d));
    const double E2 = EE * EE / (EE * EE + (e - d) * (e - d));
    const double B1 = BB * BB / (BB * BB + F * F);
    const double B2 = BB * BB / (BB * BB + (F - b) * (F - b));
    const double F1 = FF * FF / (FF * FF + (d + F) * (d + F));
    const double F2 = FF * FF / (FF * FF + (d + F - b) * (d + F - b));
    /* partials of the four bracketed blocks w.r.t. their free parameters */
    const double pA_C = ATG1 / AA;
    const double pA_A = -(ATG1 * CC) / (AA * AA) + (-A1 * (a + e) + A2 * e) * CC / (AA * AA * AA);
    const double pA_e = ((A1 - A2) * CC) / (AA * AA);
    const double pA_a = (A1 * CC) / (AA * AA);
    const double pE_G = ATG2 / EE;
    const double pE_E = -(ATG2 * GG) / (EE * EE) + (-E1 * (a + e - d) + E2 * (e - d)) * GG / (EE * EE * EE);
    const double pE_emd = ((E1 - E2) * GG) / (EE * EE);
    const double pE_a = (E1 * GG) / (EE * EE);
    const double pB_D = ATG3 / BB;
    const double pB_B = -(ATG3 * DD) / (BB * BB) + (-B1 * F + B2 * (F - b)) * DD / (BB * BB * BB);
    const double pB_fmb = ((B1 - B2) * DD) / (BB * BB);
    const double pB_b = (B1 * DD) / (BB * BB);
    const double pF_H = ATG4 / FF;
    const double pF_F = -(ATG4 * HH) / (FF * FF) + (-F1 * (d + F) + F2 * (d + F - b)) * HH / (FF * FF * FF);
    const double pF_dfb = ((F1 - F2) * HH) / (FF * FF);
    const double pF_b = (F1 * HH) / (FF * FF);
    /* chain rule onto the basis v1 = c2-c1, v2 = c4-c3, v3 = c1-c3 (:703-760) */
    const double k11 = pA_C * (-2 * F) + pA_A * (c / AA) + pA_a * 2 + pE_G * (2 * (F - b)) + pE_E * ((b + c - 2 * F) / EE) +
                       pE_a * 2 + pF_H * (-2 * b) + pF_F * (b / FF);
    const double k12 = pA_C * e + pE_G * (2 * d - e) + pE_E * ((e - d) / EE) + pE_emd * (-1) + pB_D * (-F) +
                       pF_H * (2 * d + F) + pF_F * (-(d + F) / FF) + pF_dfb;
    const double k13 = pA_C * d + pA_A * (-e / AA) + pA_e + pE_G * (-d) + pE_E * (-(e - d) / EE) + pE_emd + pB_D * b +
                       pF_H * (-b) + pF_F * (b / FF);
    const double k21 = k12;
    const double k22 = pE_G * (-2 * a) + pE_E * (a / EE) + pB_D * (2 * e) + pB_B * (c / BB) + pB_fmb * (-2) + pB_b * 2 +
                       pF_H * (-2 * (a + e)) + pF_F * ((a + c + 2 * e) / FF) + pF_dfb * (-2) + pF_b * 2;
    const double k23 = pA_C * (-a) + pE_G * a + pE_E * (-a / EE) + pB_D * (-d) + pB_B * (-F / BB) + pB_fmb + pF_H * d +
                       pF_F * (-(d + F) / FF) + pF_dfb;
    const double k31 = k13;
    const double k32 = k23;
    const double k33 = pA_A * (a / AA) + pE_E * (a / EE) + pB_B * (b / BB) + pF_F * (b / FF);
    /* derivative of JJ (:818-828) */
    const double J11 = 2 * (-b * c + F * F), J12 = 2 * (c * d - e * F), J13 = 2 * (b * e - d * F);
    const double J21 = 2 * (c * d - e * F), J22 = 2 * (-a * c + e * e), J23 = 2 * (-d * e + a * F);
    const double J31 = 2 * (b * e - d * F), J32 = 2 * (-d * e + a * F), J33 = 2 * (d * d - a * b);
    double dv[4][3];
    dv[0][0] = enFactor * (-invJJ * sumBlock * (J31 - J11) + (k31 - k11));
    dv[0][1] = enFactor * (-invJJ * sumBlock * (J32 - J12) + (k32 - k12));
    dv[0][2] = enFactor * (-invJJ * sumBlock * (J33 - J13) + (k33 - k13));
    dv[1][0] = enFactor * (-invJJ * sumBlock * (J11) + (k11));
    dv[1][1] = enFactor * (-invJJ * sumBlock * (J12) + (k12));
    dv[1][2] = enFactor * (-invJJ * sumBlock * (J13) + (k13));
    dv[2][0] = enFactor * (-invJJ * sumBlock * (-J31 - J21) + (-k31 - k21));
    dv[2][1] = enFactor * (-invJJ * sumBlock * (-J32 - J22) + (-k32 - k22));
    dv[2][2] = enFactor * (-invJJ * sumBlock * (-J33 - J23) + (-k33 - k23));
    dv[3][0] = enFactor * (-invJJ * sumBlock * (J21) + (k21));
    dv[3][1] = enFactor * (-invJJ * sumBlock * (J22) + (k22));
    dv[3][2] = enFactor * (-invJJ * sumBlock * (J23) + (k23));
    for (int bi = 0; bi < 4; bi++)
        for (int k = 0; k < 3; k++)
            out[3 * bi + k] = -(dv[bi][0] * (c2[k] - c1[k]) + dv[bi][1] * (c4[k] - c3[k]) + dv[bi][2] * (c1[k] - c3[k]));
}

/* forceN with the reference's retry ladder (:1001-1254): first without bead
 * moves; if still non-finite, with c1 and c2 moved off the plane when
 * areInPlane holds; if still non-finite the reference exits. */
static int ev_pair_force_fallback(const double *c1, const double *c2, const double *c3, const double *c4, double kr,
                                  double out[12]) {
    orc_ev_pair_force_simpson(c1, c2, c3, c4, kr, 10000, 10000, out);
    int ok = 1;
    for (int q = 0; q < 12; q++) ok &= isfinite(out[q]);
    if (ok) return 0;
    if (orc_are_in_plane(c1, c2, c3, c4)) {
        double n2[3], n1[3];
        move_out_of_plane(c1, c2, c3, c4, n2, 2, 0.01);
        move_out_of_plane(c1, c2, c3, c4, n1, 1, 0.01);
        orc_ev_pair_force_simpson(n1, n2, c3, c4, kr, 10000, 10000, out);
    } else {
        orc_ev_pair_force_simpson(c1, c2, c3, c4, kr, 10000, 10000, out);
    }
    ok = 1;
    for (int q = 0; q < 12; q++) ok &= isfinite(out[q]);
    return ok ? 0 : MDY_ERR_FORCE;
}

/* CylinderExclVolRepulsion::forces, :578-882 */
int orc_ev_forces(const double *x, double *f, int32_t n, const int32_t *bs, const double *krep, const double *eql,
                  int coplanar_mode) {
    for (int32_t i = 0; i < n; i++) {
        double kr = krep[i] * eql[2 * i] * eql[2 * i + 1];
        const double *c1 = x + bs[4 * i], *c2 = x + bs[4 * i + 1], *c3 = x + bs[4 * i + 2], *c4 = x + bs[4 * i + 3];
        double nc2[3];
        const double *c2u = c2;
        if (coplanar_mode == MDY_COPLANAR_REFERENCE && orc_are_in_plane(c1, c2, c3, c4)) {
            move_out_of_plane(c1, c2, c3, c4, nc2, 2, 0.01);
            c2u = nc2;
        }
        double out[12];
        ev_pair_force_analytic(c1, c2u, c3, c4, kr, out);
        int ok = finite3(out) && finite3(out + 3) && finite3(out + 6) && finite3(out + 9);
        if (!ok) {
            int rc = ev_pair_force_fallback(c1, c2, c3, c4, kr, out);
            if (rc) return rc;
        }
        double *f1 = f + bs[4 * i], *f2 = f + bs[4 * i + 1], *f3 = f + bs[4 * i + 2], *f4 = f + bs[4 * i + 3];
        for (int k = 0; k < 3; k++) {
            f1[k] += out[k];
            f2[k] += out[3 + k];
            f3[k] += out[6 + k];
            f4[k] += out[9 + k];
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------- */
/* monomer-sum excluded volume: CylinderVolumeMon                            */
/* (Volume/CylinderVolumeMon.hpp:31-100 energy/force, :126-185 vectorize)    */
/* Per interaction: mon[6] = minus serial, plus serial, interval of the first */
/* cylinder, then of the second; meq[2] = monomer sizes; kvol.               */
/* ------------------------------------------------------------------------- */
/* ceildiv(minus, w) * w: the first sampled monomer serial */
static int64_t first_sample(int64_t minus, int64_t w) {
    int64_t q = minus / w;       /* truncates toward zero: a ceiling below zero */
    if (q * w < minus) q++;
    return q * w;
}
/* energy(info, coords), :31-59 */
static double evm_pair_energy(const double *c11, const double *c12, const double *c21, const double *c22,
                              const int32_t *m, const double *meq, double kvol) {
    const int64_t num1 = (int64_t)m[1] - m[0], num2 = (int64_t)m[4] - m[3];
    const int64_t s1 = first_sample(m[0], m[2]), s2 = first_sample(m[3], m[5]);
    const double seg1 = m[2] * meq[0], seg2 = m[5] * meq[1];
    const double coeff = kvol * seg1 * seg2;
    double en = 0;
    for (int64_t a = s1; a < m[1]; a += m[2]) {
        const double loc1 = (a - m[0] + 0.5) / num1;
        double p1[3];
        for (int k = 0; k < 3; k++) p1[k] = c11[k] + loc1 * (c12[k] - c11[k]);
        for (int64_t b = s2; b < m[4]; b += m[5]) {
            const double loc2 = (b - m[3] + 0.5) / num2;
            double p2[3], d2 = 0;
            for (int k = 0; k < 3; k++) p2[k] = c21[k] + loc2 * (c22[k] - c21[k]);
            for (int k = 0; k < 3; k++) d2 += (p2[k] - p1[k]) * (p2[k] - p1[k]);
            en += coeff / (d2 * d2);
        }
    }
    return en;
}
/* force(info, coords, forces), :61-100: subtracts into f11..f22 */
static void evm_pair_force(const double *c11, const double *c12, const double *c21, const double *c22,
                           const int32_t *m, const double *meq, double kvol, double *f11, double *f12, double *f21,
                           double *f22) {
    const int64_t num1 = (int64_t)m[1] - m[0], num2 = (int64_t)m[4] - m[3];
    const int64_t s1 = first_sample(m[0], m[2]), s2 = first_sample(m[3], m[5]);
    const double seg1 = m[2] * meq[0], seg2 = m[5] * meq[1];
    const double coeff = kvol * seg1 * seg2;
    const double coeff4 = coeff * 4;
    for (int64_t a = s1; a < m[1]; a += m[2]) {
        const double loc1 = (a - m[0] + 0.5) / num1;
        double p1[3];
        for (int k = 0; k < 3; k++) p1[k] = c11[k] + loc1 * (c12[k] - c11[k]);
        for (int64_t b = s2; b < m[4]; b += m[5]) {
            const double loc2 = (b - m[3] + 0.5) / num2;
            double p2[3], r12[3], d2 = 0;
            for (int k = 0; k < 3; k++) p2[k] = c21[k] + loc2 * (c22[k] - c21[k]);
            for (int k = 0; k < 3; k++) r12[k] = p2[k] - p1[k];
            for (int k = 0; k < 3; k++) d2 += r12[k] * r12[k];
            const double g = coeff4 / (d2 * d2 * d2);
            for (int k = 0; k < 3; k++) {
                const double grad = g * r12[k];
                f11[k] -= (1 - loc1) * grad;
                f12[k] -= loc1 * grad;
                f21[k] -= -(1 - loc2) * grad;
                f22[k] -= -loc2 * grad;
            }
        }
    }
}
/* CylinderVolumeMon::computeEnergy, :189-195 */
double orc_evm_energy(const double *x, int32_t n, const int32_t *bs, const int32_t *mon, const double *meq,
                      const double *kvol) {
    osum U;
    os_zero(&U);
    for (int32_t i = 0; i < n; i++)
        os_add(&U, evm_pair_energy(x + bs[4 * i], x + bs[4 * i + 1], x + bs[4 * i + 2], x + bs[4 * i + 3], mon + 6 * i,
                                   meq + 2 * i, kvol[i]));
    return os_val(&U);
}
/* CylinderVolumeMon::computeForces, :196-200. The reference subtracts every
 * monomer pair's share straight into the force array; in the exact mode each
 * interaction's four bead forces start from zero and are added afterwards, the
 * association of the per-bead gather (the two differ by rounding). */
void orc_evm_forces(const double *x, double *f, int32_t n, const int32_t *bs, const int32_t *mon, const double *meq,
                    const double *kvol) {
    for (int32_t i = 0; i < n; i++) {
        const int32_t *b = bs + 4 * i;
        if (g_exact) {
            double loc[12] = {0};
            evm_pair_force(x + b[0], x + b[1], x + b[2], x + b[3], mon + 6 * i, meq + 2 * i, kvol[i], loc, loc + 3,
                           loc + 6, loc + 9);
            for (int k = 0; k < 3; k++) {
                f[b[0] + k] += loc[k];
                f[b[1] + k] += loc[3 + k];
                f[b[2] + k] += loc[6 + k];
                f[b[3] + k] += loc[9 + k];
            }
        } else {
            evm_pair_force(x + b[0], x + b[1], x + b[2], x + b[3], mon + 6 * i, meq + 2 * i, kvol[i], f + b[0],
                           f + b[1], f + b[2], f + b[3]);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* a10: BoundaryCylinderRepulsionExp (Boundary/BoundaryCylinderRepulsionExp.cpp:27-62, 65-110) */
/* PlaneBoundaryElement::distance (Structure/BoundaryElementImpl.h:71-74)      */
/* ------------------------------------------------------------------------- */
static inline double plane_dist(const double *pl, const double *p) {
    return (pl[0] * p[0] + pl[1] * p[1] + pl[2] * p[2] + pl[3]) /
           sqrt(pow(pl[0], 2) + pow(pl[1], 2) + pow(pl[2], 2));
}
double orc_bd_energy(const double *x, int32_t nplanes, const double *plane, const int32_t *nneigh,
                     const int32_t *bs, const double *krep, const double *slen) {
    osum U;
    os_zero(&U);
    int64_t cum = 0;
    for (int32_t ib = 0; ib < nplanes; ib++) {
        for (int32_t ic = 0; ic < nneigh[ib]; ic++) {
            double r = plane_dist(plane + 4 * ib, x + bs[cum + ic]);
            double R = -r / slen[cum + ic];
            double Ui = krep[cum + ic] * o_exp(R);
            os_add(&U, Ui);
        }
        cum += nneigh[ib];
    }
    return os_val(&U);
}
void orc_bd_forces(const double *x, double *f, int32_t nplanes, const double *plane, const int32_t *nneigh,
                   const int32_t *bs, const double *krep, const double *slen) {
    int64_t cum = 0;
    for (int32_t ib = 0; ib < nplanes; ib++) {
        const double *pl = plane + 4 * ib;
        for (int32_t ic = 0; ic < nneigh[ib]; ic++) {
            double r = plane_dist(pl, x + bs[cum + ic]);
            double R = -r / slen[cum + ic];
            double f0 = krep[cum + ic] * o_exp(R) / slen[cum + ic];
            double *f1 = f + bs[cum + ic];
            f1[0] += f0 * pl[0];
            f1[1] += f0 * pl[1];
            f1[2] += f0 * pl[2];
        }
        cum += nneigh[ib];
    }
}

/* Boundary load forces: ForceFieldManager::computeLoadForces
 * (ForceFieldManager.cpp:292-309) -> BoundaryCylinderRepulsion::computeLoadForces
 * (BoundaryCylinderRepulsion.cpp:167-189) -> boundaryCylinderRepulsionLoadForce
 * (:124-163): for each plane in order and each cylinder of its neighbor list
 * (midpoint of x_nl within the float cutoff, NeighborListImpl.cpp:738-751),
 * a plus end (Cylinder::isPlusEnd: the last cylinder of its filament) loads
 * its second bead, a minus end its first: dir = normalizedVector(bd - bo)
 * (Util/Math/Vec.hpp:350-374), monomer i at bd + (i*monSize)*dir,
 * proj = max(-n.dir, 0) with the raw plane normal (BoundaryElementImpl.h:97-103),
 * load = K exp(-r/lambda)/lambda (BoundaryCylinderRepulsionExp.cpp:113-118),
 * loadForces[i] += proj*load. lp/lm: n_beads x max_mon, zeroed first. */
static void load_one(const double *pl, double krep, double slen, const double *bo, const double *bd, double mon,
                     int nmon, double *out) {
    double v[3] = {bd[0] - bo[0], bd[1] - bo[1], bd[2] - bo[2]};
    double mag2 = 0.0;
    for (int k = 0; k < 3; k++) mag2 += v[k] * v[k];
    double norm = sqrt(mag2);
    double dir[3];
    for (int k = 0; k < 3; k++) dir[k] = v[k] / norm;
    for (int i = 0; i < nmon; i++) {
        double step = i * mon;
        double nc[3] = {bd[0] + step * dir[0], bd[1] + step * dir[1], bd[2] + step * dir[2]};
        double dt = 0.0;
        for (int k = 0; k < 3; k++) dt += pl[k] * dir[k];
        double proj = (-dt < 0.0) ? 0.0 : -dt;
        double r = plane_dist(pl, nc);
        double R = -r / slen;
        double load = krep * o_exp(R) / slen;
        out[i] += proj * load;
    }
}
static const mdy_cylinders *g_sort_cyl;
static int cmp_fil_pos(const void *a, const void *b) {
    int32_t i = *(const int32_t *)a, j = *(const int32_t *)b;
    const mdy_cylinders *c = g_sort_cyl;
    if (c->filament[i] != c->filament[j]) return c->filament[i] < c->filament[j] ? -1 : 1;
    if (c->position[i] != c->position[j]) return c->position[i] < c->position[j] ? -1 : 1;
    return (i > j) - (i < j);
}
void orc_load_forces(const mdy_cylinders *cyl, int32_t n_beads, const double *x_nl, const double *x,
                     const mdy_boundary *bnd, int32_t n_types, const double *mon_size, const int32_t *n_mon,
                     int32_t max_mon, double *lp, double *lm) {
    for (int64_t k = 0; k < (int64_t)n_beads * max_mon; k++) { lp[k] = 0.0; lm[k] = 0.0; }
    unsigned char *plus = (unsigned char *)calloc((size_t)(cyl->n > 0 ? cyl->n : 1), 1);
    {   /* the last cylinder of each filament: sort by (filament, position) */
        int32_t *ord = (int32_t *)malloc(sizeof(int32_t) * (size_t)(cyl->n > 0 ? cyl->n : 1));
        for (int32_t i = 0; i < cyl->n; i++) ord[i] = i;
        g_sort_cyl = cyl;
        qsort(ord, (size_t)cyl->n, sizeof(int32_t), cmp_fil_pos);
        for (int32_t k = 0; k < cyl->n; k++)
            if (k + 1 == cyl->n || cyl->filament[ord[k + 1]] != cyl->filament[ord[k]]) plus[ord[k]] = 1;
        free(ord);
    }
    for (int32_t ib = 0; ib < bnd->n_planes; ib++) {
        const double *pl = bnd->plane + 4 * ib;
        for (int32_t i = 0; i < cyl->n; i++) {
            double mc[3];
            const double *b1 = x_nl + cyl->bead1[i], *b2 = x_nl + cyl->bead2[i];
            for (int k = 0; k < 3; k++) mc[k] = b1[k] * (1.0 - 0.5) + 0.5 * b2[k];
            if (!(plane_dist(pl, mc) < bnd->cutoff)) continue;
            int ty = (cyl->type && cyl->type[i] < n_types) ? cyl->type[i] : 0;
            if (plus[i])
                load_one(pl, bnd->k[ib], bnd->screen[ib], x + cyl->bead1[i], x + cyl->bead2[i], mon_size[ty],
                         n_mon[ty], lp + (int64_t)(cyl->bead2[i] / 3) * max_mon);
            if (cyl->minus_end[i])
                load_one(pl, bnd->k[ib], bnd->screen[ib], x + cyl->bead2[i], x + cyl->bead1[i], mon_size[ty],
                         n_mon[ty], lm + (int64_t)(cyl->bead1[i] / 3) * max_mon);
        }
    }
    free(plus);
}

/* ------------------------------------------------------------------------- */
/* a13: neighbor lists                                                       */
/* ------------------------------------------------------------------------- */
/* Cylinder database midpoint 0.5*(b1+b2) (Structure/Cylinder.cpp:45-49). */
static inline void cyl_mid(const mdy_cylinders *cyl, const double *x, int32_t i, double *m) {
    const double *b1 = x + cyl->bead1[i], *b2 = x + cyl->bead2[i];
    for (int k = 0; k < 3; k++) m[k] = 0.5 * (b1[k] + b2[k]);
}

/* Predicate of HybridCylinderCylinderNL::updateNeighborsbin
 * (HybridNeighborListImpl.cpp:405-441): half list by id, same-filament
 * neighbours within minCylinderDistanceSameFilament = 2 (SysParams.h:315)
 * excluded, squared midpoint distance (double) against float rMin^2/rMax^2
 * (HybridNeighborListImpl.h:78-80). The bin stencil only prunes candidates, so
 * the brute-force pair set equals the reference's. */
int64_t orc_nl_pairs(const mdy_cylinders *cyl, const double *x, float rmax, float rmin, int32_t *pairs, int64_t cap) {
    float rmaxsq = rmax * rmax, rminsq = rmin * rmin;
    int32_t n = cyl->n;
    double *mid = (double *)malloc(sizeof(double) * 3 * (size_t)(n > 0 ? n : 1));
    for (int32_t i = 0; i < n; i++) cyl_mid(cyl, x, i, mid + 3 * i);
    int64_t cnt = 0;
    for (int32_t i = 0; i < n; i++) {
        for (int32_t j = 0; j < n; j++) {
            if (cyl->id[i] <= cyl->id[j]) continue;
            if (cyl->filament[i] == cyl->filament[j]) {
                int ds = abs(cyl->position[i] - cyl->position[j]);
                if (ds <= 2) continue;
            }
            const double *m1 = mid + 3 * i, *m2 = mid + 3 * j;
            double d2 = 0;
            for (int k = 0; k < 3; k++) d2 += (m2[k] - m1[k]) * (m2[k] - m1[k]);
            if (d2 < rminsq || d2 > rmaxsq) continue;
            if (cnt < cap) { pairs[2 * cnt] = i; pairs[2 * cnt + 1] = j; }
            cnt++;
        }
    }
    free(mid);
    return cnt;
}

/* The same pair set and order with candidates pruned by a uniform grid of
 * cells at least 1.01 rMax wide (midpoints of a listed pair differ by at most
 * rMax per axis, so their cells are neighbours); the predicate is the one
 * above, evaluated in the same order. For large networks, where the O(n^2)
 * loop is too slow for tests. */
static int cmp_i32(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

int64_t orc_nl_pairs_grid(const mdy_cylinders *cyl, const double *x, float rmax, float rmin, int32_t *pairs,
                          int64_t cap) {
    float rmaxsq = rmax * rmax, rminsq = rmin * rmin;
    int32_t n = cyl->n;
    if (n <= 0) return 0;
    double *mid = (double *)malloc(sizeof(double) * 3 * (size_t)n);
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY}, h[3];
    int64_t nc[3];
    for (int32_t i = 0; i < n; i++) {
        cyl_mid(cyl, x, i, mid + 3 * i);
        for (int k = 0; k < 3; k++) { lo[k] = smin(lo[k], mid[3 * i + k]); hi[k] = smax(hi[k], mid[3 * i + k]); }
    }
    for (int k = 0; k < 3; k++) {
        h[k] = smax(1.01 * (double)rmax, (hi[k] - lo[k]) / 512.0);
        if (!(h[k] > 0)) h[k] = 1.0;
        nc[k] = (int64_t)((hi[k] - lo[k]) / h[k]) + 1;
    }
    const int64_t ncell = nc[0] * nc[1] * nc[2];
    int64_t *cell = (int64_t *)malloc(sizeof(int64_t) * (size_t)n);
    int32_t *start = (int32_t *)calloc((size_t)ncell + 1, sizeof(int32_t));
    int32_t *list = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    int32_t *cand = (int32_t *)malloc(sizeof(int32_t) * (size_t)n);
    int64_t *ci3 = (int64_t *)malloc(sizeof(int64_t) * 3 * (size_t)n);
    for (int32_t i = 0; i < n; i++) {
        for (int k = 0; k < 3; k++) {
            int64_t c = (int64_t)((mid[3 * i + k] - lo[k]) / h[k]);
            if (c >= nc[k]) c = nc[k] - 1;
            ci3[3 * i + k] = c;
        }
        cell[i] = (ci3[3 * i] * nc[1] + ci3[3 * i + 1]) * nc[2] + ci3[3 * i + 2];
        start[cell[i] + 1]++;
    }
    for (int64_t c = 0; c < ncell; c++) start[c + 1] += start[c];
    {
        int32_t *cur = (int32_t *)malloc(sizeof(int32_t) * (size_t)ncell);
        memcpy(cur, start, sizeof(int32_t) * (size_t)ncell);
        for (int32_t i = 0; i < n; i++) list[cur[cell[i]]++] = i;
        free(cur);
    }
    int64_t cnt = 0;
    for (int32_t i = 0; i < n; i++) {
        int32_t nk = 0;
        for (int64_t dx = -1; dx <= 1; dx++)
            for (int64_t dy = -1; dy <= 1; dy++)
                for (int64_t dz = -1; dz <= 1; dz++) {
                    int64_t a = ci3[3 * i] + dx, b = ci3[3 * i + 1] + dy, c = ci3[3 * i + 2] + dz;
                    if (a < 0 || b < 0 || c < 0 || a >= nc[0] || b >= nc[1] || c >= nc[2]) continue;
                    int64_t cc = (a * nc[1] + b) * nc[2] + c;
                    for (int32_t q = start[cc]; q < start[cc + 1]; q++) {
                        int32_t j = list[q];
                        if (cyl->id[i] <= cyl->id[j]) continue;
                        if (cyl->filament[i] == cyl->filament[j]) {
                            int ds = abs(cyl->position[i] - cyl->position[j]);
                            if (ds <= 2) continue;
                        }
                        const double *m1 = mid + 3 * i, *m2 = mid + 3 * j;
                        double d2 = 0;
                        for (int k = 0; k < 3; k++) d2 += (m2[k] - m1[k]) * (m2[k] - m1[k]);
                        if (d2 < rminsq || d2 > rmaxsq) continue;
                        cand[nk++] = j;
                    }
                }
        qsort(cand, (size_t)nk, sizeof(int32_t), cmp_i32);
        for (int32_t q = 0; q < nk; q++) {
            if (cnt < cap) { pairs[2 * cnt] = i; pairs[2 * cnt + 1] = cand[q]; }
            cnt++;
        }
    }
    free(mid); free(cell); free(start); free(list); free(cand); free(ci3);
    return cnt;
}

/* CylinderExclVolume::vectorize (Volume/CylinderExclVolume.cpp:34-179):
 * drop pairs whose neighbour is the mother of the cylinder
 * (cin->getBranchingCylinder() == ci); k is the max over types when they
 * differ. */
int64_t orc_ev_vectorize(const mdy_cylinders *cyl, const int32_t *pairs, int64_t npairs, int32_t *bs, double *krep,
                         double *eql) {
    int64_t m = 0;
    for (int64_t p = 0; p < npairs; p++) {
        int32_t ci = pairs[2 * p], cn = pairs[2 * p + 1];
        if (cyl->branch_child[cn] == ci) continue;
        bs[4 * m] = cyl->bead1[ci]; bs[4 * m + 1] = cyl->bead2[ci];
        bs[4 * m + 2] = cyl->bead1[cn]; bs[4 * m + 3] = cyl->bead2[cn];
        eql[2 * m] = cyl->eq_length[ci]; eql[2 * m + 1] = cyl->eq_length[cn];
        if (cyl->type[ci] != cyl->type[cn]) krep[m] = smax(cyl->exvol_k[ci], cyl->exvol_k[cn]);
        else krep[m] = cyl->exvol_k[ci];
        m++;
    }
    return m;
}

/* CylinderVolumeMon::vectorize (Volume/CylinderVolumeMon.hpp:126-185): drop
 * mother-daughter pairs in either order; kvol is the larger constant. */
int64_t orc_evm_vectorize(const mdy_cylinders *cyl, const int32_t *pairs, int64_t npairs, int32_t *bs, int32_t *mon,
                          double *meq, double *kvol) {
    int64_t m = 0;
    for (int64_t p = 0; p < npairs; p++) {
        int32_t ci = pairs[2 * p], cn = pairs[2 * p + 1];
        if (cyl->branch_child[ci] == cn || cyl->branch_child[cn] == ci) continue;
        bs[4 * m] = cyl->bead1[ci]; bs[4 * m + 1] = cyl->bead2[ci];
        bs[4 * m + 2] = cyl->bead1[cn]; bs[4 * m + 3] = cyl->bead2[cn];
        int32_t *mm = mon + 6 * m;
        mm[0] = cyl->mon_minus[ci]; mm[1] = cyl->mon_plus[ci]; mm[2] = cyl->mon_interval[ci];
        mm[3] = cyl->mon_minus[cn]; mm[4] = cyl->mon_plus[cn]; mm[5] = cyl->mon_interval[cn];
        meq[2 * m] = cyl->mon_size[ci]; meq[2 * m + 1] = cyl->mon_size[cn];
        kvol[m] = smax(cyl->exvol_k[ci], cyl->exvol_k[cn]);
        m++;
    }
    return m;
}

/* BoundaryCylinderNL::updateNeighbors (NeighborListImpl.cpp:738-751) with the
 * float cutoff (NeighborList.h:38-39), then BoundaryCylinderRepulsion::vectorize
 * (BoundaryCylinderRepulsion.cpp:33-88): the first bead of minus-end cylinders,
 * then the second bead of every listed cylinder. */
int64_t orc_bd_vectorize(const mdy_cylinders *cyl, const double *x, const mdy_boundary *bnd, int32_t *nneigh,
                         int32_t *bs, double *krep, double *slen, int64_t cap) {
    int64_t m = 0;
    for (int32_t ib = 0; ib < bnd->n_planes; ib++) {
        const double *pl = bnd->plane + 4 * ib;
        int32_t cnt = 0;
        for (int32_t i = 0; i < cyl->n; i++) {
            double mc[3];
            const double *b1 = x + cyl->bead1[i], *b2 = x + cyl->bead2[i];
            /* Cylinder::coordinate = midPointCoordinate(b1, b2, 0.5) */
            for (int k = 0; k < 3; k++) mc[k] = b1[k] * (1.0 - 0.5) + 0.5 * b2[k];
            double dist = plane_dist(pl, mc);
            if (!(dist < bnd->cutoff)) continue;
            if (cyl->minus_end[i]) {
                if (m < cap) { bs[m] = cyl->bead1[i]; krep[m] = bnd->k[ib]; slen[m] = bnd->screen[ib]; }
                m++; cnt++;
            }
            if (m < cap) { bs[m] = cyl->bead2[i]; krep[m] = bnd->k[ib]; slen[m] = bnd->screen[ib]; }
            m++; cnt++;
        }
        nneigh[ib] = cnt;
    }
    return m;
}

/* ------------------------------------------------------------------------- */
/* a11: ForceFieldManager                                                    */
/* ------------------------------------------------------------------------- */
int orc_system_setup(orc_system *o, const mdy_system *sys, const double *x) {
    memset(o, 0, sizeof(*o));
    o->sys = *sys;
    const mdy_cylinders *cyl = &sys->cyl;
    if (sys->exvol_enabled) {
        int64_t cap = 1024;
        int32_t *pairs = NULL;
        int64_t np;
        for (;;) {
            pairs = (int32_t *)realloc(pairs, sizeof(int32_t) * 2 * (size_t)cap);
            np = orc_nl_pairs_grid(cyl, x, sys->exvol_cutoff, 0.0f, pairs, cap);
            if (np <= cap) break;
            cap = np;
        }
        o->ev_bs = (int32_t *)malloc(sizeof(int32_t) * 4 * (size_t)(np + 1));
        o->ev_krep = (double *)malloc(sizeof(double) * (size_t)(np + 1));
        o->ev_eql = (double *)malloc(sizeof(double) * 2 * (size_t)(np + 1));
        if (sys->exvol_type == MDY_EXVOL_MONOMER) {
            o->ev_mon = (int32_t *)malloc(sizeof(int32_t) * 6 * (size_t)(np + 1));
            o->ev_n = orc_evm_vectorize(cyl, pairs, np, o->ev_bs, o->ev_mon, o->ev_eql, o->ev_krep);
        } else {
            o->ev_n = orc_ev_vectorize(cyl, pairs, np, o->ev_bs, o->ev_krep, o->ev_eql);
        }
        free(pairs);
    }
    if (sys->boundary_enabled && sys->bnd.n_planes > 0) {
        int64_t cap = 2 * (int64_t)cyl->n * sys->bnd.n_planes + 1;
        o->bd_nneigh = (int32_t *)malloc(sizeof(int32_t) * (size_t)sys->bnd.n_planes);
        o->bd_bs = (int32_t *)malloc(sizeof(int32_t) * (size_t)cap);
        o->bd_krep = (double *)malloc(sizeof(double) * (size_t)cap);
        o->bd_slen = (double *)malloc(sizeof(double) * (size_t)cap);
        o->bd_n = orc_bd_vectorize(cyl, x, &sys->bnd, o->bd_nneigh, o->bd_bs, o->bd_krep, o->bd_slen, cap);
    }
    return 0;
}

void orc_system_free(orc_system *o) {
    free(o->ev_bs); free(o->ev_krep); free(o->ev_eql); free(o->ev_mon);
    free(o->bd_nneigh); free(o->bd_bs); free(o->bd_krep); free(o->bd_slen);
    memset(o, 0, sizeof(*o));
}

static void ffm_each(const orc_system *o, const double *x, double *e) {
    const mdy_ff_tables *t = &o->sys.ff;
    e[MDY_FF_FS] = orc_fs_energy(x, t->fs_n, t->fs_beadset, t->fs_k, t->fs_l0);
    e[MDY_FF_FB] = orc_fb_energy(x, t->fb_n, t->fb_beadset, t->fb_k, t->fb_theta0);
    e[MDY_FF_LS] = orc_ls_energy(x, t->ls_n, t->ls_beadset, t->ls_k, t->ls_l0, t->ls_pos1, t->ls_pos2);
    e[MDY_FF_MS] = orc_ls_energy(x, t->ms_n, t->ms_beadset, t->ms_k, t->ms_l0, t->ms_pos1, t->ms_pos2);
    e[MDY_FF_BS] = orc_bs_energy(x, t->bs_n, t->bs_beadset, t->bs_k, t->bs_l0, t->bs_pos);
    e[MDY_FF_BB] = orc_bb_energy(x, t->bb_n, t->bb_beadset, t->bb_k, t->bb_theta0);
    e[MDY_FF_BDH] = (o->sys.br_dihedral_type == MDY_BRDIH_QUADRATIC ? orc_bdhq_energy : orc_bdh_energy)(
        x, t->bdh_n, t->bdh_beadset, t->bdh_k, t->bdh_pos);
    e[MDY_FF_BP] = orc_bp_energy(x, t->bp_n, t->bp_beadset, t->bp_k, t->bp_pos);
    if (!o->sys.exvol_enabled) e[MDY_FF_EV] = 0.0;
    else if (o->ev_mon) e[MDY_FF_EV] = orc_evm_energy(x, (int32_t)o->ev_n, o->ev_bs, o->ev_mon, o->ev_eql, o->ev_krep);
    else e[MDY_FF_EV] = orc_ev_energy(x, (int32_t)o->ev_n, o->ev_bs, o->ev_krep, o->ev_eql);
    e[MDY_FF_BD] = (o->sys.boundary_enabled && o->sys.bnd.n_planes > 0)
                       ? orc_bd_energy(x, o->sys.bnd.n_planes, o->sys.bnd.plane, o->bd_nneigh, o->bd_bs, o->bd_krep,
                                       o->bd_slen)
                       : 0.0;
}

/* ForceFieldManager::computeEnergy, ForceFieldManager.cpp:96-207: fixed-order
 * sum; an energy <= -1 from any force field makes the total +inf. */
double orc_ffm_energy(const orc_system *o, const double *x, double *per_ff) {
    double e[MDY_NFF];
    mdy_xsum X;
    mdy_xs_zero(&X);
    g_total = &X;
    ffm_each(o, x, e);
    g_total = NULL;
    if (per_ff) memcpy(per_ff, e, sizeof(e));
    double energy = 0.0;
    for (int k = 0; k < MDY_NFF; k++) {
        if (e[k] <= -1) return INFINITY;
        energy += e[k];
    }
    return g_exact ? mdy_xs_value(&X) : energy;
}
/* ForceFieldManager::computeIndividualEnergies, :209-227 (no sentinel) */
double orc_ffm_energy_individual(const orc_system *o, const double *x, double *per_ff) {
    double e[MDY_NFF];
    mdy_xsum X;
    mdy_xs_zero(&X);
    g_total = &X;
    ffm_each(o, x, e);
    g_total = NULL;
    if (per_ff) memcpy(per_ff, e, sizeof(e));
    double total = 0.0;
    for (int k = 0; k < MDY_NFF; k++) total += e[k];
    return g_exact ? mdy_xs_value(&X) : total;
}
/* ForceFieldManager::computeForces, :242-290 */
int orc_ffm_forces(const orc_system *o, const double *x, double *f, int32_t nvar) {
    const mdy_ff_tables *t = &o->sys.ff;
    for (int32_t i = 0; i < nvar; i++) f[i] = 0.0;
    orc_fs_forces(x, f, t->fs_n, t->fs_beadset, t->fs_k, t->fs_l0);
    orc_fb_forces(x, f, t->fb_n, t->fb_beadset, t->fb_k, t->fb_theta0);
    orc_ls_forces(x, f, t->ls_n, t->ls_beadset, t->ls_k, t->ls_l0, t->ls_pos1, t->ls_pos2, NULL);
    orc_ls_forces(x, f, t->ms_n, t->ms_beadset, t->ms_k, t->ms_l0, t->ms_pos1, t->ms_pos2, NULL);
    orc_bs_forces(x, f, t->bs_n, t->bs_beadset, t->bs_k, t->bs_l0, t->bs_pos, NULL);
    orc_bb_forces(x, f, t->bb_n, t->bb_beadset, t->bb_k, t->bb_theta0, NULL);
    (o->sys.br_dihedral_type == MDY_BRDIH_QUADRATIC ? orc_bdhq_forces : orc_bdh_forces)(
        x, f, t->bdh_n, t->bdh_beadset, t->bdh_k, t->bdh_pos, NULL);
    orc_bp_forces(x, f, t->bp_n, t->bp_beadset, t->bp_k, t->bp_pos, NULL);
    if (o->sys.exvol_enabled && o->ev_mon) {
        orc_evm_forces(x, f, (int32_t)o->ev_n, o->ev_bs, o->ev_mon, o->ev_eql, o->ev_krep);
    } else if (o->sys.exvol_enabled) {
        int rc = orc_ev_forces(x, f, (int32_t)o->ev_n, o->ev_bs, o->ev_krep, o->ev_eql, o->coplanar_mode);
        if (rc) return rc;
    }
    if (o->sys.boundary_enabled && o->sys.bnd.n_planes > 0)
        orc_bd_forces(x, f, o->sys.bnd.n_planes, o->sys.bnd.plane, o->bd_nneigh, o->bd_bs, o->bd_krep, o->bd_slen);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* a12: CGMethod::minimize<method> + backtrackingLineSearch                  */
/* (Minimizer/CGMethod.hpp:238-373, 826-1743; CGMethod.cpp:43-360;            */
/*  MinimizationTypes.hpp:10-64)                                             */
/* ------------------------------------------------------------------------- */
typedef struct cg_state {
    int32_t nvar, ndof;
    double *force, *forcePrev, *searchDir, *coordLS, *coordMinE;
    int haveMinE;
    double minimumE, maxForceBackup, totalEnergy;
} cg_state;

static double cg_maxF(const cg_state *s) {
    double m = 0;
    for (int32_t i = 0; i < s->ndof; i++) m = smax(m, fabs(s->force[i]));
    return m;
}
static double cg_dot(const double *a, const double *b, int32_t n) {
    osum r;
    os_zero(&r);
    for (int32_t i = 0; i < n; i++) os_add(&r, a[i] * b[i]);
    return os_val(&r);
}
static double cg_FdotF(const cg_state *s) { return cg_dot(s->force, s->force, s->nvar); }  /* CGMethod.cpp:325-331 */
static double cg_FdotFprev(const cg_state *s) { return cg_dot(s->force, s->forcePrev, s->ndof); }
static double cg_SdotF(const cg_state *s) { return cg_dot(s->searchDir, s->force, s->ndof); }
static void cg_shift(cg_state *s, double beta) {
    for (int32_t i = 0; i < s->ndof; i++) s->searchDir[i] = s->force[i] + beta * s->searchDir[i];
}

static inline int call_limit_exceeded(const mdy_min_result *r) {
    return (r->energy_call_limit != 0 && r->num_energy_call > r->energy_call_limit) ||
           (r->force_call_limit != 0 && r->num_force_call > r->force_call_limit);
}
static inline int converged(const mdy_min_result *r) {
    return r->grad_inf_norm < r->grad_tol || r->energy_rel_change < r->energy_rel_tol;
}

/* backtrackingLineSearch, CGMethod.hpp:238-373 */
static double line_search(cg_state *s, const double *x, orc_energy_fn efn, void *user, double MAXDIST,
                          double maxForce, double LAMBDAMAX, double armijo, mdy_min_result *r, double *newEnergy) {
    *newEnergy = 0;
    if (maxForce == 0.0) return 0.0;
    double lambda = smin(LAMBDAMAX, MAXDIST / maxForce);
    double currentEnergy = efn(user, x, NULL, 0);
    r->num_energy_call++;
    double expectedSlope = cg_SdotF(s);
    double reducedSlope = armijo * expectedSlope;
    int done = 0;
    double energyLambda = 0.0;
    while (!done) {
        for (int32_t i = 0; i < s->ndof; i++) s->coordLS[i] = x[i] + lambda * s->searchDir[i];
        energyLambda = efn(user, s->coordLS, NULL, 0);
        r->num_energy_call++;
        double ideal = -reducedSlope * lambda;
        if (lambda < 0 || ideal > 0) { done = 1; lambda = 0.0; }
        double energyChange = energyLambda - currentEnergy;
        if (energyChange <= ideal) done = 1;
        else lambda *= 0.5;
        if (lambda <= 0.0 || lambda <= 1e-8) {
            done = 1;
            lambda = 0.0;
            energyLambda = currentEnergy;
        }
    }
    s->totalEnergy = energyLambda;
    *newEnergy = energyLambda;
    return lambda;
}

/* quadraticoptimization, CGMethod.hpp:663-765: a parabola through three
 * (lambda, E) points, refined with energy evaluations until lambda or the
 * energy changes by less than LAMBDAQUADTOL = 1e-3. The reference loops
 * without a bound; this stops after 200 refinements with the middle point. */
static double quad_opt(cg_state *s, const double *x, orc_energy_fn efn, void *user, mdy_min_result *r, double x0,
                       double x1, double x2, double y0, double y1, double y2) {
    const double btlambda = x1;
    for (int iter = 0; iter < 200; iter++) {
        double d = (x0 - x1) * (x0 - x2) * (x1 - x2);
        double a = y0 * (x1 - x2) - y1 * (x0 - x2) + y2 * (x0 - x1);
        double b = -(y0 * (x1 + x2) * (x1 - x2) - y1 * (x0 + x2) * (x0 - x2) + y2 * (x0 + x1) * (x0 - x1));
        double lq = -b / (2 * a);
        if (fabs(lq - x1) / x1 < 1e-3) return lq;
        if (a / d < 0) return btlambda;
        for (int32_t i = 0; i < s->ndof; i++) s->coordLS[i] = x[i] + lq * s->searchDir[i];
        double eq = efn(user, s->coordLS, NULL, 0);
        r->num_energy_call++;
        if (fabs(eq - y1) / y1 < 1e-3) return lq;
        if (lq < x1) {
            if (eq < y0 && eq < y1) { x1 = lq; y1 = eq; }
            else { x0 = lq; y0 = eq; }
        } else if (x1 < lq) {
            if (eq < y1 && eq < y2) { x1 = lq; y1 = eq; }
            else { x2 = lq; y2 = eq; }
        } else {
            return x1;
        }
    }
    return x1;
}

/* quadraticLineSearchV2, CGMethod.hpp:522-658: backtracking with the Armijo
 * slope 0.4 that records (lambda, E) of every step, then, if the accepted
 * step was not the first and improved the energy while an earlier step's
 * energy was higher, a quadratic refinement through (0, E0), the accepted
 * point and the first recorded point (the search loop leaves its index at 0).
 * newEnergy is the backtracking step's energy whatever lambda is returned.
 * *fatal: the ideal energy change was positive (the reference exits). */
static double line_search_quad(cg_state *s, const double *x, orc_energy_fn efn, void *user, double MAXDIST,
                               double maxForce, double LAMBDAMAX, mdy_min_result *r, double *newEnergy, int *fatal) {
    *newEnergy = 0;
    *fatal = 0;
    if (maxForce == 0.0) return 0.0;
    double lambda = smin(LAMBDAMAX, MAXDIST / maxForce);
    double currentEnergy = efn(user, x, NULL, 0);
    r->num_energy_call++;
    double energyLambda = 0.0, prevEnergy = 0.0, prevLambda = 0.0;
    /* lambdavec / energyvec of the reference */
    int nrec = 0, caprec = 0, quadstatus = 0;
    double *lrec = NULL, *erec = NULL;
    int iter = 0, conv = 0;
    while (!conv) {
        iter++;
        for (int32_t i = 0; i < s->ndof; i++) s->coordLS[i] = x[i] + lambda * s->searchDir[i];
        energyLambda = efn(user, s->coordLS, NULL, 0);
        r->num_energy_call++;
        double ideal = -0.4 * lambda * cg_SdotF(s);
        if (ideal > 0) { *fatal = 1; free(lrec); free(erec); return 0.0; }
        double change = energyLambda - currentEnergy;
        if (change <= ideal) conv = 1;
        else { prevLambda = lambda; prevEnergy = energyLambda; lambda *= 0.5; }
        double pl, pe;
        if (lambda <= 0.0 || lambda <= 1e-8) { conv = 1; lambda = 0.0; pl = 0.0; pe = currentEnergy; }
        else if (conv) { pl = lambda; pe = energyLambda; }
        else { pl = prevLambda; pe = prevEnergy; }
        if (nrec == caprec) {
            caprec = caprec ? 2 * caprec : 16;
            lrec = (double *)realloc(lrec, sizeof(double) * (size_t)caprec);
            erec = (double *)realloc(erec, sizeof(double) * (size_t)caprec);
        }
        lrec[nrec] = pl; erec[nrec] = pe;
        nrec++;
    }
    if (lambda > 0.0 && iter > 1 && energyLambda < currentEnergy) {
        int i;
        for (i = nrec - 1; i >= 1; i--)
            if (energyLambda < erec[i]) quadstatus = 1;
        /* i == 0 here: the third point is the first record */
        if (quadstatus) {
            double lq = quad_opt(s, x, efn, user, r, 0.0, lambda, lrec[i], currentEnergy, energyLambda, erec[i]);
            lambda = (lq <= 0.0 || lq <= 1e-8) ? 0.0 : lq;
        }
    }
    free(lrec); free(erec);
    s->totalEnergy = lambda > 0 ? energyLambda : currentEnergy;
    *newEnergy = energyLambda;
    return lambda;
}

static void reset_cg(cg_state *s, int method, double *lastFDotF) {
    cg_shift(s, 0);
    if (method == MDY_CG_FLETCHER_RIEVES || method == MDY_CG_POLAK_RIBIERE) *lastFDotF = cg_FdotF(s);
    if (method == MDY_CG_POLAK_RIBIERE) memcpy(s->forcePrev, s->force, sizeof(double) * (size_t)s->nvar);
}

/* CGMethod::searchDir as the last orc_minimize left it */
static double *g_last_dir = NULL;
static int32_t g_last_dir_n = 0;
int32_t orc_last_search_dir(double *out, int32_t cap) {
    if (out) memcpy(out, g_last_dir, sizeof(double) * (size_t)(cap < g_last_dir_n ? cap : g_last_dir_n));
    return g_last_dir_n;
}

static void trace_push(orc_trace *t, double lam, double e, double mf, double beta, int safe) {
    if (!t || t->n >= t->cap) return;
    t->lambda[t->n] = lam; t->energy[t->n] = e; t->maxf[t->n] = mf; t->beta[t->n] = beta; t->safe[t->n] = safe;
    t->n++;
}

int orc_minimize(const mdy_cg_params *p, double *x, int32_t nvar, int32_t ndof, orc_energy_fn efn, orc_force_fn ffn,
                 orc_recover_fn rfn, void *user, double *out_force, mdy_min_result *r, orc_trace *trace,
                 int32_t iter_cap) {
    memset(r, 0, sizeof(*r));
    r->energy_call_limit = p->energy_call_limit;
    r->grad_tol = p->grad_tol;
    r->energy_rel_tol = p->energy_rel_tol;
    r->force_call_limit = p->force_call_limit > 3 * p->n_beads_limit ? p->force_call_limit : 3 * p->n_beads_limit;
    r->grad_inf_norm = INFINITY;
    r->energy_rel_change = INFINITY;
    if (trace) trace->n = 0;

    cg_state s;
    memset(&s, 0, sizeof(s));
    s.nvar = nvar; s.ndof = ndof;
    s.force = (double *)calloc((size_t)nvar + 1, sizeof(double));
    s.forcePrev = (double *)calloc((size_t)nvar + 1, sizeof(double));
    s.searchDir = (double *)calloc((size_t)nvar + 1, sizeof(double));
    s.coordLS = (double *)calloc((size_t)nvar + 1, sizeof(double));
    s.coordMinE = (double *)calloc((size_t)nvar + 1, sizeof(double));
    s.minimumE = INFINITY;
    int status = MDY_OK;

    r->energies_before[MDY_NFF] = efn(user, x, r->energies_before, 1);
    r->num_energy_call++;
    status = ffn(user, x, s.force, nvar);
    r->num_force_call++;
    if (status) goto done;
    r->grad_inf_norm = cg_maxF(&s);
    if (!isfinite(r->grad_inf_norm)) { status = MDY_ERR_NONFINITE; goto done; }
    double prevEnergy = r->energies_before[MDY_NFF];
    int safeMode = 0;
    int numIter = 0;
    double lastFDotF = 0;
    reset_cg(&s, p->method, &lastFDotF);

    while (!call_limit_exceeded(r) && !converged(r)) {
        if (iter_cap > 0 && numIter >= iter_cap) break;
        double beta = 0, newGrad = 0, prevGrad = 0;
        numIter++;
        double newE, lambda;
        if (p->line_search == MDY_LS_QUADRATIC && !safeMode) {
            int fatal = 0;
            lambda = line_search_quad(&s, x, efn, user, p->max_dist, r->grad_inf_norm, p->lambda_max, r, &newE,
                                      &fatal);
            if (fatal) { status = MDY_ERR_QUADRATIC; goto done; }
        } else {
            lambda = line_search(&s, x, efn, user, p->max_dist, r->grad_inf_norm, p->lambda_max,
                                 safeMode ? 0 : 0.4, r, &newE);
        }
        if (lambda == 0) {
            if (p->try_to_recover) {
                if (safeMode) {
                    if (rfn) {
                        status = rfn(user, x, s.searchDir, ndof, p->max_dist);
                        if (status) goto done;
                    }
                    r->num_recovery++;
                    status = ffn(user, x, s.force, nvar);
                    r->num_force_call++;
                    if (status) goto done;
                    r->grad_inf_norm = cg_maxF(&s);
                    if (!isfinite(r->grad_inf_norm)) { status = MDY_ERR_NONFINITE; goto done; }
                    reset_cg(&s, p->method, &lastFDotF);
                    safeMode = 0;
                } else {
                    reset_cg(&s, p->method, &lastFDotF);
                    safeMode = 1;
                    r->num_safe_mode++;
                }
                trace_push(trace, 0.0, newE, r->grad_inf_norm, 0.0, safeMode);
            } else {
                r->error_flags |= MDY_ERROR_LINESEARCH;
                r->num_iter = numIter;
                trace_push(trace, 0.0, newE, r->grad_inf_norm, 0.0, safeMode);
                status = MDY_OK;
                goto done_noafter;
            }
        } else {
            for (int32_t i = 0; i < ndof; i++) x[i] += lambda * s.searchDir[i];
            status = ffn(user, x, s.force, nvar);
            r->num_force_call++;
            if (status) goto done;
            r->grad_inf_norm = cg_maxF(&s);
            if (!isfinite(r->grad_inf_norm)) { status = MDY_ERR_NONFINITE; goto done; }
            r->energy_rel_change = fabs(newE - prevEnergy) / fabs(prevEnergy);
            prevEnergy = newE;
            if (p->method == MDY_CG_FLETCHER_RIEVES) newGrad = cg_FdotF(&s);
            else if (p->method == MDY_CG_POLAK_RIBIERE) { newGrad = cg_FdotF(&s); prevGrad = cg_FdotFprev(&s); }
            if (p->method == MDY_CG_FLETCHER_RIEVES) beta = newGrad / lastFDotF;
            else if (p->method == MDY_CG_POLAK_RIBIERE) beta = smax(0.0, (newGrad - prevGrad) / lastFDotF);
            cg_shift(&s, beta);
            if (cg_SdotF(&s) <= 0) reset_cg(&s, p->method, &lastFDotF);
            if (p->method == MDY_CG_POLAK_RIBIERE) memcpy(s.forcePrev, s.force, sizeof(double) * (size_t)nvar);
            /* forceBelowRelaxedTolerance(10) && numIter > forceCallLimit/2 -> copycoordsifminimumE */
            int relaxed = 1;
            for (int32_t i = 0; i < ndof; i++)
                if (fabs(s.force[i]) > 10 * p->grad_tol) { relaxed = 0; break; }
            if (relaxed && numIter > r->force_call_limit / 2) {
                if (s.totalEnergy <= s.minimumE) {
                    s.minimumE = s.totalEnergy;
                    s.maxForceBackup = r->grad_inf_norm;
                    memcpy(s.coordMinE, x, sizeof(double) * (size_t)nvar);
                    s.haveMinE = 1;
                }
            }
            if (p->method == MDY_CG_FLETCHER_RIEVES || p->method == MDY_CG_POLAK_RIBIERE) lastFDotF = newGrad;
            trace_push(trace, lambda, newE, r->grad_inf_norm, beta, safeMode);
        }
    }
    r->num_iter = numIter;
    if (p->report_on_call_limit && call_limit_exceeded(r)) {   /* CGMethod.hpp:1589-1608 */
        if (s.haveMinE) { memcpy(x, s.coordMinE, sizeof(double) * (size_t)nvar); s.haveMinE = 0; }
    }
    r->energies_after[MDY_NFF] = efn(user, x, r->energies_after, 1);
    r->num_energy_call++;
    if (out_force) memcpy(out_force, s.force, sizeof(double) * (size_t)nvar);
    goto done_noafter;
done:
    r->num_iter = numIter;
done_noafter:
    r->status = status;
    free(g_last_dir);
    g_last_dir = (double *)malloc(sizeof(double) * ((size_t)ndof + 1));
    memcpy(g_last_dir, s.searchDir, sizeof(double) * (size_t)ndof);
    g_last_dir_n = ndof;
    free(s.force); free(s.forcePrev); free(s.searchDir); free(s.coordLS); free(s.coordMinE);
    return status;
}

/* MController::moveAlongDirectionLimitingBeadMove_ (Controller/MController.h:273-339) */
int orc_recover_move(void *user, double *x, const double *dir, int32_t ndof, double maxdist) {
    (void)user;
    double step = INFINITY;
    for (int32_t i = 0; i < ndof; i++) step = smin(step, maxdist / fabs(dir[i]));
    if (step == INFINITY) return MDY_ERR_RECOVERY;
    for (int32_t i = 0; i < ndof; i++) x[i] += step * dir[i];
    return 0;
}

static double sys_energy(void *user, const double *x, double *per_ff, int individual) {
    const orc_system *o = (const orc_system *)user;
    return individual ? orc_ffm_energy_individual(o, x, per_ff) : orc_ffm_energy(o, x, per_ff);
}
static int sys_forces(void *user, const double *x, double *f, int32_t nvar) {
    return orc_ffm_forces((const orc_system *)user, x, f, nvar);
}

int orc_minimize_system(const orc_system *o, const mdy_cg_params *p, double *x, double *out_force,
                        mdy_min_result *res, orc_trace *trace, int32_t iter_cap) {
    int32_t nvar = 3 * o->sys.n_beads;
    return orc_minimize(p, x, nvar, nvar, sys_energy, sys_forces, orc_recover_move, (void *)o, out_force, res, trace,
                        iter_cap);
}

/* ------------------------------------------------------------------------- */
/* shared numerics, exported for their own tests                             */
/* ------------------------------------------------------------------------- */
void orc_xmath_atan(const double *x, double *y, int64_t n) {
    for (int64_t i = 0; i < n; i++) y[i] = mdy_atan(x[i]);
}
void orc_xmath_exp(const double *x, double *y, int64_t n) {
    for (int64_t i = 0; i < n; i++) y[i] = mdy_exp(x[i]);
}
void orc_xmath_acos(const double *x, double *y, int64_t n) {
    for (int64_t i = 0; i < n; i++) y[i] = mdy_acos(x[i]);
}
/* exact sum of x[0..n) added in the given order, or (stride > 1) in a
 * strided interleaving merged afterwards: the result must not depend on it */
double orc_xsum(const double *x, int64_t n, int64_t stride) {
    if (stride <= 1) {
        mdy_xsum s;
        mdy_xs_zero(&s);
        for (int64_t i = 0; i < n; i++) mdy_xs_add(&s, x[i]);
        return mdy_xs_value(&s);
    }
    mdy_xsum tot;
    mdy_xs_zero(&tot);
    for (int64_t r = stride - 1; r >= 0; r--) {
        mdy_xsum s;
        mdy_xs_zero(&s);
        for (int64_t i = r; i < n; i += stride) mdy_xs_add(&s, x[i]);
        mdy_xs_norm(&s);
        mdy_xs_merge(&tot, &s);
        mdy_xs_norm(&tot);
    }
    return mdy_xs_value(&tot);
}
