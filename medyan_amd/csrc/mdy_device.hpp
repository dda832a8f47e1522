// mdy_device.hpp -- per-interaction energy and force math for gfx950.
//
// Each function evaluates ONE interaction of one force field from bead
// positions already in registers; the kernels decide where positions come
// from (x, or x + lambda * d during a line search) and where forces go.
// Expressions keep the operation order of the reference lines they cite and
// are built with -ffp-contract=off; arctangent and exponential are the shared
// mdy_atan / mdy_exp (mdy_xmath.h) and cos/sin of equilibrium angles are
// precomputed by the host, so the FP64 results equal the CPU path's bits.
//
// The math is written once over a scalar type T: double (the FP64 path),
// float (excluded-volume derivative blocks of the mixed-precision form) and
// Dual, a forward-mode dual number whose tangent part turns a force
// evaluation into a Hessian-vector product.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>

#include "mdy_xmath.h"

namespace mdy {

template <class T>
struct V3 { T x, y, z; };
using d3 = V3<double>;
using f3 = V3<float>;   // force records of the mixed-precision form
template <class R> struct vec3_of { using T = V3<R>; };

// ---- forward-mode dual number: value v and tangent d (d/dt at t = 0) -------
struct Dual {
    double v, d;
    Dual() = default;
    __device__ explicit Dual(double x) : v(x), d(0.0) {}
    __device__ Dual(double x, double t) : v(x), d(t) {}
};
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
__device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
__device__ __forceinline__ Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
__device__ __forceinline__ Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
__device__ __forceinline__ Dual operator/(double a, Dual b) {
    const double q = a / b.v;
    return Dual(q, -q * b.d / b.v);
}
__device__ __forceinline__ bool operator<(Dual a, Dual b) { return a.v < b.v; }
__device__ __forceinline__ bool operator>(Dual a, Dual b) { return a.v > b.v; }
__device__ __forceinline__ bool operator<(Dual a, double b) { return a.v < b; }
__device__ __forceinline__ bool operator>(Dual a, double b) { return a.v > b; }
__device__ __forceinline__ bool operator<(double a, Dual b) { return a < b.v; }

// scalar functions for every T (distinct names: the C library's and the
// shared ones stay the only candidates for double and float)
__device__ __forceinline__ double xsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ float xsqrt(float a) { return sqrtf(a); }
__device__ __forceinline__ Dual xsqrt(Dual a) {
    const double s = sqrt(a.v);
    return Dual(s, a.d / (2 * s));
}
__device__ __forceinline__ double xfabs(double a) { return fabs(a); }
__device__ __forceinline__ Dual xfabs(Dual a) { return a.v < 0 ? -a : a; }
__device__ __forceinline__ double xatan(double a) { return mdy_atan(a); }
__device__ __forceinline__ Dual xatan(Dual a) { return Dual(mdy_atan(a.v), a.d / (1 + a.v * a.v)); }
__device__ __forceinline__ double xacos(double a) { return mdy_acos(a); }
// d acos(v) = -dv / sqrt(1 - v^2); a constant argument (a clamped cosine) has
// no derivative, also at v = +-1
__device__ __forceinline__ Dual xacos(Dual a) {
    return Dual(mdy_acos(a.v), a.d == 0.0 ? 0.0 : -a.d / sqrt((1 - a.v) * (1 + a.v)));
}
__device__ __forceinline__ double xexp(double a) { return mdy_exp(a); }
__device__ __forceinline__ Dual xexp(Dual a) {
    const double e = mdy_exp(a.v);
    return Dual(e, a.d * e);
}
__device__ __forceinline__ bool xfinite(double a) { return isfinite(a); }
__device__ __forceinline__ bool xfinite(float a) { return isfinite(a); }
__device__ __forceinline__ bool xfinite(Dual a) { return isfinite(a.v) && isfinite(a.d); }
template <class R, class T> __device__ __forceinline__ R cvt(T a) { return (R)a; }
template <> __device__ __forceinline__ Dual cvt<Dual, Dual>(Dual a) { return a; }

__device__ __forceinline__ d3 ld3(const double *__restrict__ p) { return d3{p[0], p[1], p[2]}; }
template <class T> __device__ __forceinline__ V3<T> sub(V3<T> a, V3<T> b) { return V3<T>{a.x - b.x, a.y - b.y, a.z - b.z}; }

// std::max / std::min argument order (NaN handling identical to the host code)
template <class T> __device__ __forceinline__ T smax(T a, T b) { return (a < b) ? b : a; }
template <class T> __device__ __forceinline__ T smin(T a, T b) { return (b < a) ? b : a; }

// (v2-v1).(v4-v3), MathFunctions.h:469-476
template <class T> __device__ __forceinline__ T sprod(V3<T> v1, V3<T> v2, V3<T> v3, V3<T> v4) {
    return ((v2.x - v1.x) * (v4.x - v3.x) + (v2.y - v1.y) * (v4.y - v3.y) + (v2.z - v1.z) * (v4.z - v3.z));
}
template <class T> __device__ __forceinline__ T tpdist(V3<T> v1, V3<T> v2) {
    return xsqrt((v2.x - v1.x) * (v2.x - v1.x) + (v2.y - v1.y) * (v2.y - v1.y) + (v2.z - v1.z) * (v2.z - v1.z));
}
template <class T> __device__ __forceinline__ V3<T> midpt(V3<T> v1, V3<T> v2, double alpha) {
    double beta = 1 - alpha;
    return V3<T>{(v1.x * beta + alpha * v2.x), (v1.y * beta + alpha * v2.y), (v1.z * beta + alpha * v2.z)};
}
template <class T> __device__ __forceinline__ V3<T> cross3(V3<T> v1, V3<T> v2) {
    return V3<T>{v1.y * v2.z - v1.z * v2.y, v1.z * v2.x - v1.x * v2.z, v1.x * v2.y - v1.y * v2.x};
}
template <class T> __device__ __forceinline__ V3<T> normalize3(V3<T> v) {
    T x2 = v.x * v.x, y2 = v.y * v.y, z2 = v.z * v.z;
    T n = xsqrt(x2 + y2 + z2);
    return V3<T>{v.x / n, v.y / n, v.z / n};
}
template <class T> __device__ __forceinline__ T dot3(V3<T> a, V3<T> b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ bool are_equal0(double a) { return fabs(a - 0.0) < 1e-6; }

// ---------------------------------------------------------------------------
// a3 FilamentStretchingHarmonic (FilamentStretchingHarmonic.cpp:185-275)
// ---------------------------------------------------------------------------
template <class T> __device__ __forceinline__ T fs_energy(V3<T> c1, V3<T> c2, double k, double l0) {
    T dist = tpdist(c1, c2) - l0;
    return 0.5 * k * dist * dist;
}
// forces on c1, c2
template <class T>
__device__ __forceinline__ void fs_force(V3<T> c1, V3<T> c2, double k, double l0, V3<T> &f1, V3<T> &f2) {
    T dist = tpdist(c1, c2);
    T invL = 1 / dist;
    T f0 = k * (dist - l0) * invL;
    T Fx = -f0 * (c2.x - c1.x), Fy = -f0 * (c2.y - c1.y), Fz = -f0 * (c2.z - c1.z);
    f2 = V3<T>{Fx, Fy, Fz};
    f1 = V3<T>{-Fx, -Fy, -Fz};
}

// ---------------------------------------------------------------------------
// a4 / a8 cosine bending (FilamentBendingCosine.cpp:199-398, BranchingBendingCosine.cpp:196-411)
// ---------------------------------------------------------------------------
// cos(theta0) and sin(theta0) arrive precomputed by the host C library (the
// same values the CPU path uses): ct = cos(th0), st = sin(th0).
template <class T>
__device__ __forceinline__ T cosbend_energy(T l1l2, T L1, T L2, double k, double th0, double ct, double st) {
    T L1L2 = L1 * L2;
    T x = l1l2 / L1L2;
    if (x < -1.0) x = T(-1.0);
    else if (x > 1.0) x = T(1.0);
    if (are_equal0(th0)) return k * (1 - x);
    T cosA = x;
    T sinA = smax(xsqrt(1 - cosA * cosA), T(0.0));
    T cosAmB = cosA * ct + sinA * st;
    return k * (1 - cosAmB);
}
template <class T>
__device__ __forceinline__ void cosbend_coef(T L1sq, T L2sq, T l1l2, double kb, double th0, double ct, double st,
                                             T &A, T &B, T &C, T &ke) {
    A = 1 / xsqrt(L1sq * L2sq);
    T x = l1l2 * A;
    B = x / L1sq;
    C = x / L2sq;
    if (are_equal0(th0)) { ke = T(kb); return; }
    if (xfabs(xfabs(x) - 1.0) < 0.001) x = 0.999 * x;
    if (x < -1.0) x = T(-1.0);
    else if (x > 1.0) x = T(1.0);
    T cosp = x;
    T sinp = xsqrt(smax((1 - cosp * cosp), T(0.0)));
    T sinpmq = sinp * ct - cosp * st;
    ke = kb * sinpmq / sinp;
}
// filament: beads c1, c2, c3; r1 = c2-c1, r2 = c3-c2
template <class T>
__device__ __forceinline__ T fb_energy(V3<T> c1, V3<T> c2, V3<T> c3, double k, double th0, double ct, double st) {
    T L1 = xsqrt(sprod(c1, c2, c1, c2));
    T L2 = xsqrt(sprod(c2, c3, c2, c3));
    T l1l2 = sprod(c1, c2, c2, c3);
    return cosbend_energy(l1l2, L1, L2, k, th0, ct, st);
}
template <class T>
__device__ __forceinline__ void fb_force(V3<T> c1, V3<T> c2, V3<T> c3, double k, double th0, double ct, double st,
                                         V3<T> &f1, V3<T> &f2, V3<T> &f3) {
    T A, B, C, ke;
    cosbend_coef(sprod(c1, c2, c1, c2), sprod(c2, c3, c2, c3), sprod(c1, c2, c2, c3), k, th0, ct, st, A, B, C, ke);
    V3<T> r1 = sub(c2, c1), r2 = sub(c3, c2);
    V3<T> F1 = V3<T>{ke * (r2.x * A - r1.x * B), ke * (r2.y * A - r1.y * B), ke * (r2.z * A - r1.z * B)};
    V3<T> F2 = V3<T>{ke * (r1.x * A - r2.x * C), ke * (r1.y * A - r2.y * C), ke * (r1.z * A - r2.z * C)};
    f1 = V3<T>{-F1.x, -F1.y, -F1.z};
    f2 = V3<T>{F1.x - F2.x, F1.y - F2.y, F1.z - F2.z};
    f3 = F2;
}
// branch: mother (c1, c2), daughter (c3, c4)
template <class T>
__device__ __forceinline__ T bb_energy(V3<T> c1, V3<T> c2, V3<T> c3, V3<T> c4, double k, double th0, double ct,
                                       double st) {
    T L1 = xsqrt(sprod(c1, c2, c1, c2));
    T L2 = xsqrt(sprod(c3, c4, c3, c4));
    T l1l2 = sprod(c1, c2, c3, c4);
    return cosbend_energy(l1l2, L1, L2, k, th0, ct, st);
}
template <class T>
__device__ __forceinline__ void bb_force(V3<T> c1, V3<T> c2, V3<T> c3, V3<T> c4, double k, double th0, double ct,
                                         double st, V3<T> &f1, V3<T> &f2, V3<T> &f3, V3<T> &f4) {
    T A, B, C, ke;
    cosbend_coef(sprod(c1, c2, c1, c2), sprod(c3, c4, c3, c4), sprod(c1, c2, c3, c4), k, th0, ct, st, A, B, C, ke);
    V3<T> r1 = sub(c2, c1), r2 = sub(c4, c3);
    V3<T> F1 = V3<T>{ke * (r2.x * A - r1.x * B), ke * (r2.y * A - r1.y * B), ke * (r2.z * A - r1.z * B)};
    V3<T> F2 = V3<T>{ke * (r1.x * A - r2.x * C), ke * (r1.y * A - r2.y * C), ke * (r1.z * A - r2.z * C)};
    f1 = V3<T>{-F1.x, -F1.y, -F1.z};
    f2 = F1;
    f3 = V3<T>{-F2.x, -F2.y, -F2.z};
    f4 = F2;
}

// ---------------------------------------------------------------------------
// a5 / a6 linker and motor springs (LinkerStretchingHarmonic.cpp:191-336)
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T ls_energy(V3<T> c1, V3<T> c2, V3<T> c3, V3<T> c4, double k, double l0, double p1,
                                       double p2) {
    V3<T> v1 = midpt(c1, c2, p1), v2 = midpt(c3, c4, p2);
    T dist = tpdist(v1, v2) - l0;
    return 0.5 * k * dist * dist;
}
template <class T>
__device__ __forceinline__ void ls_force(V3<T> c1, V3<T> c2, V3<T> c3, V3<T> c4, double k, double l0, double p1,
                                         double p2, V3<T> &f1, V3<T> &f2, V3<T> &f3, V3<T> &f4) {
    V3<T> v1 = midpt(c1, c2, p1), v2 = midpt(c3, c4, p2);
    T dist = tpdist(v1, v2);
    T invL = 1 / dist;
    T f0 = k * (dist - l0) * invL;
    V3<T> Fl = V3<T>{-f0 * (v2.x - v1.x), -f0 * (v2.y - v1.y), -f0 * (v2.z - v1.z)};
    f1 = V3<T>{-Fl.x * (1 - p1), -Fl.y * (1 - p1), -Fl.z * (1 - p1)};
    f2 = V3<T>{-Fl.x * (p1), -Fl.y * (p1), -Fl.z * (p1)};
    f3 = V3<T>{Fl.x * (1 - p2), Fl.y * (1 - p2), Fl.z * (1 - p2)};
    f4 = V3<T>{Fl.x * (p2), Fl.y * (p2), Fl.z * (p2)};
}

// ---------------------------------------------------------------------------
// a7 BranchingStretchingHarmonic (BranchingStretchingHarmonic.cpp:195-318)
// ---------------------------------------------------------------------------
template <class T>
__device__ __forceinline__ T bs_energy(V3<T> c1, V3<T> c2, V3<T> c3, double k, double l0, double p) {
    V3<T> v1 = midpt(c1, c2, p);
    T dist = tpdist(v1, c3) - l0;
    return 0.5 * k * dist * dist;
}
template <class T>
__device__ __forceinline__ void bs_force(V3<T> c1, V3<T> c2, V3<T> c3, double k, double l0, double p, V3<T> &f1,
                                         V3<T> &f2, V3<T> &f3) {
    V3<T> v1 = midpt(c1, c2, p);
    T dist = tpdist(v1, c3);
    T invL = 1 / dist;
    T f0 = k * (dist - l0) * invL;
    V3<T> r = sub(c3, v1);
    f1 = V3<T>{-f0 * (r.x) * (p - 1), -f0 * (r.y) * (p - 1), -f0 * (r.z) * (p - 1)};
    f2 = V3<T>{f0 * (r.x) * p, f0 * (r.y) * p, f0 * (r.z) * p};
    f3 = V3<T>{-f0 * (r.x), -f0 * (r.y), -f0 * (r.z)};
}

// ---------------------------------------------------------------------------
// Branching position and dihedral, cosine forms. A branch point at the
// mother's plus end (pos == 1 within ZERO_PREC) is placed at 0.5 of a
// virtual cylinder whose plus end c2' = (1/0.5) (c2 - 0.5 c1) extends the
// mother, and the forces are mapped back to c1, c2 afterwards.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool are_equal1(double a) { return fabs(a - 1.0) < 1e-6; }
template <class T> __device__ __forceinline__ V3<T> branch_plus_end(V3<T> c1, V3<T> c2, double position) {
    return V3<T>{(1 / position) * (c2.x - (1 - position) * c1.x), (1 / position) * (c2.y - (1 - position) * c1.y),
                 (1 / position) * (c2.z - (1 - position) * c1.z)};
}
// BranchingPositionCosine::energy (BranchingPositionCosine.cpp:190-248);
// cq, sq: cos and sin of pi/2 from the host C library
template <class T>
__device__ __forceinline__ T bp_energy(V3<T> c1, V3<T> c2, V3<T> c3, double kpos, double pos, double cq, double sq) {
    double position = pos;
    if (are_equal1(position)) {
        position = 0.5;
        c2 = branch_plus_end(c1, c2, position);
    }
    const V3<T> mp = midpt(c1, c2, position);
    const T X = xsqrt(sprod(mp, c2, mp, c2));
    const T D = xsqrt(sprod(mp, c3, mp, c3));
    const T XD = X * D;
    T xd = sprod(mp, c2, mp, c3);
    T x = xd / XD;
    if (xfabs(xfabs(x) - 1.0) < 0.001) {
        xd = 0.999 * XD;
        x = xd / XD;
    }
    if (x < -1.0) x = T(-1.0);
    else if (x > 1.0) x = T(1.0);
    const T cosp = x;
    const T sinp = xsqrt(smax(T(1 - cosp * cosp), T(0.0)));
    const T cospminusq = cosp * cq + sinp * sq;
    return kpos * (1 - cospminusq);
}
// BranchingPositionCosine::forces (BranchingPositionCosine.cpp:306-531):
// forces on c1, c2, c3
template <class T>
__device__ __forceinline__ void bp_force(V3<T> c1, V3<T> c2, V3<T> c3, double kpos, double pos, double cq, double sq,
                                         V3<T> &f1, V3<T> &f2, V3<T> &f3) {
    double position = pos;
    if (are_equal1(position)) {
        position = 0.5;
        c2 = branch_plus_end(c1, c2, position);
    }
    const V3<T> mp = midpt(c1, c2, position);
    const T Xsq = sprod(mp, c2, mp, c2);
    const T Dsq = sprod(mp, c3, mp, c3);
    const T xd = sprod(mp, c2, mp, c3);
    const T A = 1 / xsqrt(Xsq * Dsq);
    T x = xd * A;
    const T B = x / Xsq;
    const T C = x / Dsq;
    if (xfabs(xfabs(x) - 1.0) < 0.001) x = 0.999 * x;
    if (x < -1.0) x = T(-1.0);
    else if (x > 1.0) x = T(1.0);
    const T cosp = x;
    const T sinp = xsqrt(smax(T(1 - cosp * cosp), T(0.0)));
    const T sinpminusq = sinp * cq - cosp * sq;
    const T k = kpos * sinpminusq / sinp;
    const V3<T> r1 = sub(c2, mp), r2 = sub(c3, mp);
    const V3<T> Fr1{k * (r2.x * A - r1.x * B), k * (r2.y * A - r1.y * B), k * (r2.z * A - r1.z * B)};
    const V3<T> Fr2{k * (r1.x * A - r2.x * C), k * (r1.y * A - r2.y * C), k * (r1.z * A - r2.z * C)};
    const V3<T> fmp{-Fr1.x - Fr2.x, -Fr1.y - Fr2.y, -Fr1.z - Fr2.z};
    const V3<T> f1t{(1 - position) * fmp.x, (1 - position) * fmp.y, (1 - position) * fmp.z};
    const V3<T> f2t{Fr1.x + position * fmp.x, Fr1.y + position * fmp.y, Fr1.z + position * fmp.z};
    f3 = Fr2;
    if (are_equal1(pos)) {
        const double factor = (position - 1) / position;
        f1 = V3<T>{f1t.x + f2t.x * factor, f1t.y + f2t.y * factor, f1t.z + f2t.z * factor};
        f2 = V3<T>{f2t.x * (1 / position), f2t.y * (1 / position), f2t.z * (1 / position)};
    } else {
        f1 = f1t;
        f2 = f2t;
    }
}
// BranchingDihedralCosine::energy (BranchingDihedralCosine.cpp:211-269)
template <class T>
__device__ __forceinline__ T bdh_energy(V3<T> c1, V3<T> c2, V3<T> c3, V3<T> c4, double kdih, double pos) {
    double position = pos;
    if (are_equal1(position)) {
        position = 0.5;
        c2 = branch_plus_end(c1, c2, position);
    }
    const V3<T> mp = midpt(c1, c2, position);
    const V3<T> n1 = normalize3(cross3(sub(c2, mp), sub(c3, mp)));
    const V3<T> n2 = normalize3(cross3(sub(c4, c3), sub(c3, mp)));
    const T n1n2 = n1.x * n2.x + n1.y * n2.y + n1.z * n2.z;
    return kdih * (1 - n1n2);
}
// BranchingDihedralCosine::forces (BranchingDihedralCosine.cpp:271-661),
// the LAMMPS dihedral form: forces on c1..c4
template <class T>
__device__ __forceinline__ void bdh_force(V3<T> c1, V3<T> c2, V3<T> c3, V3<T> c4, double kdih, double pos, V3<T> &f1,
                                          V3<T> &f2, V3<T> &f3, V3<T> &f4) {
    double position = pos;
    if (are_equal1(position)) {
        position = 0.5;
        c2 = branch_plus_end(c1, c2, position);
    }
    const V3<T> mp = midpt(c1, c2, position);
    const V3<T> vb1 = sub(c2, mp), vb2 = sub(c3, mp);
    const V3<T> vb2m{-vb2.x, -vb2.y, -vb2.z};
    const V3<T> vb3 = sub(c4, c3);
    const T ax = vb1.y * vb2m.z - vb1.z * vb2m.y, ay = vb1.z * vb2m.x - vb1.x * vb2m.z,
            az = vb1.x * vb2m.y - vb1.y * vb2m.x;
    const T bx = vb3.y * vb2m.z - vb3.z * vb2m.y, by = vb3.z * vb2m.x - vb3.x * vb2m.z,
            bz = vb3.x * vb2m.y - vb3.y * vb2m.x;
    const T rasq = ax * ax + ay * ay + az * az;
    const T rbsq = bx * bx + by * by + bz * bz;
    const T rgsq = vb2m.x * vb2m.x + vb2m.y * vb2m.y + vb2m.z * vb2m.z;
    const T rg = xsqrt(rgsq);
    T rginv = T(0.0), ra2inv = T(0.0), rb2inv = T(0.0);
    if (rg > 0) rginv = 1.0 / rg;
    if (rasq > 0) ra2inv = 1.0 / rasq;
    if (rbsq > 0) rb2inv = 1.0 / rbsq;
    const T rabinv = xsqrt(ra2inv * rb2inv);
    const T s = rg * rabinv * (ax * vb3.x + ay * vb3.y + az * vb3.z);
    const T fg = vb1.x * vb2m.x + vb1.y * vb2m.y + vb1.z * vb2m.z;
    const T hg = vb3.x * vb2m.x + vb3.y * vb2m.y + vb3.z * vb2m.z;
    const T fga = fg * ra2inv * rginv;
    const T hgb = hg * rb2inv * rginv;
    const T gaa = -ra2inv * rg;
    const T gbb = rb2inv * rg;
    const V3<T> dtf{gaa * ax, gaa * ay, gaa * az};
    const V3<T> dtg{fga * ax - hgb * bx, fga * ay - hgb * by, fga * az - hgb * bz};
    const V3<T> dth{gbb * bx, gbb * by, gbb * bz};
    const T df = -kdih * s;
    const V3<T> s2{df * dtg.x, df * dtg.y, df * dtg.z};
    f4 = V3<T>{df * dth.x, df * dth.y, df * dth.z};
    f3 = V3<T>{-s2.x - f4.x, -s2.y - f4.y, -s2.z - f4.z};
    if (are_equal1(pos)) {
        const V3<T> l2p{df * dtf.x, df * dtf.y, df * dtf.z};
        V3<T> l2{s2.x - l2p.x, s2.y - l2p.y, s2.z - l2p.z};
        const double factor = (position - 1) / position;
        f1 = V3<T>{l2p.x * factor, l2p.y * factor, l2p.z * factor};
        const double factor2 = (1 / position);
        f2 = V3<T>{l2.x + l2p.x * factor2, l2.y + l2p.y * factor2, l2.z + l2p.z * factor2};
    } else {
        const V3<T> l2{df * dtf.x, df * dtf.y, df * dtf.z};
        const V3<T> lmp{s2.x - l2.x, s2.y - l2.y, s2.z - l2.z};
        f1 = V3<T>{(1 - position) * lmp.x, (1 - position) * lmp.y, (1 - position) * lmp.z};
        f2 = V3<T>{l2.x + (position)*lmp.x, l2.y + (position)*lmp.y, l2.z + (position)*lmp.z};
    }
}

// BranchingDihedralQuadratic (BranchingDihedralQuadratic.cpp:26-86, 88-398):
// E = kdih theta^2 with theta the dihedral between the planes (c2, mp, c3)
// and (mp, c3, c4). The energy takes cos(theta) from the plane normals
// b1 x b2 and b3 x b2; the forces take it from the bond angles (sines held
// at >= 1e-3) and differentiate E through it: dE = eFac d(cos theta),
// eFac = -2 kdih theta / max(sin theta, 1e-5). sin(theta) is formed as
// sqrt((1 - c)(1 + c)), c = cos(theta) clamped to [-1, 1].
template <class T>
__device__ __forceinline__ T bdhq_energy(V3<T> c1, V3<T> c2, V3<T> c3, V3<T> c4, double kdih, double pos) {
    double position = pos;
    if (are_equal1(position)) {
        position = 0.5;
        c2 = branch_plus_end(c1, c2, position);
    }
    const V3<T> mp = midpt(c1, c2, position);
    const V3<T> b1 = sub(c2, mp), b2 = sub(c3, mp), b3 = sub(c4, c3);
    const V3<T> n1 = cross3(b1, b2), n2 = cross3(b3, b2);
    const T n1mag = xsqrt(n1.x * n1.x + n1.y * n1.y + n1.z * n1.z);
    const T n2mag = xsqrt(n2.x * n2.x + n2.y * n2.y + n2.z * n2.z);
    T ct = (n1.x * n2.x + n1.y * n2.y + n1.z * n2.z) / (n1mag * n2mag);
    if (ct < -1.0) ct = T(-1.0);
    if (ct > 1.0) ct = T(1.0);
    const T th = xacos(ct);
    return kdih * th * th;
}
template <class T>
__device__ __forceinline__ void bdhq_force(V3<T> c1, V3<T> c2, V3<T> c3, V3<T> c4, double kdih, double pos, V3<T> &f1,
                                           V3<T> &f2, V3<T> &f3, V3<T> &f4) {
    double p = pos;
    if (are_equal1(p)) {
        p = 0.5;
        c2 = branch_plus_end(c1, c2, p);
    }
    const V3<T> mp = midpt(c1, c2, p);
    const V3<T> b1 = sub(c2, mp), b2 = sub(c3, mp), b3 = sub(c4, c3);
    const T l1sq = b1.x * b1.x + b1.y * b1.y + b1.z * b1.z;
    const T l2sq = b2.x * b2.x + b2.y * b2.y + b2.z * b2.z;
    const T l3sq = b3.x * b3.x + b3.y * b3.y + b3.z * b3.z;
    const T l1 = xsqrt(l1sq), l2 = xsqrt(l2sq), l3 = xsqrt(l3sq);
    const T i1sq = 1.0 / l1sq, i2sq = 1.0 / l2sq, i3sq = 1.0 / l3sq;
    const T i13 = 1.0 / (l1 * l3), i12 = 1.0 / (l1 * l2), i32 = 1.0 / (l3 * l2);
    const T d13 = b1.x * b3.x + b1.y * b3.y + b1.z * b3.z;
    const T d12 = b1.x * b2.x + b1.y * b2.y + b1.z * b2.z;
    const T d32 = b3.x * b2.x + b3.y * b2.y + b3.z * b2.z;
    // bond angles <b1, b2> and <b3, b2>
    const T c12 = d12 * i12;
    const T s12 = smax(xsqrt(smax(T(1 - c12 * c12), T(0.0))), T(0.001));
    const T is12 = 1.0 / s12;
    const T is12sq = is12 * is12;
    const T c32 = d32 * i32;
    const T s32 = smax(xsqrt(smax(T(1 - c32 * c32), T(0.0))), T(0.001));
    const T is32 = 1.0 / s32;
    const T is32sq = is32 * is32;
    const T c13 = d13 * i13;
    // cos(theta) = (cos<b1,b3> - cos<b1,b2> cos<b3,b2>) / (sin<b1,b2> sin<b3,b2>)
    const T num = c13 - c12 * c32;
    const T iden = is12 * is32;
    T ct = num * iden;
    if (ct < -1.0) ct = T(-1.0);
    if (ct > 1.0) ct = T(1.0);
    const T th = xacos(ct);
    const T st = smax(xsqrt((1 - ct) * (1 + ct)), T(1e-5));
    const T eFac = -2 * kdih * th * (1.0 / st);
    // dE/db_i = sum_j e_ij b_j (e symmetric)
    const T e11 = -eFac * ct * is12sq * i1sq;
    const T e12 = eFac * i12 * (-c32 * iden + c12 * ct * is12sq);
    const T e13 = eFac * i13 * iden;
    const T e22 = eFac * i2sq * (2 * c13 * iden - ct * (is12sq + is32sq));
    const T e23 = eFac * i32 * (-c12 * iden + c32 * ct * is32sq);
    const T e33 = -eFac * ct * is32sq * i3sq;
    // b1 = c2 - mp, b2 = c3 - mp, b3 = c4 - c3, mp = (1-p) c1 + p c2
    const T a1 = (1 - p) * (e11 + e12), a2 = (1 - p) * (e12 + e22), a3 = (1 - p) * (e13 + e23);
    const T g1 = (p - 1) * e11 + p * e12, g2 = (p - 1) * e12 + p * e22, g3 = (p - 1) * e13 + p * e23;
    const T h1 = -e12 + e13, h2 = -e22 + e23, h3 = -e23 + e33;
    auto comb = [&](T u, T v, T w) {
        return V3<T>{u * b1.x + v * b2.x + w * b3.x, u * b1.y + v * b2.y + w * b3.y, u * b1.z + v * b2.z + w * b3.z};
    };
    const V3<T> fa = comb(a1, a2, a3), fg = comb(g1, g2, g3);
    f3 = comb(h1, h2, h3);
    f4 = comb(-e13, -e23, -e33);
    if (are_equal1(pos)) {   // the energy is a function of (c1, c2', c3, c4), c2' the extended plus end
        const double factor1 = (p - 1) / p, factor2 = 1 / p;
        f1 = V3<T>{fa.x + fg.x * factor1, fa.y + fg.y * factor1, fa.z + fg.z * factor1};
        f2 = V3<T>{fg.x * factor2, fg.y * factor2, fg.z * factor2};
    } else {
        f1 = fa;
        f2 = fg;
    }
}

// ---------------------------------------------------------------------------
// a10 BoundaryCylinderRepulsionExp (BoundaryCylinderRepulsionExp.cpp:27-110)
// plane = (a, b, c, d); distance per BoundaryElementImpl.h:71-74
// ---------------------------------------------------------------------------
template <class T> __device__ __forceinline__ T plane_dist(const double *__restrict__ pl, V3<T> p) {
    double a = pl[0], b = pl[1], c = pl[2], d = pl[3];
    return (a * p.x + b * p.y + c * p.z + d) / sqrt(a * a + b * b + c * c);
}
template <class T>
__device__ __forceinline__ T bd_energy(const double *__restrict__ pl, V3<T> p, double k, double sl) {
    T r = plane_dist(pl, p);
    T R = -r / sl;
    return k * xexp(R);
}
template <class T>
__device__ __forceinline__ V3<T> bd_force(const double *__restrict__ pl, V3<T> p, double k, double sl) {
    T r = plane_dist(pl, p);
    T R = -r / sl;
    T f0 = k * xexp(R) / sl;
    return V3<T>{f0 * pl[0], f0 * pl[1], f0 * pl[2]};
}

// ---------------------------------------------------------------------------
// a9 CylinderExclVolRepulsion (CylinderExclVolRepulsion.cpp:329-882)
// ---------------------------------------------------------------------------
// areInPlane, MathFunctions.h:1101-1133
__device__ __forceinline__ bool are_in_plane(d3 p1, d3 p2, d3 p3, d3 p4) {
    d3 v1 = sub(p2, p1), v2 = sub(p3, p1), v3 = sub(p4, p1);
    d3 cp = normalize3(cross3(v1, v2));
    v3 = normalize3(v3);
    return fabs(dot3(v3, cp) - 0.0) < 1e-1;
}
// movePointOutOfPlane for bead i in {1,2}, MathFunctions.h:1221-1270
__device__ __forceinline__ d3 moved_off_plane(d3 p1, d3 p2, d3 p3, d3 src, double d) {
    d3 n = normalize3(cross3(sub(p2, p1), sub(p3, p1)));
    return d3{(src.x + n.x * d), (src.y + n.y * d), (src.z + n.z * d)};
}

template <class T>
struct EvInvT {
    T a, b, c, d, e, F;
    T AA, BB, CC, DD, EE, FF, GG, HH, JJ;
    T ATG1, ATG2, ATG3, ATG4;
};
using EvInv = EvInvT<double>;

template <class T>
__device__ __forceinline__ void ev_invariants(V3<T> c1, V3<T> c2, V3<T> c3, V3<T> c4, EvInvT<T> &v) {
    v.a = sprod(c1, c2, c1, c2);
    v.b = sprod(c3, c4, c3, c4);
    v.c = sprod(c3, c1, c3, c1);
    v.d = sprod(c1, c2, c3, c4);
    v.e = sprod(c1, c2, c3, c1);
    v.F = sprod(c3, c4, c3, c1);
    T g = sprod(c1, c2, c1, c4);
    T h = sprod(c3, c4, c3, c2);
    T I = sprod(c3, c4, c1, c4);
    T sqD, sqE;
    {
        T x = c1.x - c4.x, y = c1.y - c4.y, z = c1.z - c4.z;
        sqD = (x * x + y * y + z * z);
        x = c3.x - c2.x; y = c3.y - c2.y; z = c3.z - c2.z;
        sqE = (x * x + y * y + z * z);
    }
    const T a = v.a, b = v.b, c = v.c, d = v.d, e = v.e, F = v.F;
    T ac = xsqrt(a * c), bc = xsqrt(b * c), aD = xsqrt(a * sqD), bE = xsqrt(b * sqE);
    v.AA = xsqrt(smax(T(0.0), (ac + e) * (ac - e)));
    v.BB = xsqrt(smax(T(0.0), (bc + F) * (bc - F)));
    v.CC = d * e - a * F;
    v.DD = b * e - d * F;
    v.EE = xsqrt(smax(T(0.0), (aD + g) * (aD - g)));
    v.FF = xsqrt(smax(T(0.0), (bE + h) * (bE - h)));
    v.GG = d * g - a * I;
    v.HH = v.CC + v.GG - v.DD;
    V3<T> cp = cross3(sub(c1, c2), sub(c3, c4));
    T JJ = dot3(cp, sub(c3, c1));
    v.JJ = -JJ * JJ;
    v.ATG1 = xatan((a + e) / v.AA) - xatan(e / v.AA);
    v.ATG2 = xatan((a + e - d) / v.EE) - xatan((e - d) / v.EE);
    v.ATG3 = xatan((F) / v.BB) - xatan((F - b) / v.BB);
    v.ATG4 = xatan((d + F) / v.FF) - xatan((d + F - b) / v.FF);
}

template <class T> __device__ __forceinline__ T ev_energy_inv(const EvInvT<T> &v, double kr) {
    return 0.5 * kr / v.JJ * (v.CC / v.AA * v.ATG1 + v.GG / v.EE * v.ATG2 + v.DD / v.BB * v.ATG3 + v.HH / v.FF * v.ATG4);
}
__device__ __forceinline__ double ev_energy(d3 c1, d3 c2, d3 c3, d3 c4, double kr) {
    EvInv v;
    ev_invariants(c1, c2, c3, c4, v);
    return ev_energy_inv(v, kr);
}
__device__ __forceinline__ bool ev_energy_bad(double U) { return fabs(U) == INFINITY || U != U || U < -1.0; }

// Analytic pair force; f[0..3] = forces on c1..c4. Direction vectors use the
// c2 passed in (the moved one in the coplanar case), as the reference does.
// R = T: the FP64 path (T = double) and the Hessian-vector product (T = Dual);
// T = double, R = float: the mixed-precision form, whose derivative blocks run
// in FP32 from the FP64 invariants.
template <class R, class T>
__device__ __forceinline__ void ev_force_inv_t(const EvInvT<T> &v, V3<T> c1, V3<T> c2, V3<T> c3, V3<T> c4,
                                               double kr_, V3<R> f[4]) {
    const R a = cvt<R>(v.a), b = cvt<R>(v.b), c = cvt<R>(v.c), d = cvt<R>(v.d), e = cvt<R>(v.e), F = cvt<R>(v.F);
    const R AA = cvt<R>(v.AA), BB = cvt<R>(v.BB), CC = cvt<R>(v.CC), DD = cvt<R>(v.DD), EE = cvt<R>(v.EE),
            FF = cvt<R>(v.FF), GG = cvt<R>(v.GG), HH = cvt<R>(v.HH);
    const R ATG1 = cvt<R>(v.ATG1), ATG2 = cvt<R>(v.ATG2), ATG3 = cvt<R>(v.ATG3), ATG4 = cvt<R>(v.ATG4);
    const R kr = R(kr_);
    const R invJJ = 1 / cvt<R>(v.JJ);
    const R sumBlock = CC / AA * ATG1 + GG / EE * ATG2 + DD / BB * ATG3 + HH / FF * ATG4;
    const R enFactor = kr * invJJ / 2;
    const R A1 = AA * AA / (AA * AA + (a + e) * (a + e));
    const R A2 = AA * AA / (AA * AA + e * e);
    const R E1 = EE * EE / (EE * EE + (a + e - d) * (a + e - d));
    const R E2 = EE * EE / (EE * EE + (e - d) * (e - d));
    const R B1 = BB * BB / (BB * BB + F * F);
    const R B2 = BB * BB / (BB * BB + (F - b) * (F - b));
    const R F1 = FF * FF / (FF * FF + (d + F) * (d + F));
    const R F2 = FF * FF / (FF * FF + (d + F - b) * (d + F - b));
    const R pA_C = ATG1 / AA;
    const R pA_A = -(ATG1 * CC) / (AA * AA) + (-A1 * (a + e) + A2 * e) * CC / (AA * AA * AA);
    const R pA_e = ((A1 - A2) * CC) / (AA * AA);
    const R pA_a = (A1 * CC) / (AA * AA);
    const R pE_G = ATG2 / EE;
    const R pE_E = -(ATG2 * GG) / (EE * EE) + (-E1 * (a + e - d) + E2 * (e - d)) * GG / (EE * EE * EE);
    const R pE_emd = ((E1 - E2) * GG) / (EE * EE);
    const R pE_a = (E1 * GG) / (EE * EE);
    const R pB_D = ATG3 / BB;
    const R pB_B = -(ATG3 * DD) / (BB * BB) + (-B1 * F + B2 * (F - b)) * DD / (BB * BB * BB);
    const R pB_fmb = ((B1 - B2) * DD) / (BB * BB);
    const R pB_b = (B1 * DD) / (BB * BB);
    const R pF_H = ATG4 / FF;
    const R pF_F = -(ATG4 * HH) / (FF * FF) + (-F1 * (d + F) + F2 * (d + F - b)) * HH / (FF * FF * FF);
    const R pF_dfb = ((F1 - F2) * HH) / (FF * FF);
    const R pF_b = (F1 * HH) / (FF * FF);
    const R k11 = pA_C * (-2 * F) + pA_A * (c / AA) + pA_a * 2 + pE_G * (2 * (F - b)) +
                  pE_E * ((b + c - 2 * F) / EE) + pE_a * 2 + pF_H * (-2 * b) + pF_F * (b / FF);
    const R k12 = pA_C * e + pE_G * (2 * d - e) + pE_E * ((e - d) / EE) + pE_emd * (-1) + pB_D * (-F) +
                  pF_H * (2 * d + F) + pF_F * (-(d + F) / FF) + pF_dfb;
    const R k13 = pA_C * d + pA_A * (-e / AA) + pA_e + pE_G * (-d) + pE_E * (-(e - d) / EE) + pE_emd +
                  pB_D * b + pF_H * (-b) + pF_F * (b / FF);
    const R k22 = pE_G * (-2 * a) + pE_E * (a / EE) + pB_D * (2 * e) + pB_B * (c / BB) + pB_fmb * (-2) +
                  pB_b * 2 + pF_H * (-2 * (a + e)) + pF_F * ((a + c + 2 * e) / FF) + pF_dfb * (-2) + pF_b * 2;
    const R k23 = pA_C * (-a) + pE_G * a + pE_E * (-a / EE) + pB_D * (-d) + pB_B * (-F / BB) + pB_fmb +
                  pF_H * d + pF_F * (-(d + F) / FF) + pF_dfb;
    const R k33 = pA_A * (a / AA) + pE_E * (a / EE) + pB_B * (b / BB) + pF_F * (b / FF);
    const R k21 = k12, k31 = k13, k32 = k23;
    const R J11 = 2 * (-b * c + F * F), J12 = 2 * (c * d - e * F), J13 = 2 * (b * e - d * F);
    const R J21 = 2 * (c * d - e * F), J22 = 2 * (-a * c + e * e), J23 = 2 * (-d * e + a * F);
    const R J31 = 2 * (b * e - d * F), J32 = 2 * (-d * e + a * F), J33 = 2 * (d * d - a * b);
    const R sb = -invJJ * sumBlock;
    R dv[4][3];
    dv[0][0] = enFactor * (sb * (J31 - J11) + (k31 - k11));
    dv[0][1] = enFactor * (sb * (J32 - J12) + (k32 - k12));
    dv[0][2] = enFactor * (sb * (J33 - J13) + (k33 - k13));
    dv[1][0] = enFactor * (sb * (J11) + (k11));
    dv[1][1] = enFactor * (sb * (J12) + (k12));
    dv[1][2] = enFactor * (sb * (J13) + (k13));
    dv[2][0] = enFactor * (sb * (-J31 - J21) + (-k31 - k21));
    dv[2][1] = enFactor * (sb * (-J32 - J22) + (-k32 - k22));
    dv[2][2] = enFactor * (sb * (-J33 - J23) + (-k33 - k23));
    dv[3][0] = enFactor * (sb * (J21) + (k21));
    dv[3][1] = enFactor * (sb * (J22) + (k22));
    dv[3][2] = enFactor * (sb * (J23) + (k23));
    const V3<T> u1d = sub(c2, c1), u2d = sub(c4, c3), u3d = sub(c1, c3);
    const R u1x = cvt<R>(u1d.x), u1y = cvt<R>(u1d.y), u1z = cvt<R>(u1d.z);
    const R u2x = cvt<R>(u2d.x), u2y = cvt<R>(u2d.y), u2z = cvt<R>(u2d.z);
    const R u3x = cvt<R>(u3d.x), u3y = cvt<R>(u3d.y), u3z = cvt<R>(u3d.z);
#pragma unroll
    for (int bi = 0; bi < 4; bi++) {
        f[bi].x = -(dv[bi][0] * u1x + dv[bi][1] * u2x + dv[bi][2] * u3x);
        f[bi].y = -(dv[bi][0] * u1y + dv[bi][1] * u2y + dv[bi][2] * u3y);
        f[bi].z = -(dv[bi][0] * u1z + dv[bi][1] * u2z + dv[bi][2] * u3z);
    }
}
__device__ __forceinline__ void ev_force_inv(const EvInv &v, d3 c1, d3 c2, d3 c3, d3 c4, double kr, d3 f[4]) {
    ev_force_inv_t<double>(v, c1, c2, c3, c4, kr, f);
}
__device__ __forceinline__ void ev_force(d3 c1, d3 c2, d3 c3, d3 c4, double kr, d3 f[4]) {
    EvInv v;
    ev_invariants(c1, c2, c3, c4, v);
    ev_force_inv(v, c1, c2, c3, c4, kr, f);
}
template <class T> __device__ __forceinline__ bool finite_d3(V3<T> v) {
    return xfinite(v.x) && xfinite(v.y) && xfinite(v.z);
}

// Simpson integrands (CylinderExclVolRepulsion.h:53-72)
__device__ __forceinline__ double ev_r2(double a, double b, double c, double d, double e, double F, double s, double t) {
    return c + 2 * s * e + s * s * a - 2 * t * F - 2 * s * t * d + t * t * b;
}

// ---------------------------------------------------------------------------
// Monomer-sum excluded volume (CylinderVolumeMon.hpp:31-100): the sum over
// sampled monomer pairs of coeff / d^4. Cylinder i samples the monomers m in
// [minus, plus) that are multiples of its interval, at the fraction
// (m - minus + 0.5) / (plus - minus) of the way from its first bead;
// coeff = kvol * seg1 * seg2 comes from the neighbor list. With FORCES, each
// monomer pair's gradient 4 coeff r / d^6 (r from the first point to the
// second) is shared between the two beads of each cylinder by the sample's
// fraction and subtracted, in loop order, from forces that start at zero.
// m = (minus serial, plus serial, interval, unused).
// ---------------------------------------------------------------------------
// smallest multiple of w (>= 1) that is >= a
__device__ __forceinline__ int first_multiple(int a, int w) {
    const int q = a >= 0 ? (a + w - 1) / w : -((-a) / w);
    return q * w;
}
template <class T>
__device__ __forceinline__ V3<T> along(V3<T> c1, V3<T> c2, double loc) {
    return V3<T>{c1.x + loc * (c2.x - c1.x), c1.y + loc * (c2.y - c1.y), c1.z + loc * (c2.z - c1.z)};
}
template <class T, bool FORCES>
__device__ __forceinline__ T evm_eval(V3<T> c11, V3<T> c12, V3<T> c21, V3<T> c22, int4 m1, int4 m2, double coeff,
                                      V3<T> f[4]) {
    const double n1 = (double)(m1.y - m1.x), n2 = (double)(m2.y - m2.x);
    const int s1 = first_multiple(m1.x, m1.z), s2 = first_multiple(m2.x, m2.z);
    const double coeff4 = coeff * 4;
    T en(0.0);
    if (FORCES) {
#pragma unroll
        for (int k = 0; k < 4; k++) f[k] = V3<T>{T(0.0), T(0.0), T(0.0)};
    }
    for (int a = s1; a < m1.y; a += m1.z) {
        const double loc1 = ((double)(a - m1.x) + 0.5) / n1;
        const V3<T> p1 = along(c11, c12, loc1);
        for (int b = s2; b < m2.y; b += m2.z) {
            const double loc2 = ((double)(b - m2.x) + 0.5) / n2;
            const V3<T> p2 = along(c21, c22, loc2);
            const T rx = p2.x - p1.x, ry = p2.y - p1.y, rz = p2.z - p1.z;
            const T d2 = rx * rx + ry * ry + rz * rz;
            en = en + coeff / (d2 * d2);
            if (FORCES) {
                const T g = coeff4 / (d2 * d2 * d2);
                const T gx = g * rx, gy = g * ry, gz = g * rz;
                const double w11 = 1 - loc1, w21 = -(1 - loc2), w22 = -loc2;
                f[0].x = f[0].x - w11 * gx; f[0].y = f[0].y - w11 * gy; f[0].z = f[0].z - w11 * gz;
                f[1].x = f[1].x - loc1 * gx; f[1].y = f[1].y - loc1 * gy; f[1].z = f[1].z - loc1 * gz;
                f[2].x = f[2].x - w21 * gx; f[2].y = f[2].y - w21 * gy; f[2].z = f[2].z - w21 * gz;
                f[3].x = f[3].x - w22 * gx; f[3].y = f[3].y - w22 * gy; f[3].z = f[3].z - w22 * gz;
            }
        }
    }
    return en;
}

}  // namespace mdy
