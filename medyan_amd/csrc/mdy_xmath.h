/*
 * mdy_xmath.h -- numerics shared by the HIP kernels and the CPU path so that
 * both produce the same bits:
 *
 *   mdy_atan, mdy_exp : arctangent and exponential written from their
 *     series (no vendor constants): argument reduction to a small interval,
 *     Taylor polynomial, table/split constants computed by
 *     gen_xmath_tables.py in 100-digit decimal arithmetic. Built from IEEE
 *     +,-,*,/ only (no fused multiply-add), so host and device agree exactly.
 *     Accuracy ~1 ulp (pinned against the C library in the tests).
 *     mdy_acos is built on mdy_atan (half-angle identity).
 *
 *   mdy_xsum : order-independent exact summation of doubles. Every term is
 *     placed on a fixed-point grid of unit 2^-144 (terms are truncated toward
 *     zero below it) and accumulated in seven signed 64-bit digits of radix
 *     2^48 with deferred carries; the final value is rounded to nearest even
 *     once. Any order of additions or merges gives the same bits, and for terms
 *     on the grid the result equals the correctly rounded sum (math.fsum).
 *     Magnitudes of 2^191 and above are treated as overflow (+-inf).
 *
 * Plain C99 and HIP; include after <math.h>/<stdint.h>.
 */
#ifndef MDY_XMATH_H
#define MDY_XMATH_H

#include <math.h>
#include <stdint.h>
#include <string.h>

#if defined(__HIPCC__)
#define MDY_HD __host__ __device__ static inline __attribute__((always_inline))
#else
#define MDY_HD static inline
#endif

/* ------------------------------------------------------------------------ */
/* atan                                                                      */
/* ------------------------------------------------------------------------ */
/* atan(k/8) = hi + lo, k = 1..8 (k = 0: 0) */
#define MDY_ATAN_H1 0x1.fd5ba9aac2f6ep-4
#define MDY_ATAN_H2 0x1.f5b75f92c80ddp-3
#define MDY_ATAN_H3 0x1.6f61941e4def1p-2
#define MDY_ATAN_H4 0x1.dac670561bb4fp-2
#define MDY_ATAN_H5 0x1.1e00babdefeb4p-1
#define MDY_ATAN_H6 0x1.4978fa3269ee1p-1
#define MDY_ATAN_H7 0x1.700a7c5784634p-1
#define MDY_ATAN_H8 0x1.921fb54442d18p-1
#define MDY_ATAN_L1 -0x1.cd37686760c17p-59
#define MDY_ATAN_L2 0x1.8ab6e3cf7afbdp-57
#define MDY_ATAN_L3 -0x1.c63aae6f6e918p-56
#define MDY_ATAN_L4 0x1.a2b7f222f65e2p-56
#define MDY_ATAN_L5 -0x1.928df287a668fp-58
#define MDY_ATAN_L6 0x1.2419a87f2a458p-56
#define MDY_ATAN_L7 -0x1.8c34d25aadef6p-56
#define MDY_ATAN_L8 0x1.1a62633145c07p-55
#define MDY_PIO2_HI 0x1.921fb54442d18p+0
#define MDY_PIO2_LO 0x1.1a62633145c07p-54

/* Straight-line code (selects, no branches): on the device every lane runs
 * the same instructions, which keeps the inlined copies short. */
MDY_HD double mdy_atan(double x) {
    const double ax = fabs(x);
    const int inv = ax > 1.0;
    const double y = inv ? 1.0 / ax : ax;        /* y in [0, 1]; 1/inf = 0 */
    /* nearest table point c = k/8, except below 3/32 where the series is
     * used directly: reducing y near 1/16 to t ~ -y would pass the rounding
     * error of t undiminished into a result of the same size */
    const int k = (y < 0.09375 || !(y <= 1.0)) ? 0 : (int)(y * 8.0 + 0.5);
    const double c = (double)k * 0.125;
    const double t = (y - c) / (1.0 + y * c);    /* |t| <= 3/32: atan(y) = atan(c) + atan(t) */
    const double t2 = t * t;
    /* atan(t) = t - t^3/3 + t^5/5 - ... ; |t|^17/17 < 2^-58 |t| */
    double p = -1.0 / 15.0;
    p = p * t2 + 1.0 / 13.0;
    p = p * t2 - 1.0 / 11.0;
    p = p * t2 + 1.0 / 9.0;
    p = p * t2 - 1.0 / 7.0;
    p = p * t2 + 1.0 / 5.0;
    p = p * t2 - 1.0 / 3.0;
    p = t + t * (t2 * p);
    double hi = 0.0, lo = 0.0;
    hi = k == 1 ? MDY_ATAN_H1 : hi; lo = k == 1 ? MDY_ATAN_L1 : lo;
    hi = k == 2 ? MDY_ATAN_H2 : hi; lo = k == 2 ? MDY_ATAN_L2 : lo;
    hi = k == 3 ? MDY_ATAN_H3 : hi; lo = k == 3 ? MDY_ATAN_L3 : lo;
    hi = k == 4 ? MDY_ATAN_H4 : hi; lo = k == 4 ? MDY_ATAN_L4 : lo;
    hi = k == 5 ? MDY_ATAN_H5 : hi; lo = k == 5 ? MDY_ATAN_L5 : lo;
    hi = k == 6 ? MDY_ATAN_H6 : hi; lo = k == 6 ? MDY_ATAN_L6 : lo;
    hi = k == 7 ? MDY_ATAN_H7 : hi; lo = k == 7 ? MDY_ATAN_L7 : lo;
    hi = k == 8 ? MDY_ATAN_H8 : hi; lo = k == 8 ? MDY_ATAN_L8 : lo;
    double r = hi + (lo + p);
    const double ri = (MDY_PIO2_HI - r) + MDY_PIO2_LO;
    r = inv ? ri : r;
    r = x < 0 ? -r : r;
    return (x != x) ? x : r;
}

/* acos from the half-angle identity acos(x) = 2 atan(sqrt((1 - x) / (1 + x))):
 * 1 - x (x >= 1/2) and 1 + x (x <= -1/2) are exact, so the quotient carries
 * at most ~1.5 ulp, halved by the square root, into an atan argument of the
 * same relative error; x = -1 gives 2 atan(inf) = pi. Accuracy ~2 ulp
 * (pinned against the C library in the tests). */
MDY_HD double mdy_acos(double x) {
    return 2.0 * mdy_atan(sqrt((1.0 - x) / (1.0 + x)));
}

/* ------------------------------------------------------------------------ */
/* exp                                                                       */
/* ------------------------------------------------------------------------ */
#define MDY_LN2_HI 0x1.62e42fee00000p-1
#define MDY_LN2_LO 0x1.a39ef35793c76p-33
#define MDY_INVLN2 0x1.71547652b82fep+0

MDY_HD double mdy_exp(double x) {
    if (x != x) return x;
    if (x > 709.782712893384) return INFINITY;
    if (x < -745.2) return 0.0;
    double kd = rint(x * MDY_INVLN2);
    int k = (int)kd;
    double r = (x - kd * MDY_LN2_HI) - kd * MDY_LN2_LO;   /* |r| <= ~0.35 */
    /* exp(r) = sum r^n / n!, n <= 14 (r^15/15! < 2^-80) */
    double p = 1.0 / 87178291200.0;
    p = p * r + 1.0 / 6227020800.0;
    p = p * r + 1.0 / 479001600.0;
    p = p * r + 1.0 / 39916800.0;
    p = p * r + 1.0 / 3628800.0;
    p = p * r + 1.0 / 362880.0;
    p = p * r + 1.0 / 40320.0;
    p = p * r + 1.0 / 5040.0;
    p = p * r + 1.0 / 720.0;
    p = p * r + 1.0 / 120.0;
    p = p * r + 1.0 / 24.0;
    p = p * r + 1.0 / 6.0;
    p = p * r + 0.5;
    p = p * r + 1.0;
    p = p * r + 1.0;
    /* 2^k * p in two exact steps (k may exceed the normal exponent range) */
    int k1 = k / 2, k2 = k - k1;
    return ldexp(ldexp(p, k1), k2);
}

/* ------------------------------------------------------------------------ */
/* exact summation                                                           */
/* ------------------------------------------------------------------------ */
#define MDY_XS_N 7
#define MDY_XS_BIAS 144
#define MDY_XS_MASK ((int64_t)0xFFFFFFFFFFFFll)
enum { MDY_XS_NAN = 1, MDY_XS_PINF = 2, MDY_XS_NINF = 4 };

/* Digits carry 16 spare bits: at most 2^15 additions may separate two
 * normalizations; mdy_xs_add normalizes itself every MDY_XS_RENORM adds. */
#define MDY_XS_RENORM 16384
typedef struct mdy_xsum {
    int64_t d[MDY_XS_N];
    int32_t flags;
    int32_t cnt;      /* additions since the last normalization */
} mdy_xsum;

MDY_HD uint64_t mdy_dbits(double v) {
#if defined(__HIP_DEVICE_COMPILE__)
    return (uint64_t)__double_as_longlong(v);
#else
    uint64_t u;
    memcpy(&u, &v, sizeof(u));
    return u;
#endif
}

MDY_HD void mdy_xs_zero(mdy_xsum *s) {
    for (int k = 0; k < MDY_XS_N; k++) s->d[k] = 0;
    s->flags = 0;
    s->cnt = 0;
}

/* carries into [0, 2^48) for all but the top digit */
MDY_HD void mdy_xs_norm(mdy_xsum *s) {
    for (int k = 0; k < MDY_XS_N - 1; k++) {
        int64_t c = s->d[k] >> 48;                /* floor division (arithmetic shift) */
        s->d[k] &= MDY_XS_MASK;
        s->d[k + 1] += c;
    }
    s->cnt = 0;
}

MDY_HD void mdy_xs_add(mdy_xsum *s, double v) {
    if (++s->cnt >= MDY_XS_RENORM) mdy_xs_norm(s);
    uint64_t b = mdy_dbits(v);
    int neg = (int)(b >> 63);
    int ex = (int)((b >> 52) & 0x7ff);
    uint64_t m = b & 0xFFFFFFFFFFFFFull;
    if (ex == 0x7ff) {
        s->flags |= m ? MDY_XS_NAN : (neg ? MDY_XS_NINF : MDY_XS_PINF);
        return;
    }
    int e;
    if (ex == 0) e = -1074;
    else { m |= 0x10000000000000ull; e = ex - 1075; }
    int p = e + MDY_XS_BIAS;                      /* grid position of m's unit bit */
    if (p < 0) {
        m = (-p >= 64) ? 0 : (m >> (-p));
        p = 0;
    }
    if (m == 0) return;
    if (p + 53 > 48 * (MDY_XS_N - 1) + 47) {      /* beyond the accumulator: overflow */
        s->flags |= neg ? MDY_XS_NINF : MDY_XS_PINF;
        return;
    }
    int j = p / 48, sh = p - 48 * j;
    int64_t lo = (int64_t)((m << sh) & (uint64_t)MDY_XS_MASK);
    int64_t mid = (int64_t)((sh == 0 ? (m >> 48) : (m >> (48 - sh))) & (uint64_t)MDY_XS_MASK);
    int64_t hi = (96 - sh < 64) ? (int64_t)(m >> (96 - sh)) : 0;
    if (neg) { lo = -lo; mid = -mid; hi = -hi; }
    for (int k = 0; k < MDY_XS_N; k++) {          /* unrolled selects: no indexed registers */
        int64_t add = (k == j) ? lo : (k == j + 1) ? mid : (k == j + 2) ? hi : 0;
        s->d[k] += add;
    }
}

MDY_HD void mdy_xs_merge(mdy_xsum *s, const mdy_xsum *o) {
    for (int k = 0; k < MDY_XS_N; k++) s->d[k] += o->d[k];
    s->flags |= o->flags;
}

MDY_HD double mdy_xs_value(const mdy_xsum *in) {
    if ((in->flags & MDY_XS_NAN) || ((in->flags & MDY_XS_PINF) && (in->flags & MDY_XS_NINF))) return NAN;
    if (in->flags & MDY_XS_PINF) return INFINITY;
    if (in->flags & MDY_XS_NINF) return -INFINITY;
    int64_t d[MDY_XS_N];
    for (int k = 0; k < MDY_XS_N; k++) d[k] = in->d[k];
    for (int k = 0; k < MDY_XS_N - 1; k++) { int64_t c = d[k] >> 48; d[k] &= MDY_XS_MASK; d[k + 1] += c; }
    int neg = d[MDY_XS_N - 1] < 0;
    if (neg) {
        for (int k = 0; k < MDY_XS_N; k++) d[k] = -d[k];
        for (int k = 0; k < MDY_XS_N - 1; k++) { int64_t c = d[k] >> 48; d[k] &= MDY_XS_MASK; d[k + 1] += c; }
    }
    /* highest nonzero digit h and the two below it, selected without
     * indexing (the digits stay in registers on the device) */
    int h = -1;
    for (int k = 0; k < MDY_XS_N; k++)
        if (d[k] != 0) h = k;
    if (h < 0) return 0.0;
    uint64_t top = 0, d1 = 0, d2 = 0;
    int sticky = 0;
    for (int k = 0; k < MDY_XS_N; k++) {
        top = (k == h) ? (uint64_t)d[k] : top;
        d1 = (k == h - 1) ? (uint64_t)d[k] : d1;
        d2 = (k == h - 2) ? (uint64_t)d[k] : d2;
        sticky |= (k < h - 2) && d[k] != 0;
    }
    int nb = 64 - __builtin_clzll(top);            /* bit length of the top digit (top != 0), 1..63 */
    int need = 64 - nb;                           /* bits taken from below the top digit */
    uint64_t W = (nb == 64) ? top : (top << need);
    if (need <= 48) {
        W |= (need == 0) ? 0 : (d1 >> (48 - need));
        sticky |= (d1 & ((1ull << (48 - need)) - 1)) != 0;
        sticky |= d2 != 0;
    } else {
        W |= (d1 << (need - 48)) | (d2 >> (96 - need));
        sticky |= (d2 & ((1ull << (96 - need)) - 1)) != 0;
    }
    uint64_t mant = W >> 11;
    int rbit = (int)((W >> 10) & 1);
    sticky |= (W & 0x3ff) != 0;
    int expo = 48 * h + nb - 53 - MDY_XS_BIAS;   /* exponent of mant's unit bit */
    if (rbit && (sticky || (mant & 1))) {
        mant++;
        if (mant == (1ull << 53)) { mant >>= 1; expo++; }
    }
    double r = ldexp((double)mant, expo);
    return neg ? -r : r;
}

#endif /* MDY_XMATH_H */
