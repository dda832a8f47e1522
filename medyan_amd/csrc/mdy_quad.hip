// mdy_quad.hip -- excluded-volume pairs evaluated by quadrature of their
// defining integral (MDY_COPLANAR_QUADRATURE).
//
// The cylinder-cylinder excluded-volume energy is
//     E = kr * Int_0^1 Int_0^1 ds dt / |P(s) - Q(t)|^4,
//     P(s) = c1 + s (c2 - c1),  Q(t) = c3 + t (c4 - c3),
// which CylinderExclVolRepulsion::energy (CylinderExclVolRepulsion.cpp:
// 28-120) evaluates in closed form. For nearly coplanar or nearly parallel
// cylinders that closed form is a difference of nearly equal atan terms
// divided by a vanishing triple product: the rounding of its FP64 terms is
// amplified as ~5e-14 / q, where q is the squared normalized triple product
// (v1 x v2 . v3)^2 / (|v1|^2 |v2|^2 |v3|^2), so its second derivatives are
// dominated by rounding noise there. In this mode the pairs with q < 1e-3
// (about 2-5% of the pairs of a random network; closed-form force error
// below ~5e-11 elsewhere) are evaluated instead by a composite 16-point
// Gauss-Legendre rule in s and t: m equal panels per cylinder, m =
// ceil(max length / segment distance) (at most 64), so every panel is no
// longer than the distance to the other segment and the rule's error is
// below ~1e-19 of the integral. Energy, forces, Hessian-vector products and
// directional derivatives of a pair all differentiate the same discrete
// rule, so they are mutually consistent to rounding.
//
// Work is split into chunks of at most kQuadChunk nodes (a pair with many
// panels spans several chunks, so the closest pair does not set the kernel's
// time): k_ev_classify flags the pairs and places each one's chunks, one
// block per chunk sums its nodes (each thread a fixed subset in a fixed
// order, then a fixed tree) into a chunk record, and one thread per pair adds
// its chunk records in chunk order. Results are reproducible bit for bit.
// The closed-form kernels skip the flagged pairs.
#include <cmath>

#include "mdy_eval.hpp"
#include "mdy_launch.hpp"

namespace mdy {

constexpr double kQuadMaxQ = 1e-3;   // pairs with q below this are integrated numerically
constexpr int kQuadPanels = 64;      // at most this many panels per cylinder
constexpr int kQuadChunk = 65536;    // nodes per chunk (256 per thread)

// 16-point Gauss-Legendre nodes and weights on [0, 1], ascending
struct GL16 {
    double t[16], w[16];
};

// Newton iteration on the Legendre polynomial P_16 in long double from the
// standard cosine initial guesses; weights 2 / ((1 - x^2) P_16'(x)^2).
static GL16 make_gl16() {
    GL16 g;
    const int n = 16;
    for (int i = 0; i < n; i++) {
        long double x = cosl(3.14159265358979323846264338327950288L * (i + 0.75L) / (n + 0.5L));
        long double p0 = 1, p1 = x, dp = 1;
        for (int it = 0; it < 100; it++) {
            p0 = 1;
            p1 = x;
            for (int k = 2; k <= n; k++) {
                const long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1);
            const long double dx = p1 / dp;
            x -= dx;
            if (fabsl(dx) < 1e-30L) break;
        }
        p0 = 1;
        p1 = x;
        for (int k = 2; k <= n; k++) {
            const long double p2 = ((2 * k - 1) * x * p1 - (k - 1) * p0) / k;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (x * p1 - p0) / (x * x - 1);
        g.t[i] = (double)(0.5L * (1 - x));
        g.w[i] = (double)(1 / ((1 - x * x) * dp * dp));   // (2 / ...) * 1/2 for [0, 1]
    }
    return g;
}

static const GL16 &gl16() {
    static const GL16 g = make_gl16();
    return g;
}

__device__ __forceinline__ d3 ld3(const double *x, int o) { return d3{x[o], x[o + 1], x[o + 2]}; }
__device__ __forceinline__ double dot3(d3 a, d3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }

// q = (v1 x v2 . v3)^2 / (|v1|^2 |v2|^2 |v3|^2); NaN (degenerate pair) counts as flagged
__device__ __forceinline__ bool needs_quadrature(d3 c1, d3 c2, d3 c3, d3 c4) {
    const d3 v1 = sub(c2, c1), v2 = sub(c4, c3), v3 = sub(c3, c1);
    const d3 n{v1.y * v2.z - v1.z * v2.y, v1.z * v2.x - v1.x * v2.z, v1.x * v2.y - v1.y * v2.x};
    const double tp = dot3(n, v3);
    const double q = tp * tp / (dot3(v1, v1) * dot3(v2, v2) * dot3(v3, v3));
    return !(q >= kQuadMaxQ);
}

// the distance helper is defined below; declared here for the classifier
__device__ __forceinline__ double segment_distance(d3 c1, d3 c2, d3 c3, d3 c4);

// panels per cylinder: each no longer than the segments' distance (at most kQuadPanels)
__device__ __forceinline__ int quad_panels(d3 c1, d3 c2, d3 c3, d3 c4) {
    const d3 v1 = sub(c2, c1), v2 = sub(c4, c3);
    const double L = sqrt(fmax(dot3(v1, v1), dot3(v2, v2)));
    const double dm = segment_distance(c1, c2, c3, c4);
    const double mr = dm > 0.0 ? ceil(L / dm) : (double)kQuadPanels;
    return mr >= (double)kQuadPanels ? kQuadPanels : mr <= 1.0 ? 1 : (int)mr;
}

// flag[j]; for a flagged pair the panel count pan[j], its chunk count
// cnt[j] and the index off[j] of its first chunk, taken from the running
// total *tot (atomically: the order of pairs among the chunks changes no
// result, since each chunk's nodes and each pair's sum over its chunks in
// chunk order are fixed), and the chunk -> pair map entries
__global__ void __launch_bounds__(kBlock) k_ev_classify(Dev D, const double *__restrict__ x,
                                                         unsigned char *__restrict__ flag, int *__restrict__ pan,
                                                         int *__restrict__ cnt, int *__restrict__ off,
                                                         int *__restrict__ cmap, int *__restrict__ tot, int cap) {
    // one atomic per block and sweep turn: the block's chunk counts are
    // scanned in LDS (wave prefix sums, then the wave totals)
    __shared__ int wsum[kBlock / 64];
    __shared__ int bbase;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int j0 = blockIdx.x * blockDim.x; j0 < D.n_ev; j0 += gridDim.x * blockDim.x) {
        const int j = j0 + threadIdx.x;
        int nc = 0, m = 0;
        if (j < D.n_ev) {
            const int4 b = D.ev_bs[j];
            const d3 c1 = ld3(x, b.x), c2 = ld3(x, b.y), c3 = ld3(x, b.z), c4 = ld3(x, b.w);
            const bool f = needs_quadrature(c1, c2, c3, c4);
            flag[j] = f ? 1 : 0;
            if (f) {
                m = quad_panels(c1, c2, c3, c4);
                nc = (256 * m * m + kQuadChunk - 1) / kQuadChunk;   // (16 m)^2 nodes
            }
        }
        int incl = nc;   // inclusive prefix within the wave
#pragma unroll
        for (int o = 1; o < 64; o <<= 1) {
            const int v = __shfl_up(incl, o, 64);
            if (lane >= o) incl += v;
        }
        if (lane == 63) wsum[wave] = incl;
        __syncthreads();
        if (threadIdx.x == 0) {
            int s = 0;
            for (int w = 0; w < kBlock / 64; w++) { const int v = wsum[w]; wsum[w] = s; s += v; }
            bbase = s ? atomicAdd(tot, s) : 0;
        }
        __syncthreads();
        if (j < D.n_ev) {
            if (nc) {
                const int base = bbase + wsum[wave] + incl - nc;
                pan[j] = m;
                off[j] = base;
                if (base + nc <= cap) {
                    for (int q = 0; q < nc; q++) cmap[base + q] = j;
                } else {
                    set_err(D.err, MDY_ERR_CAPACITY);   // cannot happen with a cleared total (16 chunks per pair at most)
                }
            }
            cnt[j] = nc;
        }
        __syncthreads();   // wsum and bbase are rewritten by the next turn
    }
}

__global__ void k_quad_zero(int *tot) { *tot = 0; }

__device__ __forceinline__ double clamp01(double v) { return v < 0.0 ? 0.0 : v > 1.0 ? 1.0 : v; }

// Distance between the segments c1-c2 and c3-c4: the minimum over the unit
// square of |c1 + s d1 - c3 - t d2|, from the stationary point of that
// quadratic, clamped to the square edge by edge (for parallel segments s
// starts at 0).
__device__ __forceinline__ double segment_distance(d3 c1, d3 c2, d3 c3, d3 c4) {
    const d3 d1 = sub(c2, c1), d2 = sub(c4, c3), r = sub(c1, c3);
    const double a = dot3(d1, d1), e = dot3(d2, d2), f = dot3(d2, r), c = dot3(d1, r), b = dot3(d1, d2);
    const double den = a * e - b * b;
    double s = den > 1e-14 * a * e ? clamp01((b * f - c * e) / den) : 0.0;
    double t = (b * s + f) / e;
    if (t < 0.0) {
        t = 0.0;
        s = clamp01(-c / a);
    } else if (t > 1.0) {
        t = 1.0;
        s = clamp01((b - c) / a);
    }
    const d3 p{r.x + s * d1.x - t * d2.x, r.y + s * d1.y - t * d2.y, r.z + s * d1.z - t * d2.z};
    return sqrt(dot3(p, p));
}

enum { Q_ENERGY = 0, Q_FORCE = 1, Q_HVP = 2, Q_JVP = 3 };
template <int W> struct QuadAcc { static constexpr int N = W == Q_ENERGY ? 1 : 9; };

// One block per chunk: the chunk's nodes of its pair into part[chunk]. For
// Q_ENERGY the sum of W |r|^-4; otherwise G, sum of s h, sum of t h with
// h = W * 4 r/|r|^6 (Q_FORCE, Q_JVP: minus the gradient of |r|^-4) or
// h = W grad^2 |r|^-4 dr (Q_HVP, dr the direction v along r).
template <int W>
__global__ void __launch_bounds__(kBlock) k_quad_chunks(Dev D, const double *__restrict__ x,
                                                        const double *__restrict__ v, GL16 g,
                                                        const int *__restrict__ off, const int *__restrict__ cmap,
                                                        const int *__restrict__ pan, const int *__restrict__ tot, int cap,
                                                        double *__restrict__ part) {
    constexpr int NA = QuadAcc<W>::N;
    __shared__ double lds[17 * NA + 8];
    __shared__ double gt[16], gw[16];   // the rule, indexed per node (no dynamic indexing of kernel arguments)
    if (threadIdx.x < 16) {
        gt[threadIdx.x] = g.t[threadIdx.x];
        gw[threadIdx.x] = g.w[threadIdx.x];
    }
    __syncthreads();
    const int total = min(*tot, cap);
    for (int ch = blockIdx.x; ch < total; ch += gridDim.x) {
        const int j = cmap[ch];
        const int4 b = D.ev_bs[j];
        const d3 c1 = ld3(x, b.x), c2 = ld3(x, b.y), c3 = ld3(x, b.z), c4 = ld3(x, b.w);
        const d3 v1 = sub(c2, c1), v2 = sub(c4, c3), c13 = sub(c1, c3);
        const int m = pan[j];
        const int n = 16 * m, nn = n * n;
        const int k0 = (ch - off[j]) * kQuadChunk, k1 = min(nn, k0 + kQuadChunk);
        const double im = 1.0 / (double)m;
        d3 e1{0, 0, 0}, e2{0, 0, 0}, e13{0, 0, 0};   // HVP: the direction at (c2 - c1), (c4 - c3), (c1 - c3)
        if (W == Q_HVP) {
            const d3 u1 = ld3(v, b.x), u2 = ld3(v, b.y), u3 = ld3(v, b.z), u4 = ld3(v, b.w);
            e1 = sub(u2, u1);
            e2 = sub(u4, u3);
            e13 = sub(u1, u3);
        }
        double acc[NA];
#pragma unroll
        for (int k = 0; k < NA; k++) acc[k] = 0.0;
        for (int k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
            const int i = k / n, l = k - i * n;
            const double s = ((double)(i >> 4) + gt[i & 15]) * im, ws = gw[i & 15] * im;
            const double t = ((double)(l >> 4) + gt[l & 15]) * im, wt = gw[l & 15] * im;
            const d3 r{c13.x + s * v1.x - t * v2.x, c13.y + s * v1.y - t * v2.y, c13.z + s * v1.z - t * v2.z};
            const double d2 = dot3(r, r), inv = 1.0 / d2, inv2 = inv * inv, wst = ws * wt;
            if (W == Q_ENERGY) {
                acc[0] += wst * inv2;
            } else {
                d3 h;
                if (W == Q_HVP) {
                    // d/dt of the gradient of |r|^-4 along dr: -4 dr/|r|^6 + 24 (r.dr) r/|r|^8
                    const d3 dr{e13.x + s * e1.x - t * e2.x, e13.y + s * e1.y - t * e2.y, e13.z + s * e1.z - t * e2.z};
                    const double a = -4.0 * wst * inv2 * inv, c = 24.0 * wst * inv2 * inv2 * dot3(r, dr);
                    h = d3{a * dr.x + c * r.x, a * dr.y + c * r.y, a * dr.z + c * r.z};
                } else {
                    // -grad of |r|^-4 with respect to r: 4 r / |r|^6
                    const double a = 4.0 * wst * inv2 * inv;
                    h = d3{a * r.x, a * r.y, a * r.z};
                }
                acc[0] += h.x; acc[1] += h.y; acc[2] += h.z;
                acc[3] += s * h.x; acc[4] += s * h.y; acc[5] += s * h.z;
                acc[6] += t * h.x; acc[7] += t * h.y; acc[8] += t * h.z;
            }
        }
        block_reduce<NA>(acc, lds);
        if (threadIdx.x < NA) part[(size_t)ch * NA + threadIdx.x] = acc[threadIdx.x];
    }
}

// One thread per pair: its chunk records added in chunk order. Q_FORCE /
// Q_HVP: the pair's four slots; Q_ENERGY: pe[j] = its energy; Q_JVP:
// pe[j] = grad E_pair . w (pe[j] = 0 for pairs left to the closed form).
// The running chunk total is cleared for the next call (this kernel follows
// the chunk kernel on the stream, so every chunk has run).
template <int W>
__global__ void __launch_bounds__(kBlock) k_quad_finish(Dev D, const double *__restrict__ v,
                                                        const int *__restrict__ off, const int *__restrict__ cnt,
                                                        const double *__restrict__ part, double *__restrict__ slots,
                                                        double *__restrict__ pe, int *__restrict__ tot, int cap) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *tot = 0;
    constexpr int NA = QuadAcc<W>::N;
    bool ok = true;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < D.n_ev; j += gridDim.x * blockDim.x) {
        if (!D.evq[j]) {
            if (W == Q_ENERGY || W == Q_JVP) pe[j] = 0.0;
            continue;
        }
        double acc[NA];
#pragma unroll
        for (int k = 0; k < NA; k++) acc[k] = 0.0;
        for (int q = off[j], qe = min(off[j] + cnt[j], cap); q < qe; q++)
#pragma unroll
            for (int k = 0; k < NA; k++) acc[k] += part[(size_t)q * NA + k];
        const double kr = D.ev_kr[j];
        if constexpr (W == Q_ENERGY) {
            pe[j] = kr * acc[0];
        } else {
            // bead weights of r = P(s) - Q(t): c1 (1 - s), c2 s, c3 -(1 - t), c4 -t. For
            // Q_FORCE / Q_JVP the terms were -W grad |r|^-4 (f = forces); for Q_HVP
            // W grad^2 |r|^-4 dr (f = H v)
            const d3 G{acc[0], acc[1], acc[2]}, Gs{acc[3], acc[4], acc[5]}, Gt{acc[6], acc[7], acc[8]};
            d3 f[4];
            f[0] = d3{kr * (G.x - Gs.x), kr * (G.y - Gs.y), kr * (G.z - Gs.z)};
            f[1] = d3{kr * Gs.x, kr * Gs.y, kr * Gs.z};
            f[2] = d3{-kr * (G.x - Gt.x), -kr * (G.y - Gt.y), -kr * (G.z - Gt.z)};
            f[3] = d3{-kr * Gt.x, -kr * Gt.y, -kr * Gt.z};
            ok = ok && finite_d3(f[0]) && finite_d3(f[1]) && finite_d3(f[2]) && finite_d3(f[3]);
            if constexpr (W == Q_JVP) {
                const int4 b = D.ev_bs[j];
                const int o[4] = {b.x, b.y, b.z, b.w};
                double sdot = 0.0;
#pragma unroll
                for (int q = 0; q < 4; q++) sdot -= f[q].x * v[o[q]] + f[q].y * v[o[q] + 1] + f[q].z * v[o[q] + 2];
                pe[j] = sdot;
            } else {
                const int s0 = D.slot_ev + 4 * j;
#pragma unroll
                for (int q = 0; q < 4; q++) put3(slots, s0 + q, f[q]);
            }
        }
    }
    if (!ok) set_err(D.err, MDY_ERR_FORCE);
}

// Block records [rec0 + block] of the per-pair values: pairs j = block * 256
// + thread + k * (grid * 256), each thread's share summed in order, then the
// block's fixed tree; EV entry = that double sum, and for Q_ENERGY the exact
// total in the record's accumulator.
template <int W>
__global__ void __launch_bounds__(kBlock) k_quad_records(Dev D, const double *__restrict__ pe, double *__restrict__ out,
                                                         int rec0) {
    __shared__ double lds[17 * 8 + 8];
    double s[1] = {0.0};
    mdy_xsum X;
    mdy_xs_zero(&X);
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < D.n_ev; j += gridDim.x * blockDim.x) {
        if (!D.evq[j]) continue;
        s[0] += pe[j];
        if (W == Q_ENERGY) mdy_xs_add(&X, pe[j]);
    }
    block_reduce<1>(s, lds);
    if (W == Q_ENERGY) block_reduce_xsum(X, reinterpret_cast<long long *>(lds));
    if (threadIdx.x == 0) {
        double *r = out + (size_t)(rec0 + blockIdx.x) * kRed;
        for (int k = 0; k < kNFF; k++) r[k] = k == MDY_FF_EV ? s[0] : 0.0;
        if (W == Q_ENERGY) store_xsum(out, rec0 + blockIdx.x, kXsE, X);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
bool quad_on(const Ctx &c) {
    return c.opts.coplanar_mode == MDY_COPLANAR_QUADRATURE && !c.ev_mon && c.n_ev > 0;
}

int quad_blocks(const Ctx &c) {
    const int g = c.num_cu * 4;
    return c.n_ev < g ? (c.n_ev < 1 ? 1 : c.n_ev) : g;
}

int quad_classify(Ctx &c, Dev &D, const double *d_x, hipStream_t st) {
    const size_t n = (size_t)c.n_ev;
    MDY_TRY(ensure(c.evq, n));
    MDY_TRY(ensure(c.qpan, n));
    MDY_TRY(ensure(c.qcnt, n));
    MDY_TRY(ensure(c.qoff, n));
    MDY_TRY(ensure(c.qmap, 16 * n));   // at most (16 * kQuadPanels)^2 / kQuadChunk = 16 chunks per pair
    // everything the quadrature kernels use is allocated before the chunk total
    // is taken, so nothing fails between this call and the kernel that clears it
    MDY_TRY(ensure(c.qpart, 16 * n * 9));
    MDY_TRY(ensure(c.qpe, n));
    if (!c.qtot.p) {
        MDY_TRY(ensure(c.qtot, 1));
        hipLaunchKernelGGL(k_quad_zero, dim3(1), dim3(1), 0, st, c.qtot.p);
    }
    long long g = ((long long)n + kBlock - 1) / kBlock, cap = (long long)c.num_cu * 8;
    const dim3 grid((int)(g > cap ? cap : g < 1 ? 1 : g));
    hipLaunchKernelGGL(k_ev_classify, grid, dim3(kBlock), 0, st, D, d_x, c.evq.p, c.qpan.p, c.qcnt.p, c.qoff.p,
                       c.qmap.p, c.qtot.p, (int)(16 * n));
    D.evq = c.evq.p;
    return hipGetLastError() == hipSuccess ? MDY_OK : MDY_ERR_HIP;
}

template <int W>
static void quad_launch(Ctx &c, const Dev &D, const double *d_x, const double *d_v, double *out, int rec0,
                        hipStream_t st) {
    const long long n = c.n_ev;
    const dim3 blk(kBlock), chunks(c.num_cu * 8);
    long long g = (n + kBlock - 1) / kBlock, cap = (long long)c.num_cu * 8;
    const dim3 pairs((int)(g > cap ? cap : g < 1 ? 1 : g));
    const int ccap = (int)(16 * n);
    hipLaunchKernelGGL(k_quad_chunks<W>, chunks, blk, 0, st, D, d_x, d_v, gl16(), c.qoff.p, c.qmap.p, c.qpan.p,
                       c.qtot.p, ccap, c.qpart.p);
    hipLaunchKernelGGL(k_quad_finish<W>, pairs, blk, 0, st, D, d_v, c.qoff.p, c.qcnt.p, c.qpart.p, out, c.qpe.p,
                       c.qtot.p, ccap);
    if (W == Q_ENERGY || W == Q_JVP)
        hipLaunchKernelGGL(k_quad_records<W>, dim3(quad_blocks(c)), blk, 0, st, D, c.qpe.p, out, rec0);
}

int launch_quad(Ctx &c, const Dev &D, int what, const double *d_x, const double *d_v, double *out, int rec0,
                hipStream_t st) {
    switch (what) {
        case 0: quad_launch<Q_ENERGY>(c, D, d_x, d_v, out, rec0, st); break;
        case 1: quad_launch<Q_FORCE>(c, D, d_x, d_v, out, rec0, st); break;
        case 2: quad_launch<Q_HVP>(c, D, d_x, d_v, out, rec0, st); break;
        default: quad_launch<Q_JVP>(c, D, d_x, d_v, out, rec0, st); break;
    }
    return hipGetLastError() == hipSuccess ? MDY_OK : MDY_ERR_HIP;
}

int quad_pair_count(Ctx &c, const double *d_x, hipStream_t st, int *count) {
    *count = 0;
    if (!quad_on(c)) return MDY_OK;
    Dev D = c.dev();
    MDY_TRY(quad_classify(c, D, d_x, st));
    hipLaunchKernelGGL(k_quad_zero, dim3(1), dim3(1), 0, st, c.qtot.p);   // no quadrature follows: clear the total
    std::vector<unsigned char> h((size_t)c.n_ev);
    MDY_HIP(hipMemcpyAsync(h.data(), c.evq.p, h.size(), hipMemcpyDeviceToHost, st));
    MDY_HIP(hipStreamSynchronize(st));
    int n = 0;
    for (unsigned char f : h) n += f;
    *count = n;
    return MDY_OK;
}

}  // namespace mdy
