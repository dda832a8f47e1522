// mdy_eval.hpp -- evaluation of all interactions over the unified index space,
// per-bead force gather, block reductions and the quadrature fallbacks.
// Used both by the one-shot kernels (mdy_eval.hip) and by the persistent
// conjugate-gradient kernel (mdy_cg.hip).
#pragma once
#include "mdy_device.hpp"
#include "mdy_internal.hpp"

namespace mdy {

// ---- where bead positions come from ---------------------------------------
struct PosX {  // x
    const double *__restrict__ x;
    __device__ __forceinline__ d3 operator()(int o) const { return d3{x[o], x[o + 1], x[o + 2]}; }
};
struct PosLine {  // x + lam * d   (CGMethod::calcCoordLineSearch, CGMethod.cpp:43-46)
    const double *__restrict__ x;
    const double *__restrict__ d;
    double lam;
    __device__ __forceinline__ d3 operator()(int o) const {
        return d3{x[o] + lam * d[o], x[o + 1] + lam * d[o + 1], x[o + 2] + lam * d[o + 2]};
    }
};
struct PosSpec {  // x + lam * (F + beta * d): trial along the not-yet-stored new direction
    const double *__restrict__ x;
    const double *__restrict__ F;
    const double *__restrict__ d;
    double beta, lam;
    __device__ __forceinline__ double c(int o) const {
        double s = F[o] + beta * d[o];
        return x[o] + lam * s;
    }
    __device__ __forceinline__ d3 operator()(int o) const { return d3{c(o), c(o + 1), c(o + 2)}; }
};

__device__ __forceinline__ void set_err(int *err, int code) { atomicCAS(err, 0, code); }

__device__ __forceinline__ void enqueue(int *q, int *qn, int v, int *err) {
    int s = atomicAdd(qn, 1);
    if (s < kQueueCap) q[s] = v;
    else set_err(err, MDY_ERR_CAPACITY);
}

// ---- energies ----------------------------------------------------------------
template <class P>
__device__ __forceinline__ void energy_item(const Dev &D, int i, const P &pos, double acc[kNFF]) {
    if (i < D.n_ev) {
        int4 b = D.ev_bs[i];
        double U = ev_energy(pos(b.x), pos(b.y), pos(b.z), pos(b.w), D.ev_kr[i]);
        if (ev_energy_bad(U)) enqueue(D.qe, D.qe_n, i, D.err);
        else acc[MDY_FF_EV] += U;
        return;
    }
    i -= D.n_ev;
    if (i < D.n_fs) {
        int2 b = D.fs_bs[i];
        acc[MDY_FF_FS] += fs_energy(pos(b.x), pos(b.y), D.fs_k[i], D.fs_l0[i]);
        return;
    }
    i -= D.n_fs;
    if (i < D.n_fb) {
        const int *b = D.fb_bs + 3 * i;
        acc[MDY_FF_FB] += fb_energy(pos(b[0]), pos(b[1]), pos(b[2]), D.fb_k[i], D.fb_th[i]);
        return;
    }
    i -= D.n_fb;
    if (i < D.n_lm) {
        int4 b = D.lm_bs[i];
        double U = ls_energy(pos(b.x), pos(b.y), pos(b.z), pos(b.w), D.lm_k[i], D.lm_l0[i], D.lm_p1[i], D.lm_p2[i]);
        if (i < D.n_ls) acc[MDY_FF_LS] += U;
        else acc[MDY_FF_MS] += U;
        return;
    }
    i -= D.n_lm;
    if (i < D.n_bs) {
        const int *b = D.bs_bs + 3 * i;
        acc[MDY_FF_BS] += bs_energy(pos(b[0]), pos(b[1]), pos(b[2]), D.bs_k[i], D.bs_l0[i], D.bs_p[i]);
        return;
    }
    i -= D.n_bs;
    if (i < D.n_bb) {
        int4 b = D.bb_bs[i];
        acc[MDY_FF_BB] += bb_energy(pos(b.x), pos(b.y), pos(b.z), pos(b.w), D.bb_k[i], D.bb_th[i]);
        return;
    }
    i -= D.n_bb;
    if (i < D.n_bd) {
        acc[MDY_FF_BD] += bd_energy(D.planes + 4 * D.bd_plane[i], pos(D.bd_bead[i]), D.bd_k[i], D.bd_sl[i]);
    }
}

__device__ __forceinline__ void put3(double *__restrict__ c, int slot, d3 v) {
    c[3 * slot] = v.x; c[3 * slot + 1] = v.y; c[3 * slot + 2] = v.z;
}

// ---- forces: each interaction writes its per-bead contributions to its slots
template <class P>
__device__ __forceinline__ void force_item(const Dev &D, int i, const P &pos, double *__restrict__ contrib) {
    if (i < D.n_ev) {
        int4 b = D.ev_bs[i];
        d3 c1 = pos(b.x), c2 = pos(b.y), c3 = pos(b.z), c4 = pos(b.w);
        d3 c2u = c2;
        // coplanar perturbation of the reference force path (CylinderExclVolRepulsion.cpp:610-626)
        if (D.coplanar_mode == MDY_COPLANAR_REFERENCE && are_in_plane(c1, c2, c3, c4))
            c2u = moved_off_plane(c1, c2, c3, c2, 0.01);
        d3 f[4];
        ev_force(c1, c2u, c3, c4, D.ev_kr[i], f);
        if (!(finite_d3(f[0]) && finite_d3(f[1]) && finite_d3(f[2]) && finite_d3(f[3]))) {
            enqueue(D.qf, D.qf_n, i, D.err);
            return;
        }
        int s = D.slot_ev + 4 * i;
        put3(contrib, s, f[0]); put3(contrib, s + 1, f[1]); put3(contrib, s + 2, f[2]); put3(contrib, s + 3, f[3]);
        return;
    }
    i -= D.n_ev;
    if (i < D.n_fs) {
        int2 b = D.fs_bs[i];
        d3 f1, f2;
        fs_force(pos(b.x), pos(b.y), D.fs_k[i], D.fs_l0[i], f1, f2);
        int s = D.slot_fs + 2 * i;
        put3(contrib, s, f1); put3(contrib, s + 1, f2);
        return;
    }
    i -= D.n_fs;
    if (i < D.n_fb) {
        const int *b = D.fb_bs + 3 * i;
        d3 f1, f2, f3;
        fb_force(pos(b[0]), pos(b[1]), pos(b[2]), D.fb_k[i], D.fb_th[i], f1, f2, f3);
        int s = D.slot_fb + 3 * i;
        put3(contrib, s, f1); put3(contrib, s + 1, f2); put3(contrib, s + 2, f3);
        return;
    }
    i -= D.n_fb;
    if (i < D.n_lm) {
        int4 b = D.lm_bs[i];
        d3 f1, f2, f3, f4;
        ls_force(pos(b.x), pos(b.y), pos(b.z), pos(b.w), D.lm_k[i], D.lm_l0[i], D.lm_p1[i], D.lm_p2[i], f1, f2, f3, f4);
        int s = D.slot_lm + 4 * i;
        put3(contrib, s, f1); put3(contrib, s + 1, f2); put3(contrib, s + 2, f3); put3(contrib, s + 3, f4);
        return;
    }
    i -= D.n_lm;
    if (i < D.n_bs) {
        const int *b = D.bs_bs + 3 * i;
        d3 f1, f2, f3;
        bs_force(pos(b[0]), pos(b[1]), pos(b[2]), D.bs_k[i], D.bs_l0[i], D.bs_p[i], f1, f2, f3);
        int s = D.slot_bs + 3 * i;
        put3(contrib, s, f1); put3(contrib, s + 1, f2); put3(contrib, s + 2, f3);
        return;
    }
    i -= D.n_bs;
    if (i < D.n_bb) {
        int4 b = D.bb_bs[i];
        d3 f1, f2, f3, f4;
        bb_force(pos(b.x), pos(b.y), pos(b.z), pos(b.w), D.bb_k[i], D.bb_th[i], f1, f2, f3, f4);
        int s = D.slot_bb + 4 * i;
        put3(contrib, s, f1); put3(contrib, s + 1, f2); put3(contrib, s + 2, f3); put3(contrib, s + 3, f4);
        return;
    }
    i -= D.n_bb;
    if (i < D.n_bd) {
        put3(contrib, D.slot_bd + i, bd_force(D.planes + 4 * D.bd_plane[i], pos(D.bd_bead[i]), D.bd_k[i], D.bd_sl[i]));
    }
}

// Per-bead force: the slots of bead b in reference order, starting from 0.0
// (ForceFieldManager::computeForces zero-fills, then each force field adds).
__device__ __forceinline__ d3 gather_bead(const Dev &D, const double *__restrict__ contrib, int b) {
    double fx = 0.0, fy = 0.0, fz = 0.0;
    for (int k = D.st_off[b], e = D.st_off[b + 1]; k < e; k++) {
        int s = D.st_idx[k];
        fx += contrib[3 * s]; fy += contrib[3 * s + 1]; fz += contrib[3 * s + 2];
    }
    for (int k = D.dy_off[b], e = D.dy_off[b + 1]; k < e; k++) {
        int s = D.dy_idx[k];
        fx += contrib[3 * s]; fy += contrib[3 * s + 1]; fz += contrib[3 * s + 2];
    }
    return d3{fx, fy, fz};
}

// ---- deterministic block reductions (fixed butterfly + fixed wave order) ----
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_min(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}
// Reduces N values per thread to N values (valid in thread 0, and broadcast to
// all threads through lds). op: 0 sum, 1 max, 2 min per entry given by mask bits.
template <int N>
__device__ __forceinline__ void block_reduce(double v[N], double *lds, unsigned max_mask = 0, unsigned min_mask = 0) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < N; k++) {
        double r = ((max_mask >> k) & 1) ? wave_max(v[k]) : ((min_mask >> k) & 1) ? wave_min(v[k]) : wave_sum(v[k]);
        if (lane == 0) lds[wave * N + k] = r;
    }
    __syncthreads();
    if (threadIdx.x < N) {
        int k = threadIdx.x;
        double r = lds[k];
        for (int w = 1; w < nw; w++) {
            double o = lds[w * N + k];
            r = ((max_mask >> k) & 1) ? fmax(r, o) : ((min_mask >> k) & 1) ? fmin(r, o) : r + o;
        }
        lds[16 * N + k] = r;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = lds[16 * N + k];
    __syncthreads();
}

// Reduce per-block partials red[blk * kRed + k] for k < N over nblk blocks in
// a fixed order; every calling block obtains bit-identical results.
template <int N>
__device__ __forceinline__ void grid_partials_reduce(const double *red, int nblk, double v[N], double *lds,
                                                     unsigned max_mask = 0, unsigned min_mask = 0) {
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = ((max_mask >> k) & 1) ? -INFINITY : ((min_mask >> k) & 1) ? INFINITY : 0.0;
    for (int blk = threadIdx.x; blk < nblk; blk += blockDim.x) {
#pragma unroll
        for (int k = 0; k < N; k++) {
            double o = __builtin_nontemporal_load(red + (size_t)blk * kRed + k);
            v[k] = ((max_mask >> k) & 1) ? fmax(v[k], o) : ((min_mask >> k) & 1) ? fmin(v[k], o) : v[k] + o;
        }
    }
    block_reduce<N>(v, lds, max_mask, min_mask);
}

// Energy epilogue, shared with the persistent kernel: per-FF totals from the
// block partials and the (pair-ordered) quadrature fallbacks; out[0..7] per
// FF, out[8] = computeEnergy total (+inf sentinel), out[9] = sum without sentinel.
__device__ __forceinline__ void energy_epilogue(const Dev &D, const double *red, int nblk, int nq, const double *qe_u, double *lds,
                                double e[kNFF + 2]) {
    double v[kNFF];
    grid_partials_reduce<kNFF>(red, nblk, v, lds);
    nq = min(nq, kQueueCap);
    if (nq > 0) {
        // add fallback energies in increasing pair order (selection; queues are short)
        double add = 0.0;
        bool bad = false;
        int last = -1;
        for (int r = 0; r < nq; r++) {
            int best = -1, bi = 0x7fffffff;
            for (int q = 0; q < nq; q++) {
                int pi = D.qe[q];
                if (pi > last && pi < bi) { bi = pi; best = q; }
            }
            if (best < 0) break;
            last = bi;
            double U = qe_u[best];
            if (ev_energy_bad(U)) bad = true;
            add += U;
        }
        v[MDY_FF_EV] = bad ? -1.0 : v[MDY_FF_EV] + add;
    }
    double tot = 0.0, ind = 0.0;
    bool inf = false;
    for (int k = 0; k < kNFF; k++) {
        e[k] = v[k];
        if (v[k] <= -1) inf = true;
        if (!inf) tot += v[k];
        ind += v[k];
    }
    e[kNFF] = inf ? INFINITY : tot;
    e[kNFF + 1] = ind;
}

// ---- Simpson quadrature fallbacks (CylinderExclVolRepulsion.cpp:884-1254) ----
// The reference advances its abscissae before first use, so interior node i
// sits at (i+1)/N; node values are formed here as (i+1)*(1/N).
struct SimpsonGeo { double a, b, c, d, e, F; d3 vA, vB, vC; };
__device__ __forceinline__ SimpsonGeo simpson_geo(d3 c1, d3 c2, d3 c3, d3 c4) {
    SimpsonGeo g;
    g.a = sprod(c1, c2, c1, c2); g.b = sprod(c3, c4, c3, c4); g.c = sprod(c3, c1, c3, c1);
    g.d = sprod(c1, c2, c3, c4); g.e = sprod(c1, c2, c3, c1); g.F = sprod(c3, c4, c3, c1);
    g.vA = sub(c1, c2); g.vB = sub(c3, c4); g.vC = sub(c3, c1);
    return g;
}
// energy integrand contribution of grid node (i, j) with the reference's
// bookkeeping: returns the weighted term that node adds to Termset2/3/4 (or 0)
__device__ __forceinline__ double simpson_e_node(const SimpsonGeo &g, int i, int j, int N) {
    const double ds = 1.0 / ((double)N);
    double s = (double)(i + 1) * ds, t = (double)(j + 1) * ds;
    double fs = (i % 2 == 0) ? 2.0 : 4.0, ft = (j % 2 == 0) ? 2.0 : 4.0;
    double r = 0.0;
    if (j == 0 && i > 0 && i < N) {  // Termset2 (once per row)
        double r0 = ev_r2(g.a, g.b, g.c, g.d, g.e, g.F, s, 0.0), r1 = ev_r2(g.a, g.b, g.c, g.d, g.e, g.F, s, 1.0);
        r += fs * ((1.0 / (r0 * r0)) + (1.0 / (r1 * r1)));
    }
    if (j > 0 && j < N) {
        if (i == 0) {  // Termset3
            double r0 = ev_r2(g.a, g.b, g.c, g.d, g.e, g.F, 0.0, t), r1 = ev_r2(g.a, g.b, g.c, g.d, g.e, g.F, 1.0, t);
            r += ft * ((1.0 / (r0 * r0)) + (1.0 / (r1 * r1)));
        } else if (i < N) {  // Termset4
            double r2 = ev_r2(g.a, g.b, g.c, g.d, g.e, g.F, s, t);
            r += fs * ft * (1.0 / (r2 * r2));
        }
    }
    return r;
}
__device__ __forceinline__ double simpson_e_corners(const SimpsonGeo &g) {
    double r00 = ev_r2(g.a, g.b, g.c, g.d, g.e, g.F, 0.0, 0.0), r01 = ev_r2(g.a, g.b, g.c, g.d, g.e, g.F, 0.0, 1.0);
    double r10 = ev_r2(g.a, g.b, g.c, g.d, g.e, g.F, 1.0, 0.0), r11 = ev_r2(g.a, g.b, g.c, g.d, g.e, g.F, 1.0, 1.0);
    return (1.0 / (r00 * r00)) + (1.0 / (r01 * r01)) + (1.0 / (r10 * r10)) + (1.0 / (r11 * r11));
}
__device__ __forceinline__ void fint6(const SimpsonGeo &g, double s, double t, double w, double o[6]) {
    double r2 = ev_r2(g.a, g.b, g.c, g.d, g.e, g.F, s, t);
    double x = (1.0 / (r2 * r2 * r2));
    o[0] += w * x; o[1] += w * (s * x); o[2] += w * (s * s * x);
    o[3] += w * (s * t * x); o[4] += w * (t * x); o[5] += w * (t * t * x);
}
// force integrand contributions of node (i, j) (Termset2/3/4 bookkeeping)
__device__ __forceinline__ void simpson_f_node(const SimpsonGeo &g, int i, int j, int N, double o[6]) {
    const double ds = 1.0 / ((double)N);
    double s = (double)(i + 1) * ds, t = (double)(j + 1) * ds;
    double fs = (i % 2 == 0) ? 2.0 : 4.0, ft = (j % 2 == 0) ? 2.0 : 4.0;
    if (j == 0 && i > 0 && i < N) { fint6(g, s, 0.0, fs, o); fint6(g, s, 1.0, fs, o); }
    if (j > 0 && j < N) {
        if (i == 0) { fint6(g, 0.0, t, ft, o); fint6(g, 1.0, t, ft, o); }
        else if (i < N) fint6(g, s, t, fs * ft, o);
    }
}
__device__ __forceinline__ void simpson_f_assemble(const SimpsonGeo &g, const double T[6], double kr, int N, d3 f[4]) {
    const double ds = 1.0 / ((double)N);
    double in[6];
    for (int q = 0; q < 6; q++) in[q] = (ds * ds / 9.0) * T[q];
    const double A[3] = {g.vA.x, g.vA.y, g.vA.z}, B[3] = {g.vB.x, g.vB.y, g.vB.z}, Cv[3] = {g.vC.x, g.vC.y, g.vC.z};
    double o[4][3];
    for (int k = 0; k < 3; k++) {
        o[0][k] = 4.0 * kr * (-in[0] * Cv[k] + in[1] * (Cv[k] - A[k]) + in[2] * A[k] + in[4] * B[k] - in[3] * B[k]);
        o[1][k] = 4.0 * kr * (-in[1] * Cv[k] - in[2] * A[k] + in[3] * B[k]);
        o[2][k] = 4.0 * kr * (in[0] * Cv[k] + in[1] * A[k] - in[4] * (B[k] + Cv[k]) - in[3] * A[k] + in[5] * B[k]);
        o[3][k] = 4.0 * kr * (in[4] * Cv[k] + in[3] * A[k] - in[5] * B[k]);
    }
    for (int q = 0; q < 4; q++) f[q] = d3{o[q][0], o[q][1], o[q][2]};
}

}  // namespace mdy
