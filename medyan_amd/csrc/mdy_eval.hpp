// mdy_eval.hpp -- evaluation of all interactions over the unified index space,
// per-bead force gather, exact block/grid reductions and the quadrature
// fallbacks. Used both by the one-shot kernels (mdy_eval.hip) and by the
// persistent conjugate-gradient kernel (mdy_cg.hip).
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

// one functor for all three (uniform runtime mode): a single inlined copy of
// the interaction code serves every evaluation of the persistent kernel
struct PosAny {
    enum { X = 0, LINE = 1, SPEC = 2 };
    const double *__restrict__ x;
    const double *__restrict__ d;
    const double *__restrict__ F;
    double lam, beta;
    int mode;
    __device__ __forceinline__ double c(int o) const {
        if (mode == X) return x[o];
        const double s = (mode == SPEC) ? F[o] + beta * d[o] : d[o];
        return x[o] + lam * s;
    }
    __device__ __forceinline__ d3 operator()(int o) const { return d3{c(o), c(o + 1), c(o + 2)}; }
    static __device__ PosAny at(const double *x) { return PosAny{x, nullptr, nullptr, 0.0, 0.0, X}; }
    static __device__ PosAny line(const double *x, const double *d, double lam) {
        return PosAny{x, d, nullptr, lam, 0.0, LINE};
    }
    static __device__ PosAny spec(const double *x, const double *F, const double *d, double beta, double lam) {
        return PosAny{x, d, F, lam, beta, SPEC};
    }
};

__device__ __forceinline__ void set_err(int *err, int code) { atomicCAS(err, 0, code); }

__device__ __forceinline__ void enqueue(int *q, int *qn, int v, int *err) {
    int s = atomicAdd(qn, 1);
    if (s < kQueueCap) q[s] = v;
    else set_err(err, MDY_ERR_CAPACITY);
}

__device__ __forceinline__ void put3(double *__restrict__ c, int slot, d3 v) {
    c[3 * slot] = v.x; c[3 * slot + 1] = v.y; c[3 * slot + 2] = v.z;
}
// mixed-precision force records: 3 floats per slot
__device__ __forceinline__ void put3(float *__restrict__ c, int slot, d3 v) {
    c[3 * slot] = (float)v.x; c[3 * slot + 1] = (float)v.y; c[3 * slot + 2] = (float)v.z;
}
__device__ __forceinline__ void put3(float *__restrict__ c, int slot, f3 v) {
    c[3 * slot] = v.x; c[3 * slot + 1] = v.y; c[3 * slot + 2] = v.z;
}
// force-record element type of each form
template <bool MIXED> struct SlotT { using T = double; };
template <> struct SlotT<true> { using T = float; };

// ---- the interaction sweep -------------------------------------------------
// Every interaction is an item of the index space [EV | FS | FB | LM | BS | BB |
// BD]. An item's record (its bead offsets and parameters) and its beads'
// coordinates are fetched one item ahead of the one being computed, so each
// thread's chain of dependent memory accesses overlaps with its arithmetic.
// IT_EVM: an excluded-volume pair of the monomer-sum form (same table and slots as IT_EV)
enum { IT_EV = 0, IT_FS, IT_FB, IT_LM, IT_BS, IT_BB, IT_BDH, IT_BP, IT_BD, IT_EVM, IT_NONE };
struct ItemRec {
    int type, j;      // kind and index within its table
    int4 b;           // bead offsets (x3); unused entries repeat the first
    double p0, p1, p2, p3;
};
__device__ __forceinline__ ItemRec fetch_rec(const Dev &D, int i) {
    ItemRec r;
    r.p0 = r.p1 = r.p2 = r.p3 = 0.0;
    if (i < D.n_ev) {
        r.type = D.ev_mon ? IT_EVM : IT_EV; r.j = i;
        r.b = D.ev_bs[i];
        r.p0 = D.ev_kr[i];
    } else if ((i -= D.n_ev) < D.n_fs) {
        r.type = IT_FS; r.j = i;
        const int2 b = D.fs_bs[i];
        r.b = make_int4(b.x, b.y, b.x, b.x);
        r.p0 = D.fs_k[i]; r.p1 = D.fs_l0[i];
    } else if ((i -= D.n_fs) < D.n_fb) {
        r.type = IT_FB; r.j = i;
        const int *b = D.fb_bs + 3 * i;
        r.b = make_int4(b[0], b[1], b[2], b[0]);
        r.p0 = D.fb_k[i]; r.p1 = D.fb_th[i]; r.p2 = D.fb_ct[i]; r.p3 = D.fb_st[i];
    } else if ((i -= D.n_fb) < D.n_lm) {
        r.type = IT_LM; r.j = i;
        r.b = D.lm_bs[i];
        r.p0 = D.lm_k[i]; r.p1 = D.lm_l0[i]; r.p2 = D.lm_p1[i]; r.p3 = D.lm_p2[i];
    } else if ((i -= D.n_lm) < D.n_bs) {
        r.type = IT_BS; r.j = i;
        const int *b = D.bs_bs + 3 * i;
        r.b = make_int4(b[0], b[1], b[2], b[0]);
        r.p0 = D.bs_k[i]; r.p1 = D.bs_l0[i]; r.p2 = D.bs_p[i];
    } else if ((i -= D.n_bs) < D.n_bb) {
        r.type = IT_BB; r.j = i;
        r.b = D.bb_bs[i];
        r.p0 = D.bb_k[i]; r.p1 = D.bb_th[i]; r.p2 = D.bb_ct[i]; r.p3 = D.bb_st[i];
    } else if ((i -= D.n_bb) < D.n_bdh) {
        r.type = IT_BDH; r.j = i;
        r.b = D.bdh_bs[i];
        r.p0 = D.bdh_k[i]; r.p1 = D.bdh_pos[i];
    } else if ((i -= D.n_bdh) < D.n_bp) {
        r.type = IT_BP; r.j = i;
        const int *b = D.bp_bs + 3 * i;
        r.b = make_int4(b[0], b[1], b[2], b[0]);
        r.p0 = D.bp_k[i]; r.p1 = D.bp_pos[i];
    } else if ((i -= D.n_bp) < D.n_bd) {
        r.type = IT_BD; r.j = i;
        const int bead = D.bd_bead[i];
        r.b = make_int4(bead, D.bd_plane[i], bead, bead);
        r.p0 = D.bd_k[i]; r.p1 = D.bd_sl[i];
    } else {
        r.type = IT_NONE; r.j = 0;
        r.b = make_int4(0, 0, 0, 0);
    }
    return r;
}
__device__ __forceinline__ int rec_nbeads(int type) {
    return type == IT_EV || type == IT_EVM || type == IT_LM || type == IT_BB || type == IT_BDH ? 4
         : type == IT_FB || type == IT_BS || type == IT_BP ? 3
         : type == IT_FS ? 2 : type == IT_BD ? 1 : 0;
}
// bead k's position (the unused entries of a record repeat its first bead)
__device__ __forceinline__ d3 rec_pos(const PosAny &P, const ItemRec &r, int k) {
    const int o = k == 0 ? r.b.x : k == 1 ? (r.type == IT_BD ? r.b.x : r.b.y) : k == 2 ? r.b.z : r.b.w;
    return P(o);
}

// ---- one interaction from its record and positions: energy (into the exact
// total and, if PER_FF, into the per-force-field double sums) and, if FORCES,
// its per-bead force slots. One accumulation site and one copy of the
// excluded-volume invariants and force keep the inlined code small (it runs
// from the instruction cache of every CU).
// Mixed-precision form: a pair's force runs in FP32 only when the squared
// normalized triple product q = -JJ / (a b c) of the two cylinders is at
// least this; nearer-coplanar pairs amplify FP32 rounding as ~6e-8 / q (at
// 0.03 the FP32 pair force stays within ~6e-5 of its FP64 value, about 3/4 of
// the pairs of a random network qualify).
constexpr double kMixedMinQ = 0.03;

// QUAD: the pairs flagged in D.evq are skipped (the quadrature kernels evaluate them).
template <bool MIXED = false, bool QUAD = false>
__device__ __forceinline__ void item_compute(const Dev &D, const ItemRec &r, const d3 c[4], double acc[kNFF],
                                             mdy_xsum &xs, typename SlotT<MIXED>::T *__restrict__ contrib,
                                             const bool PER_FF, const bool FORCES) {
    double U = 0.0;
    int ff = 0;
    bool add = true;
    const int j = r.j;
    if (QUAD && r.type == IT_EV && D.evq[j]) {
        add = false;
    } else if (r.type == IT_EV) {
        const double kr = r.p0;
        const d3 c1 = c[0], c2 = c[1], c3 = c[2], c4 = c[3];
        ff = MDY_FF_EV;
        EvInv v;
        d3 c2e = c2;
        // pass 0: the pair as given (energy, and force unless coplanar);
        // pass 1: the coplanar perturbation of the reference force path,
        // c2 moved off the plane (CylinderExclVolRepulsion.cpp:610-626)
#pragma clang loop unroll(disable)
        for (int pass = 0; pass < 2; pass++) {
            ev_invariants(c1, c2e, c3, c4, v);
            if (pass == 0) {
                U = ev_energy_inv(v, kr);
                if (!FORCES || !(D.coplanar_mode == MDY_COPLANAR_REFERENCE && are_in_plane(c1, c2, c3, c4))) break;
                c2e = moved_off_plane(c1, c2, c3, c2, 0.01);
            }
        }
        if (ev_energy_bad(U)) {
            enqueue(D.qe, D.qe_n, j, D.err);
            add = false;
        }
        if (FORCES) {
            bool done = false;
            if constexpr (MIXED) {
                if (-v.JJ / (v.a * v.b * v.c) >= kMixedMinQ) {
                    // well-conditioned pair: derivative blocks in FP32
                    f3 f[4];
                    ev_force_inv_t<float>(v, c1, c2e, c3, c4, kr, f);
                    if (finite_d3(f[0]) && finite_d3(f[1]) && finite_d3(f[2]) && finite_d3(f[3])) {
                        int s = D.slot_ev + 4 * j;
                        put3(contrib, s, f[0]); put3(contrib, s + 1, f[1]);
                        put3(contrib, s + 2, f[2]); put3(contrib, s + 3, f[3]);
                        done = true;
                    }
                }
            }
            if (!done) {
                d3 f[4];
                ev_force_inv(v, c1, c2e, c3, c4, kr, f);
                if (!(finite_d3(f[0]) && finite_d3(f[1]) && finite_d3(f[2]) && finite_d3(f[3]))) {
                    enqueue(D.qf, D.qf_n, j, D.err);
                } else {
                    int s = D.slot_ev + 4 * j;
                    put3(contrib, s, f[0]); put3(contrib, s + 1, f[1]); put3(contrib, s + 2, f[2]); put3(contrib, s + 3, f[3]);
                }
            }
        }
    } else if (r.type == IT_EVM) {   // monomer-pair sum, in FP64 in both forms; no fallbacks
        ff = MDY_FF_EV;
        const int2 cc = D.ev_cyl[j];
        const int4 m1 = D.cyl_mon[cc.x], m2 = D.cyl_mon[cc.y];
        d3 f[4];
        if (FORCES) {
            U = evm_eval<double, true>(c[0], c[1], c[2], c[3], m1, m2, r.p0, f);
            const int s = D.slot_ev + 4 * j;
            put3(contrib, s, f[0]); put3(contrib, s + 1, f[1]); put3(contrib, s + 2, f[2]); put3(contrib, s + 3, f[3]);
        } else {
            U = evm_eval<double, false>(c[0], c[1], c[2], c[3], m1, m2, r.p0, f);
        }
    } else if (r.type == IT_FS) {
        ff = MDY_FF_FS;
        U = fs_energy(c[0], c[1], r.p0, r.p1);
        if (FORCES) {
            d3 f1, f2;
            fs_force(c[0], c[1], r.p0, r.p1, f1, f2);
            int s = D.slot_fs + 2 * j;
            put3(contrib, s, f1); put3(contrib, s + 1, f2);
        }
    } else if (r.type == IT_FB) {
        ff = MDY_FF_FB;
        U = fb_energy(c[0], c[1], c[2], r.p0, r.p1, r.p2, r.p3);
        if (FORCES) {
            d3 f1, f2, f3;
            fb_force(c[0], c[1], c[2], r.p0, r.p1, r.p2, r.p3, f1, f2, f3);
            int s = D.slot_fb + 3 * j;
            put3(contrib, s, f1); put3(contrib, s + 1, f2); put3(contrib, s + 2, f3);
        }
    } else if (r.type == IT_LM) {
        ff = (j < D.n_ls) ? MDY_FF_LS : MDY_FF_MS;
        U = ls_energy(c[0], c[1], c[2], c[3], r.p0, r.p1, r.p2, r.p3);
        if (FORCES) {
            d3 f1, f2, f3, f4;
            ls_force(c[0], c[1], c[2], c[3], r.p0, r.p1, r.p2, r.p3, f1, f2, f3, f4);
            int s = D.slot_lm + 4 * j;
            put3(contrib, s, f1); put3(contrib, s + 1, f2); put3(contrib, s + 2, f3); put3(contrib, s + 3, f4);
        }
    } else if (r.type == IT_BS) {
        ff = MDY_FF_BS;
        U = bs_energy(c[0], c[1], c[2], r.p0, r.p1, r.p2);
        if (FORCES) {
            d3 f1, f2, f3;
            bs_force(c[0], c[1], c[2], r.p0, r.p1, r.p2, f1, f2, f3);
            int s = D.slot_bs + 3 * j;
            put3(contrib, s, f1); put3(contrib, s + 1, f2); put3(contrib, s + 2, f3);
        }
    } else if (r.type == IT_BB) {
        ff = MDY_FF_BB;
        U = bb_energy(c[0], c[1], c[2], c[3], r.p0, r.p1, r.p2, r.p3);
        if (FORCES) {
            d3 f1, f2, f3, f4;
            bb_force(c[0], c[1], c[2], c[3], r.p0, r.p1, r.p2, r.p3, f1, f2, f3, f4);
            int s = D.slot_bb + 4 * j;
            put3(contrib, s, f1); put3(contrib, s + 1, f2); put3(contrib, s + 2, f3); put3(contrib, s + 3, f4);
        }
    } else if (r.type == IT_BDH) {
        ff = MDY_FF_BDH;
        U = D.bdh_quad ? bdhq_energy(c[0], c[1], c[2], c[3], r.p0, r.p1) : bdh_energy(c[0], c[1], c[2], c[3], r.p0, r.p1);
        if (FORCES) {
            d3 f1, f2, f3, f4;
            if (D.bdh_quad) bdhq_force(c[0], c[1], c[2], c[3], r.p0, r.p1, f1, f2, f3, f4);
            else bdh_force(c[0], c[1], c[2], c[3], r.p0, r.p1, f1, f2, f3, f4);
            int s = D.slot_bdh + 4 * j;
            put3(contrib, s, f1); put3(contrib, s + 1, f2); put3(contrib, s + 2, f3); put3(contrib, s + 3, f4);
        }
    } else if (r.type == IT_BP) {
        ff = MDY_FF_BP;
        U = bp_energy(c[0], c[1], c[2], r.p0, r.p1, D.cq, D.sq);
        if (FORCES) {
            d3 f1, f2, f3;
            bp_force(c[0], c[1], c[2], r.p0, r.p1, D.cq, D.sq, f1, f2, f3);
            int s = D.slot_bp + 3 * j;
            put3(contrib, s, f1); put3(contrib, s + 1, f2); put3(contrib, s + 2, f3);
        }
    } else if (r.type == IT_BD) {
        const double *pl = D.planes + 4 * r.b.y;
        ff = MDY_FF_BD;
        U = bd_energy(pl, c[0], r.p0, r.p1);
        if (FORCES) put3(contrib, D.slot_bd + j, bd_force(pl, c[0], r.p0, r.p1));
    } else {
        add = false;
    }
    if (add) {
        mdy_xs_add(&xs, U);
        if (PER_FF) {
#pragma unroll
            for (int k = 0; k < kNFF; k++) acc[k] += (k == ff) ? U : 0.0;
        }
    }
}

// All items i = start, start + stride, ... below D.n_items() at positions P.
// (Fetching the next record ahead of the arithmetic was measured slower: the
// extra live registers cost more than the memory round trip they hide.)
template <bool MIXED = false>
__device__ __forceinline__ void sweep(const Dev &D, const PosAny &P, int start, int stride, double acc[kNFF],
                                      mdy_xsum &xs, typename SlotT<MIXED>::T *__restrict__ contrib,
                                      const bool PER_FF, const bool FORCES) {
    const int total = D.n_items();
    for (int i = start; i < total; i += stride) {
        const ItemRec r = fetch_rec(D, i);
        const int nb = rec_nbeads(r.type);
        d3 c[4];
        c[0] = rec_pos(P, r, 0);
        c[1] = nb > 1 ? rec_pos(P, r, 1) : c[0];
        c[2] = nb > 2 ? rec_pos(P, r, 2) : c[0];
        c[3] = nb > 3 ? rec_pos(P, r, 3) : c[0];
        item_compute<MIXED>(D, r, c, acc, xs, contrib, PER_FF, FORCES);
    }
}

// single item (one-shot kernels)
template <bool MIXED = false, bool QUAD = false, class P>
__device__ __forceinline__ void item(const Dev &D, int i, const P &pos, double acc[kNFF], mdy_xsum &xs,
                                     typename SlotT<MIXED>::T *__restrict__ contrib, const bool PER_FF,
                                     const bool FORCES) {
    ItemRec r = fetch_rec(D, i);
    const int nb = rec_nbeads(r.type);
    d3 c[4];
    c[0] = pos(r.b.x);
    c[1] = nb > 1 ? pos(r.b.y) : c[0];
    c[2] = nb > 2 ? pos(r.b.z) : c[0];
    c[3] = nb > 3 ? pos(r.b.w) : c[0];
    item_compute<MIXED, QUAD>(D, r, c, acc, xs, contrib, PER_FF, FORCES);
}

// Slots k0..k1-1 of a bead's row (after the ones loaded with the row's
// first), added in order. CHUNKED: kTail indices, then their kTail values, at
// a time, so a bead with many slots (up to ~60 in a 200k-cylinder network)
// costs two dependent round trips per kTail slots rather than per slot;
// otherwise one slot at a time (fewer instructions where rows rarely go past
// their first loads).
constexpr int kTail = 8;
template <bool CHUNKED, class S>
__device__ __forceinline__ void gather_tail(const int *__restrict__ idx, int k0, int k1, const S *__restrict__ contrib,
                                            double &fx, double &fy, double &fz) {
    if (!CHUNKED) {
        for (int k = k0; k < k1; k++) {
            const int s = idx[k];
            fx += contrib[3 * s]; fy += contrib[3 * s + 1]; fz += contrib[3 * s + 2];
        }
        return;
    }
#pragma clang loop unroll(disable)
    for (int k = k0; k < k1; k += kTail) {
        int ix[kTail];
#pragma unroll
        for (int j = 0; j < kTail; j++) ix[j] = (k + j < k1) ? idx[k + j] : 0;
        S v[kTail][3];
#pragma unroll
        for (int j = 0; j < kTail; j++)
            if (k + j < k1) { v[j][0] = contrib[3 * ix[j]]; v[j][1] = contrib[3 * ix[j] + 1]; v[j][2] = contrib[3 * ix[j] + 2]; }
#pragma unroll
        for (int j = 0; j < kTail; j++)
            if (k + j < k1) { fx += v[j][0]; fy += v[j][1]; fz += v[j][2]; }
    }
}
// Per-bead force: the slots of bead b in reference order, starting from 0.0
// (ForceFieldManager::computeForces zero-fills, then each force field adds).
// The first kGS static and kGD dynamic slot indices and values are loaded
// together before any addition, so a bead costs three dependent memory round
// trips (offsets, indices, values) instead of two per slot; slots beyond
// those follow in chunks in the narrow form (gather_tail). Every phase waits
// for its slowest bead, so when a
// thread gathers one bead (the minimizer on networks with fewer beads than
// threads) the wide form (12 static, 24 dynamic) covers the beads with the
// most slots; with several beads per thread the narrow form's fewer
// instructions per bead win.
template <int kGS, int kGD, class S>
__device__ __forceinline__ d3 gather_bead_w(const Dev &D, const S *__restrict__ contrib, int b) {
    const int s0 = D.st_off[b], s1 = D.st_off[b + 1], d0 = D.dy_off[b], d1 = D.dy_off[b + 1];
    int is[kGS], id[kGD];
#pragma unroll
    for (int j = 0; j < kGS; j++) is[j] = (s0 + j < s1) ? D.st_idx[s0 + j] : 0;
#pragma unroll
    for (int j = 0; j < kGD; j++) id[j] = (d0 + j < d1) ? D.dy_idx[d0 + j] : 0;
    S vs[kGS][3], vd[kGD][3];
#pragma unroll
    for (int j = 0; j < kGS; j++) {
        if (s0 + j < s1) { vs[j][0] = contrib[3 * is[j]]; vs[j][1] = contrib[3 * is[j] + 1]; vs[j][2] = contrib[3 * is[j] + 2]; }
    }
#pragma unroll
    for (int j = 0; j < kGD; j++) {
        if (d0 + j < d1) { vd[j][0] = contrib[3 * id[j]]; vd[j][1] = contrib[3 * id[j] + 1]; vd[j][2] = contrib[3 * id[j] + 2]; }
    }
    double fx = 0.0, fy = 0.0, fz = 0.0;
#pragma unroll
    for (int j = 0; j < kGS; j++)
        if (s0 + j < s1) { fx += vs[j][0]; fy += vs[j][1]; fz += vs[j][2]; }
    gather_tail<(kGS <= 8)>(D.st_idx, s0 + kGS, s1, contrib, fx, fy, fz);
#pragma unroll
    for (int j = 0; j < kGD; j++)
        if (d0 + j < d1) { fx += vd[j][0]; fy += vd[j][1]; fz += vd[j][2]; }
    gather_tail<(kGS <= 8)>(D.dy_idx, d0 + kGD, d1, contrib, fx, fy, fz);
    return d3{fx, fy, fz};
}
template <class S>
__device__ __forceinline__ d3 gather_bead(const Dev &D, const S *__restrict__ contrib, int b) {
    return gather_bead_w<8, 8>(D, contrib, b);
}

// ---- block reductions (fixed butterfly + fixed wave order: deterministic) ----
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
__device__ __forceinline__ long long wave_sum_i64(long long v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ int wave_or(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
    return v;
}
// N doubles per thread -> N block results (all threads). Entries whose bit is
// set in max_mask / min_mask are max / min reduced, the rest summed.
// lds: >= 17 * N doubles.
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
// Exact block sum of per-thread accumulators (integer digit sums are
// associative, so any tree gives the same result). lds: >= 17 * 8 int64.
__device__ __forceinline__ void block_reduce_xsum(mdy_xsum &s, long long *lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    mdy_xs_norm(&s);
#pragma unroll
    for (int k = 0; k < MDY_XS_N; k++) {
        long long r = wave_sum_i64(s.d[k]);
        if (lane == 0) lds[wave * 8 + k] = r;
    }
    int f = wave_or(s.flags);
    if (lane == 0) lds[wave * 8 + 7] = f;
    __syncthreads();
    if (threadIdx.x < 8) {
        int k = threadIdx.x;
        long long r = lds[k];
        for (int w = 1; w < nw; w++) r = (k == 7) ? (r | lds[w * 8 + k]) : r + lds[w * 8 + k];
        lds[128 + k] = r;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < MDY_XS_N; k++) s.d[k] = lds[128 + k];
    s.flags = (int)lds[128 + 7];
    mdy_xs_norm(&s);
    __syncthreads();
}
// Two block sums in one pass (one set of barriers, interleaved shuffles).
// lds: >= 17 * 16 int64.
__device__ __forceinline__ void block_reduce_xsum2(mdy_xsum &s, mdy_xsum &t, long long *lds) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nw = blockDim.x >> 6;
    mdy_xs_norm(&s);
    mdy_xs_norm(&t);
#pragma unroll
    for (int k = 0; k < MDY_XS_N; k++) {
        const long long r = wave_sum_i64(s.d[k]), q = wave_sum_i64(t.d[k]);
        if (lane == 0) { lds[wave * 16 + k] = r; lds[wave * 16 + 8 + k] = q; }
    }
    const int f = wave_or(s.flags), g = wave_or(t.flags);
    if (lane == 0) { lds[wave * 16 + 7] = f; lds[wave * 16 + 15] = g; }
    __syncthreads();
    if (threadIdx.x < 16) {
        const int k = threadIdx.x;
        long long r = lds[k];
        for (int w = 1; w < nw; w++) r = ((k & 7) == 7) ? (r | lds[w * 16 + k]) : r + lds[w * 16 + k];
        lds[256 + k] = r;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < MDY_XS_N; k++) { s.d[k] = lds[256 + k]; t.d[k] = lds[256 + 8 + k]; }
    s.flags = (int)lds[256 + 7];
    t.flags = (int)lds[256 + 15];
    mdy_xs_norm(&s);
    mdy_xs_norm(&t);
    __syncthreads();
}
// Block-wide combination of NW 64-bit words per thread through LDS: word k
// is summed, or OR-ed where bit k of or_mask is set, or unsigned-max-ed where
// bit k of max_mask is set. Every thread writes its words; one thread per
// 16-entry chunk combines it; one thread per word combines the chunks; every
// thread reads the results. Exact for integers, so the result does not depend
// on the order, and it takes three short LDS round trips instead of the
// dependent shuffle chains of a butterfly. blockDim.x must be a multiple of 16.
// lds: >= block_combine_words(NW) 64-bit words.
__host__ __device__ constexpr int block_combine_words(int nw, int nthreads = kBlock) {
    return nw * (nthreads + nthreads / 16 + 1);
}
template <int NW>
__device__ __forceinline__ void block_combine(long long w[NW], long long *lds, unsigned or_mask, unsigned max_mask) {
    const int n = blockDim.x, t = threadIdx.x, nch = n >> 4;
    long long *part = lds + NW * n;
    long long *res = part + NW * nch;
#pragma unroll
    for (int k = 0; k < NW; k++) lds[k * n + t] = w[k];
    __syncthreads();
    for (int job = t; job < NW * nch; job += n) {
        const int k = job / nch, c = job - k * nch;
        const bool isor = (or_mask >> k) & 1u, ismax = (max_mask >> k) & 1u;
        const long long *src = lds + k * n + c * 16;
        long long r = src[0];
#pragma unroll
        for (int j = 1; j < 16; j++) {
            const long long v = src[j];
            r = isor ? (r | v) : ismax ? ((unsigned long long)v > (unsigned long long)r ? v : r) : r + v;
        }
        part[k * nch + c] = r;
    }
    __syncthreads();
    if (t < NW) {
        const bool isor = (or_mask >> t) & 1u, ismax = (max_mask >> t) & 1u;
        long long r = part[t * nch];
        for (int c = 1; c < nch; c++) {
            const long long v = part[t * nch + c];
            r = isor ? (r | v) : ismax ? ((unsigned long long)v > (unsigned long long)r ? v : r) : r + v;
        }
        res[t] = r;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < NW; k++) w[k] = res[k];
    __syncthreads();
}
// Exact block sums of two per-thread accumulators through block_combine;
// with_max: also the maximum of a non-negative double m (all threads get it).
__device__ __forceinline__ void block_reduce_xsum2_lds(mdy_xsum &s, mdy_xsum &t, long long *lds, double *m = nullptr) {
    mdy_xs_norm(&s);   // digits below 2^48: a block's sum of them cannot overflow
    mdy_xs_norm(&t);
    long long w[17];
#pragma unroll
    for (int k = 0; k < MDY_XS_N; k++) { w[k] = s.d[k]; w[8 + k] = t.d[k]; }
    w[7] = s.flags;
    w[15] = t.flags;
    w[16] = m ? __double_as_longlong(*m) : 0;
    if (m) block_combine<17>(w, lds, (1u << 7) | (1u << 15), 1u << 16);
    else block_combine<16>(w, lds, (1u << 7) | (1u << 15), 0u);
#pragma unroll
    for (int k = 0; k < MDY_XS_N; k++) { s.d[k] = w[k]; t.d[k] = w[8 + k]; }
    s.flags = (int)w[7];
    t.flags = (int)w[15];
    s.cnt = 0;
    t.cnt = 0;
    mdy_xs_norm(&s);
    mdy_xs_norm(&t);
    if (m) *m = __longlong_as_double(w[16]);
}
// per-block partial records: [blk * kRed + off .. off+7] hold a normalized accumulator
__device__ __forceinline__ void store_xsum(double *red, int blk, int off, const mdy_xsum &s) {
    long long *r = reinterpret_cast<long long *>(red) + (size_t)blk * kRed + off;
#pragma unroll
    for (int k = 0; k < MDY_XS_N; k++) r[k] = s.d[k];
    r[7] = s.flags;
}
__device__ __forceinline__ void grid_reduce_xsum(const double *red, int nblk, int off, mdy_xsum &s, long long *lds) {
    mdy_xs_zero(&s);
    const long long *r = reinterpret_cast<const long long *>(red);
    for (int b = threadIdx.x; b < nblk; b += blockDim.x) {
        const long long *q = r + (size_t)b * kRed + off;
#pragma unroll
        for (int k = 0; k < MDY_XS_N; k++) s.d[k] += __builtin_nontemporal_load(q + k);
        s.flags |= (int)__builtin_nontemporal_load(q + 7);
    }
    block_reduce_xsum(s, lds);
}

// Reduce per-block partials red[blk * kRed + off + k] for k < N over nblk
// blocks in a fixed order; every calling block obtains bit-identical results.
template <int N>
__device__ __forceinline__ void grid_partials_reduce(const double *red, int nblk, double v[N], double *lds,
                                                     unsigned max_mask = 0, unsigned min_mask = 0, int off = 0) {
#pragma unroll
    for (int k = 0; k < N; k++) v[k] = ((max_mask >> k) & 1) ? -INFINITY : ((min_mask >> k) & 1) ? INFINITY : 0.0;
    for (int blk = threadIdx.x; blk < nblk; blk += blockDim.x) {
#pragma unroll
        for (int k = 0; k < N; k++) {
            double o = __builtin_nontemporal_load(red + (size_t)blk * kRed + off + k);
            v[k] = ((max_mask >> k) & 1) ? fmax(v[k], o) : ((min_mask >> k) & 1) ? fmin(v[k], o) : v[k] + o;
        }
    }
    block_reduce<N>(v, lds, max_mask, min_mask);
}

// Energy epilogue, shared by the one-shot and persistent kernels. X holds the
// exact total of the analytic energies; v (per_ff) the per-force-field double
// sums. Quadrature-fallback energies (queue qe, values qe_u) are added to
// both, in increasing pair order. e[0..7] per FF (reported values), e[8] =
// computeEnergy total (+inf when a force field returned its -1 sentinel),
// e[9] = the sum without the sentinel.
constexpr int kXsE = 16;   // exact energy total in a reduction record (after the kNFF per-FF sums)
constexpr int kXsA = 24;   // second exact accumulator in a reduction record
static_assert(kNFF <= kXsE && kXsA + 8 <= kRed, "reduction record layout");
__device__ __forceinline__ void energy_finish(mdy_xsum X, const double v_in[kNFF], int nq, const int *qe,
                                              const double *qe_u, double e[kNFF + 2]) {
    double v[kNFF];
    for (int k = 0; k < kNFF; k++) v[k] = v_in ? v_in[k] : 0.0;
    nq = min(nq, kQueueCap);
    bool bad = false;
    if (nq > 0) {
        double add = 0.0;
        int last = -1;
        for (int r = 0; r < nq; r++) {
            int best = -1, bi = 0x7fffffff;
            for (int q = 0; q < nq; q++) {
                int pi = __hip_atomic_load(qe + q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (pi > last && pi < bi) { bi = pi; best = q; }
            }
            if (best < 0) break;
            last = bi;
            double U = qe_u[best];
            if (ev_energy_bad(U)) bad = true;
            else mdy_xs_add(&X, U);
            add += U;
        }
        v[MDY_FF_EV] = bad ? -1.0 : v[MDY_FF_EV] + add;
    }
    for (int k = 0; k < kNFF; k++) e[k] = v[k];
    const double tot = mdy_xs_value(&X);
    e[kNFF] = bad ? INFINITY : tot;
    e[kNFF + 1] = bad ? tot - 1.0 : tot;
}
// the same from per-block records: per-FF doubles at [0, kNFF), exact total at kXsE
__device__ __forceinline__ void energy_epilogue(const Dev &D, const double *red, int nblk, int nq, const int *qe,
                                                const double *qe_u, double *lds, double e[kNFF + 2], bool per_ff) {
    double v[kNFF];
    if (per_ff) grid_partials_reduce<kNFF>(red, nblk, v, lds);
    mdy_xsum X;
    grid_reduce_xsum(red, nblk, kXsE, X, reinterpret_cast<long long *>(lds));
    energy_finish(X, per_ff ? v : nullptr, nq, qe, qe_u, e);
}

// ---- Simpson quadrature fallbacks (CylinderExclVolRepulsion.cpp:884-1254) ----
// The reference advances its abscissae before first use, so interior node i
// sits at (i+1)/N; node values are formed here as (i+1)*(1/N). Sums are exact
// (mdy_xsum), like the CPU path in exact mode.
struct SimpsonGeo { double a, b, c, d, e, F; d3 vA, vB, vC; };
__device__ __forceinline__ SimpsonGeo simpson_geo(d3 c1, d3 c2, d3 c3, d3 c4) {
    SimpsonGeo g;
    g.a = sprod(c1, c2, c1, c2); g.b = sprod(c3, c4, c3, c4); g.c = sprod(c3, c1, c3, c1);
    g.d = sprod(c1, c2, c3, c4); g.e = sprod(c1, c2, c3, c1); g.F = sprod(c3, c4, c3, c1);
    g.vA = sub(c1, c2); g.vB = sub(c3, c4); g.vC = sub(c3, c1);
    return g;
}
__device__ __forceinline__ double eint(const SimpsonGeo &g, double s, double t) {
    double r2 = ev_r2(g.a, g.b, g.c, g.d, g.e, g.F, s, t);
    return (1.0 / (r2 * r2));
}
// energy: terms that node (i, j) adds to Termset2/3/4 (CylinderExclVolRepulsion.cpp:923-958)
__device__ __forceinline__ void simpson_e_node(const SimpsonGeo &g, int i, int j, int N, mdy_xsum &X) {
    const double ds = 1.0 / ((double)N);
    double s = (double)(i + 1) * ds, t = (double)(j + 1) * ds;
    double fs = (i % 2 == 0) ? 2.0 : 4.0, ft = (j % 2 == 0) ? 2.0 : 4.0;
    double v = 0.0;
    if (j == 0 && i > 0 && i < N) v = fs * (eint(g, s, 0.0) + eint(g, s, 1.0));
    if (j > 0 && j < N) {
        if (i == 0) v = ft * (eint(g, 0.0, t) + eint(g, 1.0, t));
        else if (i < N) v = fs * ft * eint(g, s, t);
    }
    mdy_xs_add(&X, v);   // adding an exact 0.0 changes nothing
}
__device__ __forceinline__ double simpson_e_corners(const SimpsonGeo &g) {
    return eint(g, 0.0, 0.0) + eint(g, 0.0, 1.0) + eint(g, 1.0, 0.0) + eint(g, 1.0, 1.0);
}
// force: component q of the six integrands {1, s, s^2, st, t, t^2}/r^6
__device__ __forceinline__ double fint_q(const SimpsonGeo &g, double s, double t, int q) {
    double r2 = ev_r2(g.a, g.b, g.c, g.d, g.e, g.F, s, t);
    double x = (1.0 / (r2 * r2 * r2));
    return q == 0 ? x : q == 1 ? s * x : q == 2 ? s * s * x : q == 3 ? s * t * x : q == 4 ? t * x : t * t * x;
}
__device__ __forceinline__ void simpson_f_node(const SimpsonGeo &g, int i, int j, int N, int q, mdy_xsum &X) {
    const double ds = 1.0 / ((double)N);
    double s = (double)(i + 1) * ds, t = (double)(j + 1) * ds;
    double fs = (i % 2 == 0) ? 2.0 : 4.0, ft = (j % 2 == 0) ? 2.0 : 4.0;
    double v1 = 0.0, v2 = 0.0;   // the node's terms (an exact 0.0 adds nothing)
    if (j == 0 && i > 0 && i < N) {
        v1 = fs * fint_q(g, s, 0.0, q);
        v2 = fs * fint_q(g, s, 1.0, q);
    }
    if (j > 0 && j < N) {
        if (i == 0) {
            v1 = ft * fint_q(g, 0.0, t, q);
            v2 = ft * fint_q(g, 1.0, t, q);
        } else if (i < N) {
            v1 = fs * ft * fint_q(g, s, t, q);
        }
    }
    mdy_xs_add(&X, v1);
    mdy_xs_add(&X, v2);
}
__device__ __forceinline__ void simpson_f_corners(const SimpsonGeo &g, int q, mdy_xsum &X) {
    mdy_xs_add(&X, fint_q(g, 0.0, 0.0, q));
    mdy_xs_add(&X, fint_q(g, 0.0, 1.0, q));
    mdy_xs_add(&X, fint_q(g, 1.0, 0.0, q));
    mdy_xs_add(&X, fint_q(g, 1.0, 1.0, q));
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
