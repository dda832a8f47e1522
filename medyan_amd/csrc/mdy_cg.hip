// mdy_cg.hip -- CGMethod::minimize<method> with backtracking line search,
// resident on the device for the whole minimization
// (src/Mechanics/Minimizer/CGMethod.hpp:238-373, 826-1743; CGMethod.cpp:43-360).
//
// One cooperative launch; every block runs the same control flow and derives
// every decision from grid reductions that all blocks evaluate in the same
// fixed order (exact sums), so decisions are identical without a broadcast.
// Phases are separated by a flag-array grid barrier.
//
// Every line-search trial evaluates energy AND force slots at the trial point,
// so an accepted trial needs no second pass over the interactions. An
// accepted iteration costs two barriers:
//   [gather F at the accepted point, x += lam*d, max|F|, F.F, F.Fprev]
//   | [beta, s = F + beta*d, s.F, energy+forces at x + lam0*s (the next
//      line search's first trial)]
// The energy at the accepted point is the trial energy of the same
// coordinates (bitwise), so E(x) is never recomputed.
#include "mdy_eval.hpp"
#include "mdy_launch.hpp"

#include <cstring>

namespace mdy {

struct CgArgs {
    Dev D;
    double *x, *F0, *F1, *S0, *S1, *xmin, *contrib;
    double *red;          // 3 * nblk * kRed
    double *qe_u;         // 3 * kQueueCap
    double *spart;        // 6 * nblk * kRed
    int *qe, *qf, *qn;    // 3 * kQueueCap each; qn: 3 x 4 counters
    unsigned long long *bar;  // [0] abort flag; block b arrival flag at [64 + 16 b]
    double *out;          // result record
    long long *prof;      // phase profile: timestamp at each barrier exit (block 0)
    int *prof_kind;       // phase kind code
    long long prof_cap;
    double *tr_lam, *tr_e, *tr_mf, *tr_beta;
    int *tr_safe;
    long long tr_cap;
    // parameters
    int method, try_recover, energy_call_limit, force_call_limit;
    double max_dist, lambda_max, grad_tol, energy_rel_tol;
    long long spin_limit;
};

// result record layout (doubles)
enum {
    R_EB = 0,              // energies before [kNFF] + total
    R_EA = R_EB + kNFF + 1,  // energies after
    R_FLAGS = R_EA + kNFF + 1,
    R_NEC, R_NFC, R_NIT, R_GIN, R_ERC, R_NEE, R_NSM, R_NREC, R_STATUS, R_TRN, R_LAST, R_FOUT, R_NPH,
    R_SIZE
};

__device__ __forceinline__ int ld_i(const int *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// phase kinds recorded by the profile
enum { PK_ENERGY = 1, PK_EFALLBACK, PK_EVAL, PK_FFALLBACK, PK_GATHER, PK_REDUCE, PK_RESTORE };

// result of one evaluation: e[0..7] per FF, e[8] computeEnergy total
// (+inf sentinel), e[9] plain total; the force-fallback queue it filled
struct Eval {
    double e[kNFF + 2];
    double xa;     // extra exact accumulator (s.F), if requested
    int nqf;       // queued force fallbacks
    int qb;        // queue buffer (0..2)
};

struct Cg {
    const CgArgs &A;
    double *lds;
    int G, blk, tid, gtid, gs;
    int ph = 0;                  // phase counter (uniform)
    bool aborted = false;
    int status = MDY_OK;

    __device__ Cg(const CgArgs &a, double *l) : A(a), lds(l) {
        G = gridDim.x; blk = blockIdx.x; tid = threadIdx.x;
        gtid = blk * blockDim.x + tid; gs = G * blockDim.x;
    }
    __device__ long long *ldsl() const { return reinterpret_cast<long long *>(lds); }
    __device__ Dev dv(int b) const {
        Dev d = A.D;
        d.qe = A.qe + b * kQueueCap; d.qe_n = A.qn + 4 * b;
        d.qf = A.qf + b * kQueueCap; d.qf_n = A.qn + 4 * b + 1;
        return d;
    }
    __device__ double *red(int p) const { return A.red + (size_t)(p % 3) * G * kRed; }
    __device__ double *qeu(int b) const { return A.qe_u + (size_t)b * kQueueCap; }

    // Grid barrier. Arrival: one agent-scope release per block, then a store of
    // the phase number to the block's own flag line (no shared counter).
    // Departure: every block polls all flags itself (thread t polls blocks t,
    // t+256, ...), then one agent-scope acquire per block. A watchdog on the
    // poll sets the shared abort flag; every waiting block sees it and leaves.
    __device__ bool sync(int kind) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        const unsigned long long epoch = (unsigned long long)(ph + 1);
        unsigned long long *flags = A.bar + 64;
        if (tid == 0) {
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
            __hip_atomic_store(flags + (size_t)blk * 16, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        __shared__ int s_abort;
        if (tid == 0) s_abort = 0;
        __syncthreads();
        int ab = 0;
        for (int b = tid; b < G; b += blockDim.x) {
            long long spins = 0;
            while (__hip_atomic_load(flags + (size_t)b * 16, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < epoch) {
                __builtin_amdgcn_s_sleep(1);
                if ((++spins & 255) == 0) {
                    if (__hip_atomic_load(A.bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) { ab = 1; break; }
                    if (spins > A.spin_limit) {
                        __hip_atomic_store(A.bar, 1ull, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        ab = 1;
                        break;
                    }
                }
            }
            if (ab) break;
        }
        if (ab) s_abort = 1;
        __syncthreads();
        if (tid == 0) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __syncthreads();
        if (blk == 0 && tid == 0 && ph < A.prof_cap) {
            A.prof[ph] = wall_clock64();
            A.prof_kind[ph] = kind;
        }
        ph++;
        if (s_abort) aborted = true;
        // recycle the queue counters of buffer (ph+1)%3: last written two phases ago,
        // read by every block before the barrier just passed
        if (blk == 0 && tid < 4) A.qn[4 * ((ph + 1) % 3) + tid] = 0;
        return !aborted;
    }

    // ---- all interactions at `pos`: energy (exact total; per FF if PER_FF) and,
    // if FORCES, the force slots. xa: optional per-thread exact accumulator
    // reduced alongside (its total lands in ev.xa). Returns false on abort;
    // a queue overflow sets `status`.
    template <bool PER_FF, bool FORCES, class P>
    __device__ bool eval(const P &pos, Eval &ev, mdy_xsum *xa) {
        const int p0 = ph, qb = p0 % 3;
        Dev d = dv(qb);
        double acc[kNFF];
#pragma unroll
        for (int k = 0; k < kNFF; k++) acc[k] = 0.0;
        mdy_xsum xs;
        mdy_xs_zero(&xs);
        const int total = d.n_items();
        for (int i = gtid; i < total; i += gs) item<PER_FF, FORCES>(d, i, pos, acc, xs, A.contrib);
        double *r = red(p0);
        if (PER_FF) {
            block_reduce<kNFF>(acc, lds);
            if (tid < kNFF) r[(size_t)blk * kRed + tid] = acc[tid];
        }
        block_reduce_xsum(xs, ldsl());
        if (tid == 0) store_xsum(r, blk, kXsE, xs);
        if (xa) {
            block_reduce_xsum(*xa, ldsl());
            if (tid == 0) store_xsum(r, blk, kXsA, *xa);
        }
        if (!sync(FORCES ? PK_EVAL : PK_ENERGY)) return false;
        const int nqe = ld_i(d.qe_n);
        ev.nqf = FORCES ? ld_i(d.qf_n) : 0;
        ev.qb = qb;
        if (nqe > kQueueCap || ev.nqf > kQueueCap) {
            status = MDY_ERR_CAPACITY;
            return true;
        }
        if (nqe > 0) {  // quadrature fallback energies, one block per queued pair
            const int N = 400;
            for (int q = blk; q < nqe; q += G) {
                const int i = ld_i(d.qe + q);
                int4 b = d.ev_bs[i];
                SimpsonGeo g = simpson_geo(pos(b.x), pos(b.y), pos(b.z), pos(b.w));
                mdy_xsum X;
                mdy_xs_zero(&X);
                if (tid == 0) mdy_xs_add(&X, simpson_e_corners(g));
                for (int node = tid; node < (N + 1) * (N + 1); node += blockDim.x)
                    simpson_e_node(g, node / (N + 1), node % (N + 1), N, X);
                block_reduce_xsum(X, ldsl());
                if (tid == 0) {
                    const double ds = 1.0 / ((double)N);
                    qeu(qb)[q] = d.ev_kr[i] * (ds * ds / 9.0) * mdy_xs_value(&X);
                }
            }
            if (!sync(PK_EFALLBACK)) return false;
        }
        energy_epilogue(d, r, G, nqe, d.qe, qeu(qb), lds, ev.e, PER_FF);
        if (xa) {
            mdy_xsum X;
            grid_reduce_xsum(r, G, kXsA, X, ldsl());
            ev.xa = mdy_xs_value(&X);
        }
        return true;
    }

    // ---- quadrature forces for the pairs an accepted evaluation queued
    // (CylinderExclVolRepulsion.cpp:1001-1254, 1017-1022 retry with the first
    // cylinder moved off the plane). The slots are written by block 0.
    template <class P>
    __device__ bool force_fallback(const P &pos, const Eval &ev) {
        const int *qf = A.qf + ev.qb * kQueueCap;
        for (int q = 0; q < ev.nqf; q++) {
            const int i = ld_i(qf + q);
            int4 b = A.D.ev_bs[i];
            bool done = false;
            for (int moved = 0; moved < 2 && !done; moved++) {
                d3 c1 = pos(b.x), c2 = pos(b.y), c3 = pos(b.z), c4 = pos(b.w);
                const bool inplane = are_in_plane(c1, c2, c3, c4);
                if (moved) {
                    d3 n2 = moved_off_plane(c1, c2, c3, c2, 0.01), n1 = moved_off_plane(c1, c2, c3, c1, 0.01);
                    c1 = n1; c2 = n2;
                }
                SimpsonGeo g = simpson_geo(c1, c2, c3, c4);
                const int N = 10000;
                const long long nodes = (long long)(N + 1) * (N + 1);
                for (int comp = 0; comp < 6; comp++) {
                    mdy_xsum X;
                    mdy_xs_zero(&X);
                    if (gtid == 0) simpson_f_corners(g, comp, X);
                    for (long long nd = gtid; nd < nodes; nd += gs)
                        simpson_f_node(g, (int)(nd / (N + 1)), (int)(nd % (N + 1)), N, comp, X);
                    block_reduce_xsum(X, ldsl());
                    if (tid == 0) store_xsum(A.spart + (size_t)comp * G * kRed, blk, 0, X);
                }
                if (!sync(PK_FFALLBACK)) return false;
                double T[6];
                for (int comp = 0; comp < 6; comp++) {
                    mdy_xsum X;
                    grid_reduce_xsum(A.spart + (size_t)comp * G * kRed, G, 0, X, ldsl());
                    T[comp] = mdy_xs_value(&X);
                }
                d3 f[4];
                simpson_f_assemble(g, T, A.D.ev_kr[i], N, f);
                const bool ok = finite_d3(f[0]) && finite_d3(f[1]) && finite_d3(f[2]) && finite_d3(f[3]);
                if (ok) {
                    if (blk == 0 && tid == 0) {
                        int s = A.D.slot_ev + 4 * i;
                        put3(A.contrib, s, f[0]); put3(A.contrib, s + 1, f[1]);
                        put3(A.contrib, s + 2, f[2]); put3(A.contrib, s + 3, f[3]);
                    }
                    done = true;
                } else if (moved || !inplane) {
                    status = MDY_ERR_FORCE;   // the reference exits here
                    return true;
                }
                if (!sync(PK_FFALLBACK)) return false;   // spart reuse / slot visibility
            }
        }
        return true;
    }

    // ---- per-bead forces from the slots into Fout; x += lam*d when `move`.
    // Returns max|F_i|, F.F and F.Fprev (Fprev may be null), sums exact.
    __device__ bool gather(double *Fout, const double *Fprev, bool move, const double *dvec, double lam, double &maxF,
                           double &FF, double &FFp) {
        const int p1 = ph;
        double m[1] = {0.0};
        mdy_xsum sff, sfp;
        mdy_xs_zero(&sff);
        mdy_xs_zero(&sfp);
        for (int b = gtid; b < A.D.nb; b += gs) {
            d3 f = gather_bead(A.D, A.contrib, b);
            const int o = 3 * b;
            Fout[o] = f.x; Fout[o + 1] = f.y; Fout[o + 2] = f.z;
            m[0] = fmax(m[0], fmax(fabs(f.x), fmax(fabs(f.y), fabs(f.z))));
            mdy_xs_add(&sff, f.x * f.x); mdy_xs_add(&sff, f.y * f.y); mdy_xs_add(&sff, f.z * f.z);
            if (Fprev) {
                mdy_xs_add(&sfp, f.x * Fprev[o]); mdy_xs_add(&sfp, f.y * Fprev[o + 1]);
                mdy_xs_add(&sfp, f.z * Fprev[o + 2]);
            }
            if (move) {
                A.x[o] = A.x[o] + lam * dvec[o];
                A.x[o + 1] = A.x[o + 1] + lam * dvec[o + 1];
                A.x[o + 2] = A.x[o + 2] + lam * dvec[o + 2];
            }
        }
        block_reduce<1>(m, lds, 0x1u);
        block_reduce_xsum(sff, ldsl());
        double *r = red(p1);
        if (tid == 0) { r[(size_t)blk * kRed] = m[0]; store_xsum(r, blk, 8, sff); }
        if (Fprev) {
            block_reduce_xsum(sfp, ldsl());
            if (tid == 0) store_xsum(r, blk, 16, sfp);
        }
        if (!sync(PK_GATHER)) return false;
        grid_partials_reduce<1>(r, G, m, lds, 0x1u);
        maxF = m[0];
        mdy_xsum X;
        grid_reduce_xsum(r, G, 8, X, ldsl());
        FF = mdy_xs_value(&X);
        FFp = 0.0;
        if (Fprev) {
            grid_reduce_xsum(r, G, 16, X, ldsl());
            FFp = mdy_xs_value(&X);
        }
        return true;
    }
};

__global__ void __launch_bounds__(kBlock) k_cg(const CgArgs *__restrict__ Ap) {
    __shared__ double lds[17 * 16];
    const CgArgs &A = *Ap;
    Cg K(A, lds);
    double *R = A.out;
    const bool leader = (K.blk == 0 && K.tid == 0);
    unsigned flags = 0;
    int nEC = 0, nFC = 0, nIter = 0, nEE = 0, nSM = 0, nRec = 0;
    long long trn = 0;
    double gradInf = INFINITY, relE = INFINITY;
    Eval ev;
    bool after = true;   // compute energiesAfter / output force
    int cur = 0, sidx = 0;
#define FB(i) ((i) ? A.F1 : A.F0)
#define SB(i) ((i) ? A.S1 : A.S0)
    const double *dvec = nullptr;
    double maxF = 0, FF = 0, FFp = 0, FFcur = 0;

    // energiesBefore = computeIndividualEnergies(x); computeForces(x) from the same pass
    if (!K.eval<true, true>(PosX{A.x}, ev, nullptr)) goto aborted;
    if (K.status) { after = false; goto finish; }
    nEC++; nEE++;
    if (leader) {
        for (int k = 0; k < kNFF; k++) R[R_EB + k] = ev.e[k];
        R[R_EB + kNFF] = ev.e[kNFF + 1];
    }
    {
        double E0 = ev.e[kNFF];           // computeEnergy at the same x (sentinel form)
        double prevEnergy = ev.e[kNFF + 1];
        if (ev.nqf && !K.force_fallback(PosX{A.x}, ev)) goto aborted;
        if (K.status) { after = false; goto finish; }
        if (!K.gather(FB(cur), nullptr, false, nullptr, 0.0, maxF, FF, FFp)) goto aborted;
        nFC++;
        gradInf = maxF;
        if (!isfinite(gradInf)) { K.status = MDY_ERR_NONFINITE; after = false; goto finish; }
        FFcur = FF;
        // resetCGParams: s = F, lastFDotF = F.F, Fprev = F
        dvec = FB(cur);
        double lastFDotF = FF, slope = FF;
        int safeMode = 0;
        double minimumE = INFINITY;
        bool haveMin = false;
        bool specValid = false;
        double specLam = 0;
        Eval spec;
        auto callLimit = [&]() {
            return (A.energy_call_limit != 0 && nEC > A.energy_call_limit) ||
                   (A.force_call_limit != 0 && nFC > A.force_call_limit);
        };
        auto converged = [&]() { return gradInf < A.grad_tol || relE < A.energy_rel_tol; };

        while (!callLimit() && !converged()) {
            nIter++;
            const double armijo = safeMode ? 0.0 : 0.4;
            double lam = 0.0, newE = 0.0;
            if (maxF != 0.0) {
                lam = smin(A.lambda_max, A.max_dist / maxF);
                nEC++;   // currentEnergy = computeEnergy(x): E0, known bitwise from the accepting trial
                const double reducedSlope = armijo * slope;
                double Et = 0.0;
                for (;;) {
                    if (specValid && specLam == lam) {
                        ev = spec;
                    } else {
                        if (!K.eval<false, true>(PosLine{A.x, dvec, lam}, ev, nullptr)) goto aborted;
                        nEE++;
                        if (K.status) { after = false; goto finish; }
                    }
                    Et = ev.e[kNFF];
                    specValid = false;
                    nEC++;
                    const double ideal = -reducedSlope * lam;
                    bool done = false;
                    if (lam < 0 || ideal > 0) { done = true; lam = 0.0; }
                    const double change = Et - E0;
                    if (change <= ideal) done = true;
                    else lam *= 0.5;
                    if (lam <= 0.0 || lam <= 1e-8) { done = true; lam = 0.0; Et = E0; }
                    if (done) break;
                }
                newE = Et;
            }
            specValid = false;
            if (lam == 0.0) {
                if (A.try_recover) {
                    if (safeMode) {
                        // MController::moveAlongDirectionLimitingBeadMove_: step = min maxDist/|d_i|
                        double v[1] = {INFINITY};
                        for (int o = K.gtid; o < A.D.ndof; o += K.gs) v[0] = smin(v[0], A.max_dist / fabs(dvec[o]));
                        block_reduce<1>(v, lds, 0u, 0x1u);
                        const int p1 = K.ph;
                        if (K.tid == 0) K.red(p1)[(size_t)K.blk * kRed] = v[0];
                        if (!K.sync(PK_REDUCE)) goto aborted;
                        double st[1];
                        grid_partials_reduce<1>(K.red(p1), K.G, st, lds, 0u, 0x1u);
                        if (st[0] == INFINITY) { K.status = MDY_ERR_RECOVERY; after = false; goto finish; }
                        const PosLine pl{A.x, dvec, st[0]};
                        if (!K.eval<false, true>(pl, ev, nullptr)) goto aborted;
                        nEE++;
                        if (K.status) { after = false; goto finish; }
                        if (ev.nqf && !K.force_fallback(pl, ev)) goto aborted;
                        if (K.status) { after = false; goto finish; }
                        double *Fn = FB(1 - cur);
                        if (!K.gather(Fn, FB(cur), true, dvec, st[0], maxF, FF, FFp)) goto aborted;
                        nFC++; nRec++;
                        gradInf = maxF;
                        if (!isfinite(gradInf)) { K.status = MDY_ERR_NONFINITE; after = false; goto finish; }
                        cur = 1 - cur;
                        FFcur = FF;
                        dvec = FB(cur); lastFDotF = FF; slope = FF;
                        safeMode = 0;
                        E0 = ev.e[kNFF];   // computeEnergy at the moved coordinates
                    } else {
                        dvec = FB(cur); lastFDotF = FFcur; slope = FFcur;
                        safeMode = 1;
                        nSM++;
                    }
                    if (leader && trn < A.tr_cap) {
                        A.tr_lam[trn] = 0.0; A.tr_e[trn] = newE; A.tr_mf[trn] = gradInf; A.tr_beta[trn] = 0.0;
                        A.tr_safe[trn] = safeMode;
                    }
                    trn++;
                } else {
                    flags |= MDY_ERROR_LINESEARCH;
                    if (leader && trn < A.tr_cap) {
                        A.tr_lam[trn] = 0.0; A.tr_e[trn] = newE; A.tr_mf[trn] = gradInf; A.tr_beta[trn] = 0.0;
                        A.tr_safe[trn] = safeMode;
                    }
                    trn++;
                    after = false;   // the reference returns before energiesAfter / ptrOutForce
                    goto finish;
                }
                continue;
            }
            // accepted step: the force slots hold the forces at x + lam*d
            {
                const PosLine pl{A.x, dvec, lam};
                if (ev.nqf && !K.force_fallback(pl, ev)) goto aborted;
                if (K.status) { after = false; goto finish; }
                double *Fn = FB(1 - cur);
                const double *Fp = FB(cur);
                if (!K.gather(Fn, Fp, true, dvec, lam, maxF, FF, FFp)) goto aborted;
                nFC++;
                gradInf = maxF;
                if (!isfinite(gradInf)) { K.status = MDY_ERR_NONFINITE; after = false; goto finish; }
                E0 = newE;
                relE = fabs(newE - prevEnergy) / fabs(prevEnergy);
                prevEnergy = newE;
                const double newGrad = FF, prevGrad = FFp;
                double beta = 0.0;
                if (A.method == MDY_CG_FLETCHER_RIEVES) beta = newGrad / lastFDotF;
                else if (A.method == MDY_CG_POLAK_RIBIERE) beta = smax(0.0, (newGrad - prevGrad) / lastFDotF);
                // s = F + beta*s, its slope, and (speculatively) the first trial of the next line search
                const bool cont = !(callLimit() || converged());
                const bool more = cont && maxF != 0.0;
                const double lam1 = more ? smin(A.lambda_max, A.max_dist / maxF) : 0.0;
                double *Sn = SB(1 - sidx);
                double sdF = 0.0;
                if (cont) {
                    mdy_xsum sd;
                    mdy_xs_zero(&sd);
                    for (int b = K.gtid; b < A.D.nb; b += K.gs) {
                        const int o = 3 * b;
                        for (int k = 0; k < 3; k++) {
                            double s = Fn[o + k] + beta * dvec[o + k];
                            Sn[o + k] = s;
                            mdy_xs_add(&sd, s * Fn[o + k]);
                        }
                    }
                    if (more) {
                        if (!K.eval<false, true>(PosSpec{A.x, Fn, dvec, beta, lam1}, spec, &sd)) goto aborted;
                        nEE++;
                        if (K.status) { after = false; goto finish; }
                        specValid = true; specLam = lam1;
                        sdF = spec.xa;
                    } else {
                        block_reduce_xsum(sd, K.ldsl());
                        const int p1 = K.ph;
                        if (K.tid == 0) store_xsum(K.red(p1), K.blk, kXsA, sd);
                        if (!K.sync(PK_REDUCE)) goto aborted;
                        mdy_xsum X;
                        grid_reduce_xsum(K.red(p1), K.G, kXsA, X, K.ldsl());
                        sdF = mdy_xs_value(&X);
                    }
                }
                cur = 1 - cur;
                sidx = 1 - sidx;
                FFcur = newGrad;
                if (sdF <= 0) {  // resetCGParams
                    dvec = FB(cur); lastFDotF = newGrad; slope = newGrad;
                    specValid = false;
                } else {
                    dvec = Sn; slope = sdF;
                }
                // forceBelowRelaxedTolerance(10) && numIter > forceCallLimit/2 -> keep lowest-energy coordinates
                if (maxF <= 10 * A.grad_tol && nIter > A.force_call_limit / 2 && newE <= minimumE) {
                    minimumE = newE;
                    haveMin = true;
                    for (int o = K.gtid; o < A.D.ndof; o += K.gs) A.xmin[o] = A.x[o];
                }
                if (A.method != MDY_CG_STEEPEST) lastFDotF = newGrad;
                if (leader && trn < A.tr_cap) {
                    A.tr_lam[trn] = lam; A.tr_e[trn] = newE; A.tr_mf[trn] = gradInf; A.tr_beta[trn] = beta;
                    A.tr_safe[trn] = safeMode;
                }
                trn++;
            }
        }
        if (callLimit() && haveMin) {
            for (int o = K.gtid; o < A.D.ndof; o += K.gs) A.x[o] = A.xmin[o];
            if (!K.sync(PK_RESTORE)) goto aborted;
        }
    }
finish:
    if (after) {
        if (!K.eval<true, false>(PosX{A.x}, ev, nullptr)) goto aborted;
        nEC++; nEE++;
        if (leader) {
            for (int k = 0; k < kNFF; k++) R[R_EA + k] = ev.e[k];
            R[R_EA + kNFF] = ev.e[kNFF + 1];
        }
    }
    if (leader) {
        R[R_FLAGS] = flags; R[R_NEC] = nEC; R[R_NFC] = nFC; R[R_NIT] = nIter; R[R_GIN] = gradInf; R[R_ERC] = relE;
        R[R_NEE] = nEE; R[R_NSM] = nSM; R[R_NREC] = nRec; R[R_STATUS] = K.status; R[R_TRN] = (double)trn;
        R[R_LAST] = after ? 1.0 : 0.0; R[R_FOUT] = (double)cur; R[R_NPH] = (double)K.ph;
    }
    return;
aborted:
    if (leader) R[R_STATUS] = MDY_ERR_TIMEOUT;
#undef FB
#undef SB
}

// ---------------------------------------------------------------------------
int run_minimize(Ctx &c, const mdy_cg_params *p, double *d_x, double *d_force_out, mdy_min_result *res) {
    hipStream_t st = c.stream;
    const int ndof = 3 * c.nb;
    MDY_TRY(ensure(c.F0, ndof)); MDY_TRY(ensure(c.F1, ndof));
    MDY_TRY(ensure(c.S0, ndof)); MDY_TRY(ensure(c.S1, ndof));
    MDY_TRY(ensure(c.xmin, ndof));
    MDY_TRY(ensure(c.prof, c.prof_cap + 1));
    MDY_TRY(ensure(c.prof_kind, c.prof_cap + 1));
    MDY_TRY(ensure(c.cg_args, sizeof(CgArgs) / sizeof(double) + 1));
    MDY_TRY(ensure(c.cg_out, R_SIZE));
    MDY_TRY(ensure(c.qe, 3 * kQueueCap)); MDY_TRY(ensure(c.qf, 3 * kQueueCap));
    MDY_TRY(ensure(c.qn, 16));
    MDY_TRY(ensure(c.qe_u, 3 * kQueueCap));
    MDY_TRY(ensure(c.tr_lam, c.trace_cap + 1)); MDY_TRY(ensure(c.tr_e, c.trace_cap + 1));
    MDY_TRY(ensure(c.tr_mf, c.trace_cap + 1)); MDY_TRY(ensure(c.tr_beta, c.trace_cap + 1));
    MDY_TRY(ensure(c.tr_safe, c.trace_cap + 1));
    // persistent grid: all blocks resident
    int per_cu = 0;
    MDY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void *>(k_cg), kBlock, 0));
    if (per_cu < 1) return MDY_ERR_LAUNCH;
    if (c.opts.blocks_per_cu > 0 && c.opts.blocks_per_cu < per_cu) per_cu = c.opts.blocks_per_cu;
    const int G = per_cu * c.num_cu;
    c.coop_blocks = G;
    MDY_TRY(ensure(c.bar, 64 + 16 * (size_t)G));
    MDY_TRY(ensure(c.red, (size_t)3 * G * kRed));
    MDY_TRY(ensure(c.simpson_part, (size_t)6 * G * kRed));
    MDY_HIP(hipMemsetAsync(c.bar.p, 0, (64 + 16 * (size_t)G) * sizeof(unsigned long long), st));
    MDY_HIP(hipMemsetAsync(c.qn.p, 0, 16 * sizeof(int), st));
    MDY_HIP(hipMemsetAsync(c.err.p, 0, sizeof(int), st));
    CgArgs A;
    A.D = c.dev();
    A.x = d_x; A.F0 = c.F0.p; A.F1 = c.F1.p; A.S0 = c.S0.p; A.S1 = c.S1.p; A.xmin = c.xmin.p;
    A.contrib = c.contrib.p; A.red = c.red.p; A.qe_u = c.qe_u.p; A.spart = c.simpson_part.p;
    A.qe = c.qe.p; A.qf = c.qf.p; A.qn = c.qn.p; A.bar = c.bar.p; A.out = c.cg_out.p;
    A.tr_lam = c.tr_lam.p; A.tr_e = c.tr_e.p; A.tr_mf = c.tr_mf.p; A.tr_beta = c.tr_beta.p; A.tr_safe = c.tr_safe.p;
    A.tr_cap = c.trace_cap;
    A.prof = c.prof.p; A.prof_kind = c.prof_kind.p; A.prof_cap = c.prof_cap;
    A.method = p->method; A.try_recover = p->try_to_recover; A.energy_call_limit = p->energy_call_limit;
    A.force_call_limit = p->force_call_limit > 3 * p->n_beads_limit ? p->force_call_limit : 3 * p->n_beads_limit;
    A.max_dist = p->max_dist; A.lambda_max = p->lambda_max; A.grad_tol = p->grad_tol;
    A.energy_rel_tol = p->energy_rel_tol;
    A.spin_limit = 20000000LL;   // polls (about 10 s) before the watchdog fires
    hipEvent_t e0, e1;
    MDY_HIP(hipEventCreate(&e0));
    MDY_HIP(hipEventCreate(&e1));
    MDY_HIP(hipMemcpyAsync(c.cg_args.p, &A, sizeof(CgArgs), hipMemcpyHostToDevice, st));
    MDY_HIP(hipEventRecord(e0, st));
    const CgArgs *Ad = reinterpret_cast<const CgArgs *>(c.cg_args.p);
    void *args[] = {&Ad};
    hipError_t le = hipLaunchCooperativeKernel(reinterpret_cast<const void *>(k_cg), dim3(G), dim3(kBlock), args, 0, st);
    if (le != hipSuccess) {
        (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
        return MDY_ERR_LAUNCH;
    }
    MDY_HIP(hipEventRecord(e1, st));
    double R[R_SIZE];
    MDY_HIP(hipMemcpyAsync(R, c.cg_out.p, sizeof(R), hipMemcpyDeviceToHost, st));
    MDY_HIP(hipStreamSynchronize(st));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    c.last_ms = ms;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    std::memset(res, 0, sizeof(*res));
    for (int k = 0; k <= kNFF; k++) { res->energies_before[k] = R[R_EB + k]; res->energies_after[k] = R[R_EA + k]; }
    res->error_flags = (uint32_t)R[R_FLAGS];
    res->num_energy_call = (int32_t)R[R_NEC];
    res->energy_call_limit = p->energy_call_limit;
    res->num_force_call = (int32_t)R[R_NFC];
    res->force_call_limit = A.force_call_limit;
    res->num_iter = (int32_t)R[R_NIT];
    res->grad_inf_norm = R[R_GIN];
    res->grad_tol = p->grad_tol;
    res->energy_rel_change = R[R_ERC];
    res->energy_rel_tol = p->energy_rel_tol;
    res->num_energy_eval = (int32_t)R[R_NEE];
    res->num_safe_mode = (int32_t)R[R_NSM];
    res->num_recovery = (int32_t)R[R_NREC];
    res->status = (int32_t)R[R_STATUS];
    c.trace_n = (long long)R[R_TRN];
    c.prof_n = (long long)R[R_NPH];
    if (d_force_out && R[R_LAST] != 0.0) {
        const double *src = ((int)R[R_FOUT] == 0) ? c.F0.p : c.F1.p;
        MDY_HIP(hipMemcpyAsync(d_force_out, src, sizeof(double) * ndof, hipMemcpyDeviceToDevice, st));
        MDY_HIP(hipStreamSynchronize(st));
    }
    return res->status == MDY_ERR_TIMEOUT ? MDY_ERR_TIMEOUT : MDY_OK;
}

}  // namespace mdy
