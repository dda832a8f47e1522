// mdy_cg.hip -- CGMethod::minimize<method> with backtracking line search,
// resident on the device for the whole minimization
// (src/Mechanics/Minimizer/CGMethod.hpp:238-373, 826-1743; CGMethod.cpp:43-360).
//
// One cooperative launch; every block runs the same control flow and derives
// every decision from grid reductions that all blocks evaluate in the same
// fixed order, so decisions are identical without a broadcast. Phases are
// separated by a grid barrier (atomic arrival counter + agent-scope fences).
// An accepted iteration costs three barriers:
//   [forces at x+lam*d -> slots] | [gather F, x += lam*d, max|F|, F.F, F.Fprev]
//   | [beta, s = F + beta*d, s.F, first trial energy of the next line search]
// The energy at the accepted point is the line-search trial energy of the same
// coordinates (bitwise), so E(x) is not recomputed.
#include "mdy_eval.hpp"
#include "mdy_launch.hpp"

#include <cstring>

namespace mdy {

struct CgArgs {
    Dev D;
    double *x, *F0, *F1, *S0, *S1, *xmin, *contrib;
    double *red;          // 3 * nblk * kRed
    double *qe_u;         // 3 * kQueueCap
    double *spart;        // nblk * kRed
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
enum { PK_ENERGY = 1, PK_EFALLBACK, PK_FORCE, PK_FFALLBACK, PK_GATHER, PK_REDUCE, PK_RESTORE };

struct Cg {
    const CgArgs &A;
    double *lds;
    int G, blk, tid, gtid, gs;
    int ph = 0;                  // phase counter (uniform)
    bool aborted = false;

    __device__ Cg(const CgArgs &a, double *l) : A(a), lds(l) {
        G = gridDim.x; blk = blockIdx.x; tid = threadIdx.x;
        gtid = blk * blockDim.x + tid; gs = G * blockDim.x;
    }
    __device__ Dev dv(int p) const {
        Dev d = A.D;
        int b = p % 3;
        d.qe = A.qe + b * kQueueCap; d.qe_n = A.qn + 4 * b;
        d.qf = A.qf + b * kQueueCap; d.qf_n = A.qn + 4 * b + 1;
        return d;
    }
    __device__ double *red(int p) const { return A.red + (size_t)(p % 3) * G * kRed; }
    __device__ double *qeu(int p) const { return A.qe_u + (size_t)(p % 3) * kQueueCap; }

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

    // ---- energy of all interactions at positions `pos`; e[kNFF]=sentinel total, e[kNFF+1]=plain sum
    // extra: optional per-thread value reduced alongside (returned in *extra_out)
    template <class P>
    __device__ bool energy(const P &pos, double e[kNFF + 2], double extra_in = 0.0, double *extra_out = nullptr) {
        const int p0 = ph;
        Dev d = dv(p0);
        double acc[kNFF + 1];
#pragma unroll
        for (int k = 0; k < kNFF; k++) acc[k] = 0.0;
        acc[kNFF] = extra_in;
        const int total = d.n_items();
        for (int i = gtid; i < total; i += gs) energy_item(d, i, pos, acc);
        block_reduce<kNFF + 1>(acc, lds);
        if (tid < kNFF + 1) red(p0)[(size_t)blk * kRed + tid] = acc[tid];
        if (!sync(PK_ENERGY)) return false;
        const int nq = min(ld_i(d.qe_n), kQueueCap);
        if (nq > 0) {  // quadrature fallback, one block per queued pair
            const int N = 400;
            for (int q = blk; q < nq; q += G) {
                const int i = ld_i(d.qe + q);
                int4 b = d.ev_bs[i];
                SimpsonGeo g = simpson_geo(pos(b.x), pos(b.y), pos(b.z), pos(b.w));
                double a1[1] = {0.0};
                for (int node = tid; node < (N + 1) * (N + 1); node += blockDim.x)
                    a1[0] += simpson_e_node(g, node / (N + 1), node % (N + 1), N);
                block_reduce<1>(a1, lds);
                if (tid == 0) {
                    const double ds = 1.0 / ((double)N);
                    qeu(p0)[q] = d.ev_kr[i] * (ds * ds / 9.0) * (simpson_e_corners(g) + a1[0]);
                }
            }
            if (!sync(PK_EFALLBACK)) return false;
        }
        energy_epilogue(d, red(p0), G, nq, qeu(p0), lds, e);
        if (extra_out) {
            double v[1];
            // the extra column sits at index kNFF of the same partials
            v[0] = 0.0;
            for (int b = tid; b < G; b += blockDim.x) v[0] += __builtin_nontemporal_load(red(p0) + (size_t)b * kRed + kNFF);
            block_reduce<1>(v, lds);
            *extra_out = v[0];
        }
        return true;
    }

    // ---- forces at `pos` into Fout; optionally x += lam*d for every bead.
    // Returns max|F_i|, F.F and F.Fprev (Fprev may be null).
    template <class P>
    __device__ bool forces(const P &pos, double *Fout, const double *Fprev, bool move, const double *dvec, double lam,
                           double &maxF, double &FF, double &FFp, int &status) {
        const int p0 = ph;
        Dev d = dv(p0);
        const int total = d.n_items();
        for (int i = gtid; i < total; i += gs) force_item(d, i, pos, A.contrib);
        if (!sync(PK_FORCE)) return false;
        const int nq = ld_i(d.qf_n);
        if (nq > kQueueCap) { status = MDY_ERR_CAPACITY; return true; }
        for (int q = 0; q < nq; q++) {
            const int i = ld_i(d.qf + q);
            int4 b = d.ev_bs[i];
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
                double o[6] = {0, 0, 0, 0, 0, 0};
                const long long nodes = (long long)(N + 1) * (N + 1);
                for (long long nd = gtid; nd < nodes; nd += gs) simpson_f_node(g, (int)(nd / (N + 1)), (int)(nd % (N + 1)), N, o);
                block_reduce<6>(o, lds);
                if (tid < 6) A.spart[(size_t)blk * kRed + tid] = o[tid];
                if (!sync(PK_FFALLBACK)) return false;
                double T[6];
                grid_partials_reduce<6>(A.spart, G, T, lds);
                double cn[6] = {0, 0, 0, 0, 0, 0};
                fint6(g, 0.0, 0.0, 1.0, cn); fint6(g, 0.0, 1.0, 1.0, cn);
                fint6(g, 1.0, 0.0, 1.0, cn); fint6(g, 1.0, 1.0, 1.0, cn);
                for (int k = 0; k < 6; k++) T[k] = cn[k] + T[k];
                d3 f[4];
                simpson_f_assemble(g, T, d.ev_kr[i], N, f);
                const bool ok = finite_d3(f[0]) && finite_d3(f[1]) && finite_d3(f[2]) && finite_d3(f[3]);
                if (ok) {
                    if (blk == 0 && tid == 0) {
                        int s = d.slot_ev + 4 * i;
                        put3(A.contrib, s, f[0]); put3(A.contrib, s + 1, f[1]);
                        put3(A.contrib, s + 2, f[2]); put3(A.contrib, s + 3, f[3]);
                    }
                    done = true;
                } else if (moved || !inplane) {
                    status = MDY_ERR_FORCE;
                    return true;
                }
                if (!sync(PK_FFALLBACK)) return false;   // spart reuse / contrib visibility
            }
        }
        // gather + optional move + dot products
        double v[3] = {0.0, 0.0, 0.0};  // max|F|, F.F, F.Fprev
        for (int b = gtid; b < d.nb; b += gs) {
            d3 f = gather_bead(d, A.contrib, b);
            const int o = 3 * b;
            Fout[o] = f.x; Fout[o + 1] = f.y; Fout[o + 2] = f.z;
            v[0] = fmax(v[0], fmax(fabs(f.x), fmax(fabs(f.y), fabs(f.z))));
            v[1] += f.x * f.x; v[1] += f.y * f.y; v[1] += f.z * f.z;
            if (Fprev) { v[2] += f.x * Fprev[o]; v[2] += f.y * Fprev[o + 1]; v[2] += f.z * Fprev[o + 2]; }
            if (move) {
                A.x[o] = A.x[o] + lam * dvec[o];
                A.x[o + 1] = A.x[o + 1] + lam * dvec[o + 1];
                A.x[o + 2] = A.x[o + 2] + lam * dvec[o + 2];
            }
        }
        block_reduce<3>(v, lds, 0x1u);
        const int p1 = ph;
        if (tid < 3) red(p1)[(size_t)blk * kRed + tid] = v[tid];
        if (!sync(PK_GATHER)) return false;
        double r[3];
        grid_partials_reduce<3>(red(p1), G, r, lds, 0x1u);
        maxF = r[0]; FF = r[1]; FFp = r[2];
        return true;
    }
};

__global__ void __launch_bounds__(kBlock) k_cg(const CgArgs *__restrict__ Ap) {
    __shared__ double lds[17 * 16];
    const CgArgs &A = *Ap;
    Cg K(A, lds);
    double *R = A.out;
    const bool leader = (K.blk == 0 && K.tid == 0);
    int status = MDY_OK;
    unsigned flags = 0;
    int nEC = 0, nFC = 0, nIter = 0, nEE = 0, nSM = 0, nRec = 0;
    long long trn = 0;
    double gradInf = INFINITY, relE = INFINITY;
    double e[kNFF + 2];
    bool after = true;   // compute energiesAfter / output force
    int cur = 0, sidx = 0;
#define FB(i) ((i) ? A.F1 : A.F0)
#define SB(i) ((i) ? A.S1 : A.S0)
    const double *dvec = nullptr;
    double maxF = 0, FF = 0, FFp = 0, FFcur = 0;

    // energiesBefore = computeIndividualEnergies(x)
    if (!K.energy(PosX{A.x}, e)) goto aborted;
    nEC++; nEE++;
    if (leader) for (int k = 0; k < kNFF; k++) R[R_EB + k] = e[k];
    if (leader) R[R_EB + kNFF] = e[kNFF + 1];
    {
        double E0 = e[kNFF];           // computeEnergy at the same x (sentinel form)
        bool E0valid = true;
        double prevEnergy = e[kNFF + 1];
        if (!K.forces(PosX{A.x}, FB(cur), nullptr, false, nullptr, 0.0, maxF, FF, FFp, status)) goto aborted;
        nFC++;
        if (status) { after = false; goto finish; }
        gradInf = maxF;
        if (!isfinite(gradInf)) { status = MDY_ERR_NONFINITE; after = false; goto finish; }
        FFcur = FF;
        // resetCGParams: s = F, lastFDotF = F.F, Fprev = F
        dvec = FB(cur);
        double lastFDotF = FF, slope = FF;
        int safeMode = 0;
        double minimumE = INFINITY;
        bool haveMin = false;
        bool specValid = false;
        double specLam = 0, specE = 0;
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
                if (!E0valid) {
                    if (!K.energy(PosX{A.x}, e)) goto aborted;
                    nEE++;
                    E0 = e[kNFF];
                    E0valid = true;
                }
                nEC++;
                const double reducedSlope = armijo * slope;
                double Et = 0.0;
                for (;;) {
                    if (specValid && specLam == lam) {
                        Et = specE;
                    } else {
                        if (!K.energy(PosLine{A.x, dvec, lam}, e)) goto aborted;
                        nEE++;
                        Et = e[kNFF];
                    }
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
                        if (st[0] == INFINITY) { status = MDY_ERR_RECOVERY; after = false; goto finish; }
                        double *Fn = FB(1 - cur);
                        if (!K.forces(PosLine{A.x, dvec, st[0]}, Fn, FB(cur), true, dvec, st[0], maxF, FF, FFp, status))
                            goto aborted;
                        nFC++; nRec++;
                        if (status) { after = false; goto finish; }
                        gradInf = maxF;
                        if (!isfinite(gradInf)) { status = MDY_ERR_NONFINITE; after = false; goto finish; }
                        cur = 1 - cur;
                        FFcur = FF;
                        dvec = FB(cur); lastFDotF = FF; slope = FF;
                        safeMode = 0;
                        E0valid = false;
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
            // accepted step: x += lam*d, forces at the new x
            {
                double *Fn = FB(1 - cur);
                const double *Fp = FB(cur);
                if (!K.forces(PosLine{A.x, dvec, lam}, Fn, Fp, true, dvec, lam, maxF, FF, FFp, status)) goto aborted;
                nFC++;
                if (status) { after = false; goto finish; }
                gradInf = maxF;
                if (!isfinite(gradInf)) { status = MDY_ERR_NONFINITE; after = false; goto finish; }
                E0 = newE;
                E0valid = true;
                relE = fabs(newE - prevEnergy) / fabs(prevEnergy);
                prevEnergy = newE;
                const double newGrad = FF, prevGrad = FFp;
                double beta = 0.0;
                if (A.method == MDY_CG_FLETCHER_RIEVES) beta = newGrad / lastFDotF;
                else if (A.method == MDY_CG_POLAK_RIBIERE) beta = smax(0.0, (newGrad - prevGrad) / lastFDotF);
                // s = F + beta*s, its slope, and (speculatively) the first trial of the next line search
                const bool more = !(callLimit() || converged()) && maxF != 0.0;
                const double lam1 = more ? smin(A.lambda_max, A.max_dist / maxF) : 0.0;
                double *Sn = SB(1 - sidx);
                double sdot = 0.0;
                for (int b = K.gtid; b < A.D.nb; b += K.gs) {
                    const int o = 3 * b;
                    for (int k = 0; k < 3; k++) {
                        double s = Fn[o + k] + beta * dvec[o + k];
                        Sn[o + k] = s;
                        sdot += s * Fn[o + k];
                    }
                }
                double sdF = 0.0;
                if (more) {
                    if (!K.energy(PosSpec{A.x, Fn, dvec, beta, lam1}, e, sdot, &sdF)) goto aborted;
                    nEE++;
                    specValid = true; specLam = lam1; specE = e[kNFF];
                } else {
                    double v[1] = {sdot};
                    block_reduce<1>(v, lds);
                    const int p1 = K.ph;
                    if (K.tid == 0) K.red(p1)[(size_t)K.blk * kRed] = v[0];
                    if (!K.sync(PK_REDUCE)) goto aborted;
                    double r1[1];
                    grid_partials_reduce<1>(K.red(p1), K.G, r1, lds);
                    sdF = r1[0];
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
        if (!K.energy(PosX{A.x}, e)) goto aborted;
        nEC++; nEE++;
        if (leader) {
            for (int k = 0; k < kNFF; k++) R[R_EA + k] = e[k];
            R[R_EA + kNFF] = e[kNFF + 1];
        }
    }
    if (leader) {
        R[R_FLAGS] = flags; R[R_NEC] = nEC; R[R_NFC] = nFC; R[R_NIT] = nIter; R[R_GIN] = gradInf; R[R_ERC] = relE;
        R[R_NEE] = nEE; R[R_NSM] = nSM; R[R_NREC] = nRec; R[R_STATUS] = status; R[R_TRN] = (double)trn;
        R[R_LAST] = after ? 1.0 : 0.0; R[R_FOUT] = (double)cur; R[R_NPH] = (double)K.ph;
    }
    return;
aborted:
    if (leader) R[R_STATUS] = MDY_ERR_TIMEOUT;
}

// ---------------------------------------------------------------------------
__global__ void k_cg_copy(const double *src, double *dst, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = src[i];
}

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
    MDY_TRY(ensure(c.simpson_part, (size_t)G * kRed));
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
    A.spin_limit = 20000000LL;   // seconds of polling before the watchdog fires
    hipEvent_t e0, e1;
    MDY_HIP(hipEventCreate(&e0));
    MDY_HIP(hipEventCreate(&e1));
    MDY_HIP(hipEventRecord(e0, st));
    MDY_HIP(hipMemcpyAsync(c.cg_args.p, &A, sizeof(CgArgs), hipMemcpyHostToDevice, st));
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
