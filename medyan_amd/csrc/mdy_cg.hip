// mdy_cg.hip -- CGMethod::minimize<method> with the backtracking or the
// quadratic line search, resident on the device for the whole minimization
// (src/Mechanics/Minimizer/CGMethod.hpp:238-373, 522-765, 826-1743;
// CGMethod.cpp:43-360).
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

#include <algorithm>
#include <cstring>
#include <utility>
#include <vector>

namespace mdy {

struct CgArgs {
    Dev D;
    double *x, *F0, *F1, *S0, *S1, *xmin, *contrib;
    double *red;          // 3 * nblk * kRed
    double *qe_u;         // 3 * kQueueCap
    double *spart;        // 6 * nblk * kRed
    int *qe, *qf, *qn;    // 3 * kQueueCap each; qn: 3 x 4 counters
    unsigned long long *bar;  // [0] abort flag; block b arrival flag at [64 + 16 b]
    unsigned long long *acc;  // 3 x kAccWords phase accumulators
    double *out;          // result record
    long long *prof;      // phase profile, 5 words per phase: barrier exit (block 0), latest
                          // arrival (all blocks), block 0's work start, end of its item
                          // loop, and its arrival
    int *prof_kind;       // phase kind code
    long long prof_cap;
    double *tr_lam, *tr_e, *tr_mf, *tr_beta;
    int *tr_safe;
    long long tr_cap;
    // parameters
    int method, line_search, try_recover, energy_call_limit, force_call_limit, report_call_limit;
    double max_dist, lambda_max, grad_tol, energy_rel_tol;
    long long spin_limit;
    int grid;             // blocks of this minimization's cluster (batched launch)
    // batched passes: suspend at the top of an iteration once this many force
    // calls are done (0: never), and resume from the state a suspension saved
    int suspend_calls, resume;
};

// result record layout (doubles)
enum {
    R_EB = 0,              // energies before [kNFF] + total
    R_EA = R_EB + kNFF + 1,  // energies after
    R_FLAGS = R_EA + kNFF + 1,
    R_NEC, R_NFC, R_NIT, R_GIN, R_ERC, R_NEE, R_NSM, R_NREC, R_STATUS, R_TRN, R_LAST, R_FOUT, R_NPH,
    R_T0, R_T1,            // 100 MHz wall clock at the leader's start and end
    R_DIR, R_PEND, R_BETA, R_DPREV,   // search direction at the end (buffer ids 0 F0, 1 F1, 2 S0, 3 S1)
    R_SUSP,                // 1: the run suspended at the top of an iteration (batched passes)
    R_RS,                  // the state a suspension saves (26 words)
    R_SIZE = R_RS + 26
};

__device__ __forceinline__ int ld_i(const int *p) { return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT); }

// phase accumulator sets: 3 rotating sets of kAccWords 64-bit words
constexpr int kAccWords = 32;
constexpr int kAcc0 = 0, kAcc1 = 8, kAccMax = 16, kAccDone = 19;

// phase kinds recorded by the profile
enum { PK_ENERGY = 1, PK_EFALLBACK, PK_EVAL, PK_FFALLBACK, PK_GATHER, PK_REDUCE, PK_RESTORE };

// result of one hot-loop evaluation: the computeEnergy total (+inf sentinel),
// the extra accumulator, and the force-fallback queue it filled
struct Eval {
    double E;
    double xa;     // s.F, when requested
    int nqf;       // queued force fallbacks
    int qb;        // queue buffer (0..2)
};

template <bool MIXED>
struct Cg {
    using Slot = typename SlotT<MIXED>::T;   // force records: double, or float in the mixed form
    const CgArgs &A;
    double *lds;
    unsigned long long *accl;    // LDS copy of the words a phase's decisions read
    int G, blk, tid, gtid, gs;
    int ph = 0;                  // phase counter (uniform)
    // batched passes by completions: the launch's count of finished
    // minimizations (nullable), and its value as of the last gather (uniform)
    int *done_ctr = nullptr;
    int done_seen = 0;
    bool aborted = false;
    int status = MDY_OK;

    // blk: this block's index within the cluster of G blocks that runs the
    // minimization (the whole grid, or one network's share of a batch)
    __device__ Cg(const CgArgs &a, double *l, unsigned long long *al, int blk_, int G_) : A(a), lds(l), accl(al) {
        G = G_; blk = blk_; tid = threadIdx.x;
        gtid = blk * blockDim.x + tid; gs = G * blockDim.x;
    }
    __device__ long long *ldsl() const { return reinterpret_cast<long long *>(lds); }
    __device__ Slot *slots() const { return reinterpret_cast<Slot *>(A.contrib); }
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
    __attribute__((always_inline)) __device__ bool sync(int kind) {
        __builtin_amdgcn_s_waitcnt(0);
        __syncthreads();
        const unsigned long long epoch = (unsigned long long)(ph + 1);
        unsigned long long *flags = A.bar + 64;
        if (tid == 0) {
            if (ph < A.prof_cap) {
                const unsigned long long t = wall_clock64();
                atomicMax(reinterpret_cast<unsigned long long *>(A.prof) + 5 * (size_t)ph + 1, t);
                if (blk == 0) A.prof[5 * (size_t)ph + 4] = (long long)t;
            }
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
                __builtin_amdgcn_s_sleep(64);   // ~4k cycles between polls: less L2 traffic while others work
                if (__builtin_expect((++spins & 255) == 0, 0)) {
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
            A.prof[5 * (size_t)ph] = wall_clock64();
            A.prof_kind[ph] = kind;
        }
        ph++;
        if (s_abort) aborted = true;
        // recycle the queue counters of buffer (ph+1)%3: last written two phases ago,
        // read by every block before the barrier just passed
        if (blk == 0 && tid < 4) A.qn[4 * ((ph + 1) % 3) + tid] = 0;
        // and the accumulator set of that phase (read by every block right
        // after its barrier, which every block has since left)
        if (blk == 0 && tid < kAccWords)
            __hip_atomic_store(A.acc + (size_t)((ph + 1) % 3) * kAccWords + tid, 0ull, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        return !aborted;
    }

    // profile marks of block 0 inside a phase: 0 = work start, 1 = end of the main loop
    __device__ void mark(int i) const {
        if (blk == 0 && tid == 0 && ph < A.prof_cap) A.prof[5 * (size_t)ph + 2 + i] = wall_clock64();
    }

    // ---- phase accumulators. Every exact sum a decision needs (energy, s.F,
    // F.F, F.Fprev) is reduced per block, then added digit by digit with
    // 64-bit integer atomics into one of three rotating sets (phase % 3):
    // integer addition is associative, so the total is independent of the
    // arrival order and every block reads the same bits right after the
    // barrier. Max |F| uses an integer max on the bit pattern (values >= 0).
    // A set is read right after its phase's barrier and zeroed by block 0 one
    // barrier later.
    __device__ unsigned long long *accs(int p) const { return A.acc + (size_t)(p % 3) * kAccWords; }
    // all threads hold the block total `s` (after block_reduce_xsum)
    __attribute__((always_inline)) __device__ void acc_push(unsigned long long *base, const mdy_xsum &s) const {
        if (tid < 8) {
            const long long v = tid == 0 ? s.d[0] : tid == 1 ? s.d[1] : tid == 2 ? s.d[2] : tid == 3 ? s.d[3]
                              : tid == 4 ? s.d[4] : tid == 5 ? s.d[5] : tid == 6 ? s.d[6] : (long long)s.flags;
            if (v != 0) {
                if (tid < 7) atomicAdd(base + tid, (unsigned long long)v);
                else atomicOr(base + 7, (unsigned long long)v);
            }
        }
    }
    __attribute__((always_inline)) __device__ double acc_value(const unsigned long long *base) const {
        mdy_xsum X;
#pragma unroll
        for (int k = 0; k < MDY_XS_N; k++)
            X.d[k] = (long long)__hip_atomic_load(base + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        X.flags = (int)__hip_atomic_load(base + 7, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        X.cnt = 0;
        return mdy_xs_value(&X);
    }
    // After a barrier every thread needs the same accumulator words: the
    // first wave loads them once per block into LDS (one request instead of
    // one per wave to the same L2 lines), then every thread reads LDS.
    // extra0/extra1 (nullable): counters loaded into words 17 and 18.
    __attribute__((always_inline)) __device__ void acc_fetch(const unsigned long long *base, const int *extra0,
                                                             const int *extra1) const {
        if (tid < 17) accl[tid] = __hip_atomic_load(base + tid, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        else if (tid == 17 && extra0) accl[17] = (unsigned long long)(unsigned)ld_i(extra0);
        else if (tid == 18 && extra1) accl[18] = (unsigned long long)(unsigned)ld_i(extra1);
        __syncthreads();
    }
    __attribute__((always_inline)) __device__ mdy_xsum acc_local(int off) const {
        mdy_xsum X;
#pragma unroll
        for (int k = 0; k < MDY_XS_N; k++) X.d[k] = (long long)accl[off + k];
        X.flags = (int)accl[off + 7];
        X.cnt = 0;
        return X;
    }
    __attribute__((always_inline)) __device__ mdy_xsum acc_read(const unsigned long long *base) const {
        mdy_xsum X;
#pragma unroll
        for (int k = 0; k < MDY_XS_N; k++)
            X.d[k] = (long long)__hip_atomic_load(base + k, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        X.flags = (int)__hip_atomic_load(base + 7, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        X.cnt = 0;
        return X;
    }

    // ---- quadrature fallback energies of the pairs queued in buffer qb, one
    // block per pair (CylinderExclVolRepulsion.cpp:884-999, 442-481)
    __attribute__((always_inline)) __device__ bool energy_fallback(const PosAny &pos, const Dev &d, int nqe, int qb) {
        const int N = 400;
        for (int q = blk; q < nqe; q += G) {
            const int i = ld_i(d.qe + q);
            int4 b = d.ev_bs[i];
            SimpsonGeo g = simpson_geo(pos(b.x), pos(b.y), pos(b.z), pos(b.w));
            mdy_xsum Y;
            mdy_xs_zero(&Y);
            if (tid == 0) mdy_xs_add(&Y, simpson_e_corners(g));
            for (int node = tid; node < (N + 1) * (N + 1); node += blockDim.x)
                simpson_e_node(g, node / (N + 1), node % (N + 1), N, Y);
            block_reduce_xsum(Y, ldsl());
            if (tid == 0) {
                const double ds = 1.0 / ((double)N);
                qeu(qb)[q] = d.ev_kr[i] * (ds * ds / 9.0) * mdy_xs_value(&Y);
            }
        }
        return sync(PK_EFALLBACK);
    }

    // ---- hot loop: all interactions at `pos`, energy (exact total) and force
    // slots. sd: this thread's share of s.F (zero without a new direction),
    // reduced together with the energy into the phase's second accumulator;
    // with_sd: its value lands in ev.xa. Returns false on abort; a queue
    // overflow sets `status`.
    __attribute__((always_inline)) __device__ bool eval(const PosAny &pos, Eval &ev, bool with_sd, mdy_xsum &sd) {
        const int p0 = ph, qb = p0 % 3;
        Dev d = dv(qb);
        mdy_xsum xs;
        mdy_xs_zero(&xs);
        const int total = d.n_items();
        mark(0);
        sweep<MIXED>(d, pos, gtid, gs, nullptr, xs, slots(), false, true);
        mark(1);
        block_reduce_xsum2_lds(xs, sd, ldsl());
        acc_push(accs(p0) + kAcc0, xs);
        acc_push(accs(p0) + kAcc1, sd);
        if (!sync(PK_EVAL)) return false;
        acc_fetch(accs(p0), d.qe_n, d.qf_n);
        mdy_xsum X = acc_local(kAcc0);
        if (with_sd) {
            mdy_xsum Y = acc_local(kAcc1);
            ev.xa = mdy_xs_value(&Y);
        }
        const int nqe = (int)accl[17];
        ev.nqf = (int)accl[18];
        ev.qb = qb;
        if (__builtin_expect(nqe > kQueueCap || ev.nqf > kQueueCap, 0)) {
            status = MDY_ERR_CAPACITY;
            return true;
        }
        if (__builtin_expect(nqe > 0, 0) && !energy_fallback(pos, d, nqe, qb)) return false;
        double e[kNFF + 2];
        energy_finish(X, nullptr, nqe, d.qe, qeu(qb), e);
        ev.E = e[kNFF];
        return true;
    }

    // ---- computeIndividualEnergies at x (energies only, per force field):
    // out[0..7] per force field, out[8] their sum; returns the computeEnergy
    // total (+inf sentinel) in E and the plain total in Eplain.
    __attribute__((always_inline)) __device__ bool eval_ff(const PosAny &pos, double *out, double &E, double &Eplain) {
        const int p0 = ph, qb = p0 % 3;
        Dev d = dv(qb);
        double acc[kNFF];
#pragma unroll
        for (int k = 0; k < kNFF; k++) acc[k] = 0.0;
        mdy_xsum xs;
        mdy_xs_zero(&xs);
        const int total = d.n_items();
        sweep<MIXED>(d, pos, gtid, gs, acc, xs, nullptr, true, false);
        double *r = red(p0);
        block_reduce<kNFF>(acc, lds);
        if (tid < kNFF) r[(size_t)blk * kRed + tid] = acc[tid];
        block_reduce_xsum(xs, ldsl());
        acc_push(accs(p0) + kAcc0, xs);
        if (!sync(PK_ENERGY)) return false;
        mdy_xsum X = acc_read(accs(p0) + kAcc0);
        const int nqe = ld_i(d.qe_n);
        if (nqe > kQueueCap) {
            status = MDY_ERR_CAPACITY;
            return true;
        }
        if (__builtin_expect(nqe > 0, 0) && !energy_fallback(pos, d, nqe, qb)) return false;
        double v[kNFF], e[kNFF + 2];
        grid_partials_reduce<kNFF>(r, G, v, lds);
        energy_finish(X, v, nqe, d.qe, qeu(qb), e);
        if (blk == 0 && tid == 0) {
            for (int k = 0; k < kNFF; k++) out[k] = e[k];
            out[kNFF] = e[kNFF + 1];
        }
        E = e[kNFF];
        Eplain = e[kNFF + 1];
        return true;
    }

    // ---- quadrature forces for the pairs an accepted evaluation queued
    // (CylinderExclVolRepulsion.cpp:1001-1254, 1017-1022 retry with the first
    // cylinder moved off the plane). The slots are written by block 0.
    __attribute__((always_inline)) __device__ bool force_fallback(const PosAny &pos, const Eval &ev) {
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
                double T0 = 0, T1 = 0, T2 = 0, T3 = 0, T4 = 0, T5 = 0;
#pragma clang loop unroll(disable)
                for (int comp = 0; comp < 6; comp++) {
                    mdy_xsum X;
                    grid_reduce_xsum(A.spart + (size_t)comp * G * kRed, G, 0, X, ldsl());
                    const double t = mdy_xs_value(&X);
                    T0 = comp == 0 ? t : T0; T1 = comp == 1 ? t : T1; T2 = comp == 2 ? t : T2;
                    T3 = comp == 3 ? t : T3; T4 = comp == 4 ? t : T4; T5 = comp == 5 ? t : T5;
                }
                const double T[6] = {T0, T1, T2, T3, T4, T5};
                d3 f[4];
                simpson_f_assemble(g, T, A.D.ev_kr[i], N, f);
                const bool ok = finite_d3(f[0]) && finite_d3(f[1]) && finite_d3(f[2]) && finite_d3(f[3]);
                if (ok) {
                    if (blk == 0 && tid == 0) {
                        int s = A.D.slot_ev + 4 * i;
                        put3(slots(), s, f[0]); put3(slots(), s + 1, f[1]);
                        put3(slots(), s + 2, f[2]); put3(slots(), s + 3, f[3]);
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
    template <bool WIDE, bool BATCH>
    __attribute__((always_inline)) __device__ bool gather(double *Fout, const double *Fprev, bool move, const double *dvec, double lam, double &maxF,
                           double &FF, double &FFp) {
        const int p1 = ph;
        double m[1] = {0.0};
        mdy_xsum sff, sfp;
        mdy_xs_zero(&sff);
        mdy_xs_zero(&sfp);
        mark(0);
        for (int b = gtid; b < A.D.nb; b += gs) {
            d3 f = WIDE ? gather_bead_w<12, 24>(A.D, slots(), b) : gather_bead_w<8, 8>(A.D, slots(), b);
            const int o = 3 * b;
            Fout[o] = f.x; Fout[o + 1] = f.y; Fout[o + 2] = f.z;
            m[0] = fmax(m[0], fmax(fabs(f.x), fmax(fabs(f.y), fabs(f.z))));
#pragma clang loop unroll(disable)
            for (int k = 0; k < 3; k++) {
                const double fk = k == 0 ? f.x : k == 1 ? f.y : f.z;
                mdy_xs_add(&sff, fk * fk);
                if (Fprev) mdy_xs_add(&sfp, fk * Fprev[o + k]);
            }
            if (move) {
                A.x[o] = A.x[o] + lam * dvec[o];
                A.x[o + 1] = A.x[o + 1] + lam * dvec[o + 1];
                A.x[o + 2] = A.x[o + 2] + lam * dvec[o + 2];
            }
        }
        mark(1);
        block_reduce_xsum2_lds(sff, sfp, ldsl(), m);   // F.F and F.Fprev together, and max|F|
        unsigned long long *a = accs(p1);
        if (tid == 0) atomicMax(a + kAccMax, (unsigned long long)__double_as_longlong(m[0]));
        // the cluster's blocks all read the launch's completion count as one
        // block saw it, through this phase's accumulators
        if (BATCH && done_ctr && blk == 0 && tid == 0)
            atomicMax(a + kAccDone, (unsigned long long)(unsigned)ld_i(done_ctr));
        acc_push(a + kAcc0, sff);
        acc_push(a + kAcc1, sfp);
        if (!sync(PK_GATHER)) return false;
        acc_fetch(a, nullptr, nullptr);
        maxF = __longlong_as_double((long long)accl[kAccMax]);
        if (BATCH && done_ctr)
            done_seen = (int)__hip_atomic_load(a + kAccDone, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        double v2[2];
#pragma clang loop unroll(disable)
        for (int k = 0; k < 2; k++) {
            mdy_xsum Y = acc_local(k == 0 ? kAcc0 : kAcc1);
            v2[k] = mdy_xs_value(&Y);
        }
        FF = v2[0];
        FFp = Fprev ? v2[1] : 0.0;
        return true;
    }
};

// s = F + beta * d for every bead into Sn, and this thread's share of s.F
// into sd (reduced with the phase's energy, or on its own without a trial)
template <bool MIXED>
__device__ __attribute__((always_inline)) void spec_direction(Cg<MIXED> &K, const double *Fn, const double *dprev,
                                                              double beta, double *Sn, mdy_xsum &sd) {
    // thread t takes beads gs-1-t, 2gs-1-t, ...: the first threads carry the
    // most excluded-volume items in the sweep that follows in the same phase
    for (int b = K.gs - 1 - K.gtid; b < K.A.D.nb; b += K.gs) {
        const int o = 3 * b;
        for (int k = 0; k < 3; k++) {
            double s = Fn[o + k] + beta * dprev[o + k];
            Sn[o + k] = s;
            mdy_xs_add(&sd, s * Fn[o + k]);
        }
    }
}

// The minimizer as a state machine: every loop turn performs at most one kind
// of device work (each with a single call site, so each is inlined once) and
// then advances the CGMethod::minimize control flow
// (CGMethod.hpp:826-1743, backtrackingLineSearch CGMethod.cpp:43-360).
// The first trial of a line search that follows an accepted step is taken
// along s = F + beta*d while s itself is formed in the same phase; if s.F <= 0
// (resetCGParams) that trial is discarded and redone along F.
enum Op { OP_NONE, OP_EVAL_FF, OP_EVAL, OP_GATHER, OP_MINSTEP, OP_SDONLY, OP_RESTORE };
enum Stage {
    ST_INIT_FF, ST_INIT_F, ST_INIT_G, ST_ITER, ST_TRIAL, ST_SD, ST_QUAD, ST_QUAD_EVAL, ST_QUAD_DONE, ST_QUAD_AT,
    ST_LS_DONE, ST_ACCEPT, ST_REC_STEP, ST_REC_EVAL, ST_REC_GATHER, ST_FINAL, ST_FINAL_DONE, ST_FINISH
};

// One minimization run by a cluster of G blocks (blk = index within it).
// QUAD: compiled with the quadratic line-search stages (a separate kernel, so
// the backtracking kernel carries none of their state in registers).
// WIDE: the wide per-bead gather (one bead per thread). BATCH: the batched
// launch's passes ended by completions.
template <bool MIXED, bool QUAD, bool WIDE, bool BATCH>
__device__ __attribute__((always_inline)) void cg_run(const CgArgs &A, double *lds, int blk, int G,
                                                      int *done = nullptr, int suspend_done = 0) {
    __shared__ unsigned long long accl[24];
    Cg<MIXED> K(A, lds, accl, blk, G);
    if (BATCH) K.done_ctr = done;
    double *R = A.out;
    const bool leader = (K.blk == 0 && K.tid == 0);
    if (leader) R[R_T0] = (double)wall_clock64();
#define FB(i) ((i) ? A.F1 : A.F0)
#define SB(i) ((i) ? A.S1 : A.S0)
    unsigned flags = 0;
    int nEC = 0, nFC = 0, nIter = 0, nEE = 0, nSM = 0, nRec = 0;
    long long trn = 0;
    double gradInf = INFINITY, relE = INFINITY;
    bool after = true;   // compute energiesAfter / output force
    bool suspended = false;
    int cur = 0, sidx = 0;
    const double *dvec = nullptr;
    double maxF = 0, FF = 0, FFp = 0, FFcur = 0;
    double E0 = 0, prevEnergy = 0, lastFDotF = 0, slope = 0;
    int safeMode = 0;
    double minimumE = INFINITY;
    bool haveMin = false;
    double lam = 0, newE = 0, armijo = 0, reducedSlope = 0, sdF = 0, stepRec = 0, beta = 0;
    // quadratic line search (quadraticLineSearchV2): trial count, the last
    // rejected trial, the first record, the highest later record, the
    // parabola's bracket, its refinement count, and the last evaluated step
    bool quadLS = false;
    int qIter = 0, qRef = 0;
    double qPrevL = 0, qPrevE = 0, qL0 = 0, qE0 = 0, qMax = 0, qBt = 0;
    double qx0 = 0, qx1 = 0, qx2 = 0, qy0 = 0, qy1 = 0, qy2 = 0, qLam = 0;
    double lamEval = 0, Eacc = 0;    // step of the last trial and the energy at the accepted point
    bool pendingSpec = false;        // an accepted step's s is not formed yet
    const double *dprev = nullptr;   // the direction s extends
    double *Snew = nullptr;          // where s goes
    Eval ev{0.0, 0.0, 0, 0};
    double Eff = 0, Effp = 0;
    // the device work of the next turn and its operands
    int op = OP_EVAL_FF, stage = ST_INIT_FF;
    PosAny P = PosAny::at(A.x);
    double *ffout = R + R_EB;
    double *gOut = nullptr;
    const double *gPrev = nullptr;
    bool gMove = false;
    auto callLimit = [&]() {
        return (A.energy_call_limit != 0 && nEC > A.energy_call_limit) ||
               (A.force_call_limit != 0 && nFC > A.force_call_limit);
    };
    auto converged = [&]() { return gradInf < A.grad_tol || relE < A.energy_rel_tol; };
    // Suspension: everything the iterations carry from one to the next; the
    // vectors (x, forces, directions, minimum-energy backup, trace) stay in
    // the network's buffers. The phase counter restarts at 0 with cleared
    // accumulators, which changes no value.
    auto bufid = [&](const double *q) {
        return q == A.F0 ? 0.0 : q == A.F1 ? 1.0 : q == A.S0 ? 2.0 : q == A.S1 ? 3.0 : -1.0;
    };
    auto bufptr = [&](double id) -> double * {
        return id == 0.0 ? A.F0 : id == 1.0 ? A.F1 : id == 2.0 ? A.S0 : id == 3.0 ? A.S1 : nullptr;
    };
    if (leader) R[R_SUSP] = 0.0;
    if (A.resume) {
        const double *S = R + R_RS;
        flags = (unsigned)S[0]; nEC = (int)S[1]; nFC = (int)S[2]; nIter = (int)S[3]; nEE = (int)S[4];
        nSM = (int)S[5]; nRec = (int)S[6]; trn = (long long)S[7]; gradInf = S[8]; relE = S[9];
        cur = (int)S[10]; sidx = (int)S[11]; dvec = bufptr(S[12]); maxF = S[13]; FFcur = S[14]; E0 = S[15];
        prevEnergy = S[16]; lastFDotF = S[17]; slope = S[18]; safeMode = (int)S[19]; minimumE = S[20];
        haveMin = S[21] != 0.0; beta = S[22]; pendingSpec = S[23] != 0.0; dprev = bufptr(S[24]);
        Snew = bufptr(S[25]);
        op = OP_NONE; stage = ST_ITER;
    }
    auto trace = [&](double l, double e, double b) {
        if (leader && trn < A.tr_cap) {
            A.tr_lam[trn] = l; A.tr_e[trn] = e; A.tr_mf[trn] = gradInf; A.tr_beta[trn] = b;
            A.tr_safe[trn] = safeMode;
        }
        trn++;
    };

    for (;;) {
        // ---- device work
        bool ok = true;
        switch (op) {
        case OP_EVAL_FF:
            ok = K.eval_ff(P, ffout, Eff, Effp);
            break;
        case OP_EVAL:
        case OP_SDONLY: {
            const int p1 = K.ph;
            mdy_xsum sd;
            mdy_xs_zero(&sd);
            if (P.mode == PosAny::SPEC) spec_direction(K, P.F, P.d, P.beta, Snew, sd);
            if (op == OP_EVAL) {
                ok = K.eval(P, ev, P.mode == PosAny::SPEC, sd);
                nEE++;
            } else {   // s.F only (max|F| = 0: no trial)
                block_reduce_xsum(sd, K.ldsl());
                K.acc_push(K.accs(p1) + kAcc1, sd);
                ok = K.sync(PK_REDUCE);
                if (ok) sdF = K.acc_value(K.accs(p1) + kAcc1);
            }
            break;
        }
        case OP_GATHER:
            if (__builtin_expect(ev.nqf != 0, 0)) ok = K.force_fallback(P, ev);
            if (ok && !K.status) ok = K.template gather<WIDE, BATCH>(gOut, gPrev, gMove, P.d, P.lam, maxF, FF, FFp);
            break;
        case OP_MINSTEP: {
            // MController::moveAlongDirectionLimitingBeadMove_: step = min maxDist/|d_i|
            double v[1] = {INFINITY};
            for (int o = K.gtid; o < A.D.ndof; o += K.gs) v[0] = smin(v[0], A.max_dist / fabs(dvec[o]));
            block_reduce<1>(v, lds, 0u, 0x1u);
            const int p1 = K.ph;
            if (K.tid == 0) K.red(p1)[(size_t)K.blk * kRed] = v[0];
            ok = K.sync(PK_REDUCE);
            if (ok) {
                grid_partials_reduce<1>(K.red(p1), K.G, v, lds, 0u, 0x1u);
                stepRec = v[0];
            }
            break;
        }
        case OP_RESTORE:
            for (int o = K.gtid; o < A.D.ndof; o += K.gs) A.x[o] = A.xmin[o];
            ok = K.sync(PK_RESTORE);
            break;
        default:
            break;
        }
        if (__builtin_expect(!ok, 0)) goto aborted;
        op = OP_NONE;
        if (__builtin_expect(K.status != 0, 0) && stage != ST_FINISH) { after = false; stage = ST_FINISH; }

        // ---- control
    control:
        switch (stage) {
        case ST_INIT_FF:   // energiesBefore = computeIndividualEnergies(x); E(x) for the first line search
            nEC++; nEE++;
            E0 = Eff; prevEnergy = Effp;
            op = OP_EVAL; P = PosAny::at(A.x); stage = ST_INIT_F;
            break;
        case ST_INIT_F:    // forces at x
            op = OP_GATHER; gOut = FB(cur); gPrev = nullptr; gMove = false; stage = ST_INIT_G;
            break;
        case ST_INIT_G:
            nFC++;
            gradInf = maxF;
            if (__builtin_expect(!isfinite(gradInf), 0)) { K.status = MDY_ERR_NONFINITE; after = false; stage = ST_FINISH; break; }
            FFcur = FF;
            dvec = FB(cur); lastFDotF = FF; slope = FF;   // resetCGParams
            stage = ST_ITER;
            break;
        case ST_ITER:
            if (callLimit() || converged()) {
                if (callLimit() && haveMin && A.report_call_limit) op = OP_RESTORE;
                stage = ST_FINAL;
                break;
            }
            if ((A.suspend_calls > 0 && nFC >= A.suspend_calls) ||
                (BATCH && suspend_done > 0 && K.done_seen >= suspend_done)) {   // this pass ends; the next resumes here
                if (leader) {
                    double *S = R + R_RS;
                    S[0] = flags; S[1] = nEC; S[2] = nFC; S[3] = nIter; S[4] = nEE; S[5] = nSM; S[6] = nRec;
                    S[7] = (double)trn; S[8] = gradInf; S[9] = relE; S[10] = cur; S[11] = sidx; S[12] = bufid(dvec);
                    S[13] = maxF; S[14] = FFcur; S[15] = E0; S[16] = prevEnergy; S[17] = lastFDotF; S[18] = slope;
                    S[19] = safeMode; S[20] = minimumE; S[21] = haveMin ? 1.0 : 0.0; S[22] = beta;
                    S[23] = pendingSpec ? 1.0 : 0.0; S[24] = bufid(dprev); S[25] = bufid(Snew);
                    R[R_SUSP] = 1.0;
                }
                suspended = true;
                after = false;
                stage = ST_FINISH;
                break;
            }
            nIter++;
            armijo = safeMode ? 0.0 : 0.4;
            quadLS = QUAD && A.line_search == MDY_LS_QUADRATIC && !safeMode;
            qIter = 0;
            qMax = -INFINITY;
            lam = 0.0; newE = 0.0;
            if (maxF != 0.0) {
                lam = smin(A.lambda_max, A.max_dist / maxF);
                nEC++;   // currentEnergy = computeEnergy(x): E0, known bitwise from the accepting trial
                op = OP_EVAL;
                P = pendingSpec ? PosAny::spec(A.x, FB(cur), dprev, beta, lam) : PosAny::line(A.x, dvec, lam);
                reducedSlope = armijo * slope;
                stage = ST_TRIAL;
            } else if (pendingSpec) {
                op = OP_SDONLY; P = PosAny::spec(A.x, FB(cur), dprev, beta, 0.0); stage = ST_SD;
            } else {
                stage = ST_LS_DONE;
            }
            break;
        case ST_SD:        // s.F without a trial (max|F| = 0)
            pendingSpec = false;
            if (sdF <= 0) { dvec = FB(cur); lastFDotF = FFcur; slope = FFcur; }
            else { dvec = Snew; slope = sdF; }
            stage = ST_LS_DONE;
            break;
        case ST_TRIAL: {   // one backtracking step
            if (P.mode == PosAny::SPEC) {
                pendingSpec = false;
                sdF = ev.xa;
                if (sdF <= 0) {   // resetCGParams: redo this trial along F
                    dvec = FB(cur); lastFDotF = FFcur; slope = FFcur;
                    reducedSlope = armijo * slope;
                    op = OP_EVAL; P = PosAny::line(A.x, dvec, lam);
                    break;
                }
                dvec = Snew; slope = sdF;
                reducedSlope = armijo * slope;
            }
            double Et = ev.E;
            nEC++;
            if (quadLS) {   // CGMethod.hpp:567-617: Armijo slope 0.4, every trial recorded
                lamEval = lam; Eacc = Et;
                qIter++;
                const double idealq = -0.4 * lam * slope;
                if (idealq > 0) { K.status = MDY_ERR_QUADRATIC; after = false; stage = ST_FINISH; break; }
                bool conv = false;
                if (Et - E0 <= idealq) conv = true;
                else { qPrevL = lam; qPrevE = Et; lam *= 0.5; }
                double recL, recE;
                if (lam <= 0.0 || lam <= 1e-8) { conv = true; lam = 0.0; recL = 0.0; recE = E0; }
                else if (conv) { recL = lam; recE = Et; }
                else { recL = qPrevL; recE = qPrevE; }
                if (qIter == 1) { qL0 = recL; qE0 = recE; }
                else qMax = fmax(qMax, recE);
                if (!conv) { op = OP_EVAL; P = PosAny::line(A.x, dvec, lam); break; }
                newE = Et;
                // CGMethod.hpp:621-645: refine when an earlier record lies above
                // the accepted energy; the third point is the first record
                if (lam > 0.0 && qIter > 1 && Et < E0 && Et < qMax) {
                    qx0 = 0.0; qx1 = lam; qx2 = qL0; qy0 = E0; qy1 = Et; qy2 = qE0;
                    qBt = lam; qRef = 0;
                    stage = ST_QUAD;
                } else {
                    stage = ST_LS_DONE;
                }
                break;
            }
            const double ideal = -reducedSlope * lam;
            bool done = false;
            if (lam < 0 || ideal > 0) { done = true; lam = 0.0; }
            const double change = Et - E0;
            if (change <= ideal) done = true;
            else lam *= 0.5;
            if (lam <= 0.0 || lam <= 1e-8) { done = true; lam = 0.0; Et = E0; }
            if (!done) { op = OP_EVAL; P = PosAny::line(A.x, dvec, lam); break; }
            newE = Et;
            if (lam != 0.0) {   // accepted: gather at once (ST_LS_DONE's accepting branch)
                op = OP_GATHER; P = PosAny::line(A.x, dvec, lam);
                gOut = FB(1 - cur); gPrev = FB(cur); gMove = true;
                stage = ST_ACCEPT;
            } else {
                stage = ST_LS_DONE;
            }
            break;
        }
        case ST_QUAD: {   // quadraticoptimization, CGMethod.hpp:663-765 (at most 200 refinements)
            if constexpr (!QUAD) break;
            if (qRef >= 200) { qLam = qx1; stage = ST_QUAD_DONE; break; }
            qRef++;
            const double d = (qx0 - qx1) * (qx0 - qx2) * (qx1 - qx2);
            const double a = qy0 * (qx1 - qx2) - qy1 * (qx0 - qx2) + qy2 * (qx0 - qx1);
            const double b = -(qy0 * (qx1 + qx2) * (qx1 - qx2) - qy1 * (qx0 + qx2) * (qx0 - qx2) +
                               qy2 * (qx0 + qx1) * (qx0 - qx1));
            qLam = -b / (2 * a);
            if (fabs(qLam - qx1) / qx1 < 1e-3) { stage = ST_QUAD_DONE; break; }
            if (a / d < 0) { qLam = qBt; stage = ST_QUAD_DONE; break; }
            op = OP_EVAL; P = PosAny::line(A.x, dvec, qLam); stage = ST_QUAD_EVAL;
            break;
        }
        case ST_QUAD_EVAL: {
            if constexpr (!QUAD) break;
            const double eq = ev.E;
            nEC++;
            lamEval = qLam; Eacc = eq;
            if (fabs(eq - qy1) / qy1 < 1e-3) { stage = ST_QUAD_DONE; break; }
            if (qLam < qx1) {
                if (eq < qy0 && eq < qy1) { qx1 = qLam; qy1 = eq; }
                else { qx0 = qLam; qy0 = eq; }
            } else if (qx1 < qLam) {
                if (eq < qy1 && eq < qy2) { qx1 = qLam; qy1 = eq; }
                else { qx2 = qLam; qy2 = eq; }
            } else {
                qLam = qx1; stage = ST_QUAD_DONE; break;
            }
            stage = ST_QUAD;
            break;
        }
        case ST_QUAD_DONE:   // newE stays the backtracking energy (CGMethod.hpp:655)
            if constexpr (!QUAD) break;
            lam = (qLam <= 0.0 || qLam <= 1e-8) ? 0.0 : qLam;
            if (lam != 0.0 && lam != lamEval) {   // forces and energy at the accepted point
                op = OP_EVAL; P = PosAny::line(A.x, dvec, lam); stage = ST_QUAD_AT;
                break;
            }
            stage = ST_LS_DONE;
            break;
        case ST_QUAD_AT:
            if constexpr (!QUAD) break;
            Eacc = ev.E;
            stage = ST_LS_DONE;
            break;
        case ST_LS_DONE:
            if (lam == 0.0) {
                if (!A.try_recover) {
                    flags |= MDY_ERROR_LINESEARCH;
                    trace(0.0, newE, 0.0);
                    after = false;   // the reference returns before energiesAfter / ptrOutForce
                    stage = ST_FINISH;
                } else if (safeMode) {
                    op = OP_MINSTEP; stage = ST_REC_STEP;
                } else {
                    dvec = FB(cur); lastFDotF = FFcur; slope = FFcur;
                    safeMode = 1;
                    nSM++;
                    trace(0.0, newE, 0.0);
                    stage = ST_ITER;
                }
            } else {   // accepted: the force slots hold the forces at x + lam*d
                op = OP_GATHER; P = PosAny::line(A.x, dvec, lam);
                gOut = FB(1 - cur); gPrev = FB(cur); gMove = true;
                stage = ST_ACCEPT;
            }
            break;
        case ST_REC_STEP:
            if (stepRec == INFINITY) { K.status = MDY_ERR_RECOVERY; after = false; stage = ST_FINISH; break; }
            op = OP_EVAL; P = PosAny::line(A.x, dvec, stepRec); stage = ST_REC_EVAL;
            break;
        case ST_REC_EVAL:
            op = OP_GATHER; gOut = FB(1 - cur); gPrev = FB(cur); gMove = true; stage = ST_REC_GATHER;
            break;
        case ST_REC_GATHER:
            nFC++; nRec++;
            gradInf = maxF;
            if (__builtin_expect(!isfinite(gradInf), 0)) { K.status = MDY_ERR_NONFINITE; after = false; stage = ST_FINISH; break; }
            cur = 1 - cur;
            FFcur = FF;
            dvec = FB(cur); lastFDotF = FF; slope = FF;
            safeMode = 0;
            E0 = ev.E;   // computeEnergy at the moved coordinates
            trace(0.0, newE, 0.0);
            stage = ST_ITER;
            break;
        case ST_ACCEPT: {
            nFC++;
            gradInf = maxF;
            if (__builtin_expect(!isfinite(gradInf), 0)) { K.status = MDY_ERR_NONFINITE; after = false; stage = ST_FINISH; break; }
            E0 = quadLS ? Eacc : newE;   // E(x) of the next line search
            relE = fabs(newE - prevEnergy) / fabs(prevEnergy);
            prevEnergy = newE;
            const double newGrad = FF, prevGrad = FFp;
            beta = 0.0;
            if (A.method == MDY_CG_FLETCHER_RIEVES) beta = newGrad / lastFDotF;
            else if (A.method == MDY_CG_POLAK_RIBIERE) beta = smax(0.0, (newGrad - prevGrad) / lastFDotF);
            cur = 1 - cur;
            FFcur = newGrad;
            dprev = dvec; Snew = SB(1 - sidx); sidx = 1 - sidx;
            pendingSpec = true;   // s = F + beta*d is formed with the next trial
            // forceBelowRelaxedTolerance(10) && numIter > forceCallLimit/2 -> keep lowest-energy coordinates
            if (maxF <= 10 * A.grad_tol && nIter > A.force_call_limit / 2 && newE <= minimumE) {
                minimumE = newE;
                haveMin = true;
                for (int o = K.gtid; o < A.D.ndof; o += K.gs) A.xmin[o] = A.x[o];
            }
            if (A.method != MDY_CG_STEEPEST) lastFDotF = newGrad;
            trace(lam, newE, beta);
            stage = ST_ITER;
            goto control;   // no device work in between
        }
        case ST_FINAL:
            if (after) { op = OP_EVAL_FF; P = PosAny::at(A.x); ffout = R + R_EA; stage = ST_FINAL_DONE; }
            else stage = ST_FINISH;
            break;
        case ST_FINAL_DONE:
            nEC++; nEE++;
            stage = ST_FINISH;
            break;
        default:
            break;
        }
        if (stage == ST_FINISH && op == OP_NONE) break;
    }
    if (leader) {
        R[R_FLAGS] = flags; R[R_NEC] = nEC; R[R_NFC] = nFC; R[R_NIT] = nIter; R[R_GIN] = gradInf; R[R_ERC] = relE;
        R[R_NEE] = nEE; R[R_NSM] = nSM; R[R_NREC] = nRec; R[R_STATUS] = K.status; R[R_TRN] = (double)trn;
        R[R_LAST] = after ? 1.0 : 0.0; R[R_FOUT] = (double)cur; R[R_NPH] = (double)K.ph;
        // the reference shifts the direction right after an accepted step; here
        // that shift is pending until the next trial
        R[R_DIR] = bufid(dvec); R[R_PEND] = pendingSpec ? 1.0 : 0.0; R[R_BETA] = beta; R[R_DPREV] = bufid(dprev);
        R[R_T1] = (double)wall_clock64();
        if (BATCH && done && !suspended) atomicAdd(done, 1);
    }
    return;
aborted:
    if (leader) R[R_STATUS] = MDY_ERR_TIMEOUT;
#undef FB
#undef SB
}

// The arguments come by value: pointers read from the kernel-argument segment
// are known to be global, so every access is a global (not flat) access.
// LDS: the shuffle-free exact reductions need block_combine_words(17) words
constexpr int kCgLds = block_combine_words(17);
static_assert(kCgLds >= 17 * 16, "LDS of the other block reductions");

template <bool MIXED, bool QUAD, bool WIDE>
__global__ void __launch_bounds__(kBlock) k_cg(const CgArgs Av) {
    __shared__ double lds[kCgLds];
    cg_run<MIXED, QUAD, WIDE, false>(Av, lds, blockIdx.x, gridDim.x);
}

// Batched form: independent minimizations in one cooperative launch. Block b
// belongs to network map[b].x as its block map[b].y; each network's cluster
// runs its own state machine with its own barrier, reductions and buffers.
// Pointers loaded from generic memory are flat to the compiler; loaded from
// the constant address space (as kernel arguments are) they are known to be
// global, so the arguments are read through a constant-space view and every
// access through them uses global instructions, as in the single-network
// kernel.
// done (nullable): the launch's count of finished minimizations; each
// suspends at the top of an iteration once the count reaches suspend_done (>0).
template <bool MIXED, bool QUAD>
__global__ void __launch_bounds__(kBlock) k_cg_batch(const CgArgs *__restrict__ all, const int2 *__restrict__ map,
                                                     int *done, int suspend_done) {
    __shared__ double lds[kCgLds];
    const int2 m = map[blockIdx.x];
#if defined(__HIP_DEVICE_COMPILE__)
    typedef __attribute__((address_space(4))) const CgArgs ConstArgs;
    CgArgs A = ((ConstArgs *)all)[m.x];
#else
    CgArgs A = all[m.x];
#endif
    cg_run<MIXED, QUAD, false, true>(A, lds, m.y, A.grid, done, suspend_done);
}

template <bool WIDE>
static const void *cg_kernel_w(bool mixed, bool quad) {
    if (mixed) return quad ? reinterpret_cast<const void *>(k_cg<true, true, WIDE>)
                           : reinterpret_cast<const void *>(k_cg<true, false, WIDE>);
    return quad ? reinterpret_cast<const void *>(k_cg<false, true, WIDE>)
                : reinterpret_cast<const void *>(k_cg<false, false, WIDE>);
}
static const void *cg_kernel(bool mixed, bool quad, bool wide) {
    return wide ? cg_kernel_w<true>(mixed, quad) : cg_kernel_w<false>(mixed, quad);
}
static const void *cg_batch_kernel(bool mixed, bool quad) {
    if (mixed) return quad ? reinterpret_cast<const void *>(k_cg_batch<true, true>)
                           : reinterpret_cast<const void *>(k_cg_batch<true, false>);
    return quad ? reinterpret_cast<const void *>(k_cg_batch<false, true>)
                : reinterpret_cast<const void *>(k_cg_batch<false, false>);
}

// ---------------------------------------------------------------------------
// Buffers and arguments of one minimization run by a cluster of G blocks on
// stream st (the counters and flags it relies on are cleared there).
static int prepare_cg(Ctx &c, const mdy_cg_params *p, double *d_x, int G, hipStream_t st, CgArgs &A) {
    const int ndof = 3 * c.nb;
    MDY_TRY(ensure(c.F0, ndof)); MDY_TRY(ensure(c.F1, ndof));
    MDY_TRY(ensure(c.S0, ndof)); MDY_TRY(ensure(c.S1, ndof));
    MDY_TRY(ensure(c.xmin, ndof));
    MDY_TRY(ensure(c.prof, 5 * (size_t)c.prof_cap + 5));
    MDY_TRY(ensure(c.prof_kind, c.prof_cap + 1));
    MDY_TRY(ensure(c.cg_out, R_SIZE));
    MDY_TRY(ensure(c.qe, 3 * kQueueCap)); MDY_TRY(ensure(c.qf, 3 * kQueueCap));
    MDY_TRY(ensure(c.qn, 16));
    MDY_TRY(ensure(c.qe_u, 3 * kQueueCap));
    MDY_TRY(ensure(c.tr_lam, c.trace_cap + 1)); MDY_TRY(ensure(c.tr_e, c.trace_cap + 1));
    MDY_TRY(ensure(c.tr_mf, c.trace_cap + 1)); MDY_TRY(ensure(c.tr_beta, c.trace_cap + 1));
    MDY_TRY(ensure(c.tr_safe, c.trace_cap + 1));
    MDY_TRY(ensure(c.bar, 64 + 16 * (size_t)G));
    MDY_TRY(ensure(c.acc, 3 * (size_t)kAccWords));
    MDY_TRY(ensure(c.red, (size_t)3 * G * kRed));
    MDY_TRY(ensure(c.simpson_part, (size_t)6 * G * kRed));
    MDY_HIP(hipMemsetAsync(c.bar.p, 0, (64 + 16 * (size_t)G) * sizeof(unsigned long long), st));
    MDY_HIP(hipMemsetAsync(c.qn.p, 0, 16 * sizeof(int), st));
    MDY_HIP(hipMemsetAsync(c.acc.p, 0, 3 * kAccWords * sizeof(unsigned long long), st));
    MDY_HIP(hipMemsetAsync(c.err.p, 0, sizeof(int), st));
    if (c.prof_cap > 0) MDY_HIP(hipMemsetAsync(c.prof.p, 0, (5 * (size_t)c.prof_cap + 5) * sizeof(long long), st));
    A.D = c.dev();
    A.x = d_x; A.F0 = c.F0.p; A.F1 = c.F1.p; A.S0 = c.S0.p; A.S1 = c.S1.p; A.xmin = c.xmin.p;
    A.contrib = c.contrib.p; A.red = c.red.p; A.qe_u = c.qe_u.p; A.spart = c.simpson_part.p;
    A.qe = c.qe.p; A.qf = c.qf.p; A.qn = c.qn.p; A.bar = c.bar.p; A.acc = c.acc.p; A.out = c.cg_out.p;
    A.tr_lam = c.tr_lam.p; A.tr_e = c.tr_e.p; A.tr_mf = c.tr_mf.p; A.tr_beta = c.tr_beta.p; A.tr_safe = c.tr_safe.p;
    A.tr_cap = c.trace_cap;
    A.prof = c.prof.p; A.prof_kind = c.prof_kind.p; A.prof_cap = c.prof_cap;
    A.method = p->method; A.line_search = p->line_search; A.try_recover = p->try_to_recover; A.energy_call_limit = p->energy_call_limit;
    A.force_call_limit = p->force_call_limit > 3 * p->n_beads_limit ? p->force_call_limit : 3 * p->n_beads_limit;
    A.max_dist = p->max_dist; A.lambda_max = p->lambda_max; A.grad_tol = p->grad_tol;
    A.energy_rel_tol = p->energy_rel_tol;
    A.report_call_limit = p->report_on_call_limit;
    A.spin_limit = 20000000LL;   // polls (about 10 s) before the watchdog fires
    A.grid = G;
    A.suspend_calls = 0;
    A.resume = 0;
    return MDY_OK;
}

// MinimizationResult from a result record
static void collect(Ctx &c, const double *R, const mdy_cg_params *p, const CgArgs &A, mdy_min_result *res) {
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
    c.have_dir = res->status == MDY_OK;
    c.dir_buf = (int)R[R_DIR]; c.dir_pend = R[R_PEND] != 0.0; c.dir_beta = R[R_BETA];
    c.dir_prev = (int)R[R_DPREV]; c.dir_cur = (int)R[R_FOUT];
}

// s = F + beta d and the exact s.F, in one block (cgShiftSearchDir and
// resetCGParams, CGMethod.cpp:325-360: the same expression and the same exact
// sum as the kernel's direction update)
__global__ void __launch_bounds__(1024) k_final_dir(const double *__restrict__ F, const double *__restrict__ d,
                                                     double beta, int n, double *__restrict__ s, double *sf) {
    __shared__ long long lds[136];
    mdy_xsum x;
    mdy_xs_zero(&x);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const double v = F[i] + beta * d[i];
        s[i] = v;
        mdy_xs_add(&x, v * F[i]);
    }
    block_reduce_xsum(x, lds);
    if (threadIdx.x == 0) sf[0] = mdy_xs_value(&x);
}

int final_search_dir(Ctx &c, double *h_out) {
    if (!c.have_dir) return MDY_ERR_STATE;
    hipStream_t st = c.stream;
    const size_t n = 3 * (size_t)c.nb;
    auto buf = [&](int id) -> const double * {
        return id == 0 ? c.F0.p : id == 1 ? c.F1.p : id == 2 ? c.S0.p : id == 3 ? c.S1.p : nullptr;
    };
    const double *src = buf(c.dir_buf);
    if (c.dir_pend) {
        const double *F = buf(c.dir_cur), *d = buf(c.dir_prev);
        if (!F || !d) return MDY_ERR_STATE;
        MDY_TRY(ensure(c.xt, n));
        MDY_TRY(ensure(c.out, 8));
        hipLaunchKernelGGL(k_final_dir, dim3(1), dim3(1024), 0, st, F, d, c.dir_beta, (int)n, c.xt.p, c.out.p);
        MDY_HIP(hipGetLastError());
        double sf = 0;
        MDY_HIP(hipMemcpyAsync(&sf, c.out.p, sizeof(double), hipMemcpyDeviceToHost, st));
        MDY_HIP(hipStreamSynchronize(st));
        src = (sf <= 0) ? F : c.xt.p;   // resetCGParams when s.F <= 0
    }
    if (!src) return MDY_ERR_STATE;
    MDY_HIP(hipMemcpyAsync(h_out, src, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    MDY_HIP(hipStreamSynchronize(st));
    return MDY_OK;
}

static int copy_force_out(Ctx &c, const double *R, double *d_force_out, hipStream_t st) {
    if (d_force_out && R[R_LAST] != 0.0) {
        const double *src = ((int)R[R_FOUT] == 0) ? c.F0.p : c.F1.p;
        MDY_HIP(hipMemcpyAsync(d_force_out, src, sizeof(double) * 3 * (size_t)c.nb, hipMemcpyDeviceToDevice, st));
    }
    return MDY_OK;
}

// two timing events, destroyed on every return path
struct EventPair {
    hipEvent_t a = nullptr, b = nullptr;
    ~EventPair() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

int run_minimize(Ctx &c, const mdy_cg_params *p, double *d_x, double *d_force_out, mdy_min_result *res) {
    hipStream_t st = c.stream;
    // persistent grid: all blocks resident
    int per_cu = 0;
    const bool mixed = c.opts.precision == MDY_PREC_MIXED;
    const bool quad = p->line_search == MDY_LS_QUADRATIC;
    const void *kern = cg_kernel(mixed, quad, false);
    MDY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, kBlock, 0));
    if (per_cu < 1) return MDY_ERR_LAUNCH;
    if (c.opts.blocks_per_cu > 0 && c.opts.blocks_per_cu < per_cu) per_cu = c.opts.blocks_per_cu;
    const int G = per_cu * c.num_cu;
    if ((long long)c.nb <= (long long)G * kBlock) {   // one bead per thread in the gather: the wide form
        int per_cu_w = 0;
        const void *kw = cg_kernel(mixed, quad, true);
        MDY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu_w, kw, kBlock, 0));
        if (per_cu_w >= per_cu) kern = kw;
    }
    c.coop_blocks = G;
    CgArgs A;
    MDY_TRY(prepare_cg(c, p, d_x, G, st, A));
    EventPair ev;
    MDY_HIP(hipEventCreate(&ev.a));
    MDY_HIP(hipEventCreate(&ev.b));
    MDY_HIP(hipEventRecord(ev.a, st));
    void *args[] = {&A};
    if (hipLaunchCooperativeKernel(kern, dim3(G), dim3(kBlock), args, 0, st) != hipSuccess) return MDY_ERR_LAUNCH;
    MDY_HIP(hipEventRecord(ev.b, st));
    double R[R_SIZE];
    MDY_HIP(hipMemcpyAsync(R, c.cg_out.p, sizeof(R), hipMemcpyDeviceToHost, st));
    MDY_HIP(hipStreamSynchronize(st));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ev.a, ev.b);
    c.last_ms = ms;
    collect(c, R, p, A, res);
    MDY_TRY(copy_force_out(c, R, d_force_out, st));
    MDY_HIP(hipStreamSynchronize(st));
    return res->status == MDY_ERR_TIMEOUT ? MDY_ERR_TIMEOUT : MDY_OK;
}

// Batched minimizations: the resident blocks (threads_per_block threads each)
// are split into one cluster per network, in proportion to its weight (by
// default its per-evaluation work, interactions + 2 x beads), at least one
// block each; networks beyond the resident block count run in further
// launches. With first_pass_calls > 0 the minimizations run in passes: pass
// k suspends every minimization at the top of its first iteration after
// first_pass_calls * 4^(k-1) force calls (one below its own limit and not
// converged), saving its state; the next pass resumes it exactly there, with
// all the blocks shared among the networks still running; a pass in which
// every network gets at least kUncappedShare blocks runs uncapped. A
// minimization's trajectory does not depend on its cluster size, so the
// results equal those of one uncapped pass; the passes only move the blocks
// of finished networks to the slow ones. With first_pass_calls = -k < 0 the
// passes end by completions instead: a launch of m > 1 networks counts its
// finished minimizations on the device, and once max(1, m / k) have finished
// the others suspend at the top of their next iteration (each reads the
// count once per iteration); the next pass re-splits all the blocks among
// them, down to one network on the whole grid. cs[i] owns network i's
// tables, neighbor lists and CG buffers; cs[0] the launch arrays.
constexpr int kUncappedShare = 64;   // blocks per network at which a pass runs uncapped
constexpr int kCapGrowth = 4;        // force-call cap growth from one pass to the next (x2 measured slower)
constexpr int kMaxClusterThreads = 8192;    // threads of one network's cluster at most
int run_minimize_batch(Ctx **cs, int n, const mdy_cg_params *p, double *const *d_x, double *const *d_force_out,
                       mdy_min_result *res, double *net_ms, int threads_per_block, const double *weights,
                       int first_pass_calls, hipStream_t st, double *total_ms) {
    Ctx &c0 = *cs[0];
    const bool mixed = c0.opts.precision == MDY_PREC_MIXED;
    bool quad = false;
    for (int i = 0; i < n; i++) quad = quad || p[i].line_search == MDY_LS_QUADRATIC;
    const void *kern = cg_batch_kernel(mixed, quad);
    const int tpb = threads_per_block;
    int per_cu = 0;
    MDY_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, tpb, 0));
    if (per_cu < 1) return MDY_ERR_LAUNCH;
    const int Gmax = per_cu * c0.num_cu;
    std::vector<char> resume(n, 0);   // suspended in an earlier pass
    EventPair ev;
    MDY_HIP(hipEventCreate(&ev.a));
    MDY_HIP(hipEventCreate(&ev.b));
    hipEvent_t e0 = ev.a, e1 = ev.b;
    double kernel_ms = 0;
    double t_origin = -1.0;   // device clock at the first launch's start
    std::vector<CgArgs> args;
    std::vector<int2> map;
    std::vector<int> gsz, todo(n), next;
    for (int i = 0; i < n; i++) todo[i] = i;
    long long cap = first_pass_calls > 0 ? first_pass_calls : 0;
    const int by_done = first_pass_calls < 0 ? -first_pass_calls : 0;
    int rc = MDY_OK;
    if (by_done) rc = ensure(c0.batch_done, 1);
    while (!todo.empty() && rc == MDY_OK) {
        next.clear();
        const int nt = (int)todo.size();
        // few networks left: each gets enough blocks that a larger cluster
        // would not be faster, so this pass runs them to the end
        if ((long long)nt * kUncappedShare <= Gmax) cap = 0;
        for (int lo = 0; lo < nt && rc == MDY_OK; lo += Gmax) {
            const int m = (nt - lo < Gmax) ? nt - lo : Gmax;
            // cluster sizes: proportional shares of the blocks left after one each
            double tot = 0;
            std::vector<double> w(m);
            for (int i = 0; i < m; i++) {
                const int k = todo[lo + i];
                const Ctx &ci = *cs[k];
                w[i] = weights ? weights[k] : (double)ci.dev().n_items() + 2.0 * ci.nb;
                if (!(w[i] > 0.0)) w[i] = 0.0;
                tot += w[i];
            }
            gsz.assign(m, 1);
            const int spare = Gmax - m;
            int used = 0;
            std::vector<std::pair<double, int>> frac;
            for (int i = 0; i < m; i++) {
                const double share = tot > 0 ? spare * w[i] / tot : 0.0;
                const int whole = (int)share;
                gsz[i] += whole;
                used += whole;
                frac.push_back({share - whole, i});
            }
            std::sort(frac.begin(), frac.end(),
                      [](const std::pair<double, int> &a, const std::pair<double, int> &b) {
                          return a.first > b.first || (a.first == b.first && a.second < b.second);
                      });
            for (int k = 0; k < spare - used && k < m; k++) gsz[frac[k].second]++;
            // a cluster's barrier slows down with its block count: beyond
            // kMaxClusterThreads threads a network gets no faster
            const int gcap = std::max(1, kMaxClusterThreads / tpb);
            for (int i = 0; i < m; i++) gsz[i] = std::min(gsz[i], gcap);
            args.assign(m, CgArgs{});
            map.clear();
            for (int i = 0; i < m && rc == MDY_OK; i++) {
                const int k = todo[lo + i];
                rc = prepare_cg(*cs[k], &p[k], d_x[k], gsz[i], st, args[i]);
                args[i].resume = resume[k];
                if (cap > 0 && cap < (long long)args[i].force_call_limit) args[i].suspend_calls = (int)cap;
                for (int b = 0; b < gsz[i]; b++) map.push_back(make_int2(i, b));
            }
            if (rc != MDY_OK) break;
            const int G = (int)map.size();
            rc = ensure(c0.batch_args, (size_t)m * sizeof(CgArgs));
            if (rc == MDY_OK) rc = ensure(c0.batch_map, (size_t)G);
            if (rc != MDY_OK) break;
            if (hipMemcpyAsync(c0.batch_args.p, args.data(), (size_t)m * sizeof(CgArgs), hipMemcpyHostToDevice,
                               st) != hipSuccess ||
                hipMemcpyAsync(c0.batch_map.p, map.data(), (size_t)G * sizeof(int2), hipMemcpyHostToDevice, st) !=
                    hipSuccess) {
                rc = MDY_ERR_HIP;
                break;
            }
            if (by_done && hipMemsetAsync(c0.batch_done.p, 0, sizeof(int), st) != hipSuccess) {
                rc = MDY_ERR_HIP;
                break;
            }
            const CgArgs *pa = reinterpret_cast<const CgArgs *>(c0.batch_args.p);
            const int2 *pm = c0.batch_map.p;
            int *done = by_done && m > 1 ? c0.batch_done.p : nullptr;
            int suspend_done = done ? std::max(1, m / by_done) : 0;
            void *kargs[] = {&pa, &pm, &done, &suspend_done};
            (void)hipEventRecord(e0, st);
            if (hipLaunchCooperativeKernel(kern, dim3(G), dim3(tpb), kargs, 0, st) != hipSuccess) {
                rc = MDY_ERR_LAUNCH;
                break;
            }
            (void)hipEventRecord(e1, st);
            // the host arrays are reused by the next launch: wait for this one
            if (hipStreamSynchronize(st) != hipSuccess) { rc = MDY_ERR_HIP; break; }
            float ms = 0;
            (void)hipEventElapsedTime(&ms, e0, e1);
            kernel_ms += ms;
            for (int i = 0; i < m && rc == MDY_OK; i++) {
                const int k = todo[lo + i];
                Ctx &ci = *cs[k];
                double R[R_SIZE];
                if (hipMemcpy(R, ci.cg_out.p, sizeof(R), hipMemcpyDeviceToHost) != hipSuccess) {
                    rc = MDY_ERR_HIP;
                    break;
                }
                if (t_origin < 0 || R[R_T0] < t_origin) t_origin = R[R_T0];
                ci.coop_blocks = gsz[i];
                if (R[R_SUSP] != 0.0 && R[R_STATUS] == MDY_OK) {   // resumes in the next pass
                    resume[k] = 1;
                    next.push_back(k);
                    continue;
                }
                collect(ci, R, &p[k], args[i], &res[k]);
                if (net_ms) net_ms[k] = (R[R_T1] - t_origin) * 1e-5;   // 100 MHz ticks
                rc = copy_force_out(ci, R, d_force_out ? d_force_out[k] : nullptr, st);
            }
        }
        todo.swap(next);
        if (cap > 0) cap *= kCapGrowth;
    }
    (void)hipStreamSynchronize(st);
    if (total_ms) *total_ms = kernel_ms;
    return rc;
}

}  // namespace mdy
