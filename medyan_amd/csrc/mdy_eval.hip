// mdy_eval.hip -- one-shot energy / force evaluation kernels behind
// mdy_energy(_dev) and mdy_forces(_dev) (ForceFieldManager::computeEnergy and
// ::computeForces), including the excluded-volume quadrature fallbacks.
#include <mutex>

#include "mdy_eval.hpp"
#include "mdy_launch.hpp"

namespace mdy {

// Three waves per SIMD (168 registers, ~5 words per lane spilled in one
// instantiation): 7-8 % faster than two at 40k and 200k cylinders.
// energies: per-FF double sums at [0, kNFF), exact total at kXsE
template <class P, bool QUAD = false>
__global__ void __launch_bounds__(kBlock, 3) k_energy(Dev D, P pos, double *red) {
    __shared__ double lds[17 * kNFF + 8];
    double acc[kNFF];
#pragma unroll
    for (int k = 0; k < kNFF; k++) acc[k] = 0.0;
    mdy_xsum xs;
    mdy_xs_zero(&xs);
    const int total = D.n_items(), gs = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gs)
        item<false, QUAD>(D, i, pos, acc, xs, nullptr, true, false);
    block_reduce<kNFF>(acc, lds);
    block_reduce_xsum(xs, reinterpret_cast<long long *>(lds));
    if (threadIdx.x < kNFF) red[(size_t)blockIdx.x * kRed + threadIdx.x] = acc[threadIdx.x];
    if (threadIdx.x == 0) store_xsum(red, blockIdx.x, kXsE, xs);
}

// One block per queued pair: 2-D Simpson energy (energyN, 400 x 400 intervals).
template <class P>
__global__ void __launch_bounds__(kBlock) k_energy_simpson(Dev D, P pos, double *qe_u) {
    __shared__ long long lds[17 * 8];
    const int nq = min(*D.qe_n, kQueueCap);
    const int N = 400;
    for (int q = blockIdx.x; q < nq; q += gridDim.x) {
        const int i = D.qe[q];
        int4 b = D.ev_bs[i];
        SimpsonGeo g = simpson_geo(pos(b.x), pos(b.y), pos(b.z), pos(b.w));
        mdy_xsum X;
        mdy_xs_zero(&X);
        if (threadIdx.x == 0) mdy_xs_add(&X, simpson_e_corners(g));
        for (int node = threadIdx.x; node < (N + 1) * (N + 1); node += blockDim.x)
            simpson_e_node(g, node / (N + 1), node % (N + 1), N, X);
        block_reduce_xsum(X, lds);
        if (threadIdx.x == 0) {
            const double ds = 1.0 / ((double)N);
            qe_u[q] = D.ev_kr[i] * (ds * ds / 9.0) * mdy_xs_value(&X);
        }
        __syncthreads();
    }
}

// out: the kNFF energies, the total (+inf for an invalid pair), and when
// n_out = kNFF + 2 the total without the sentinel
__global__ void __launch_bounds__(kBlock) k_energy_final(Dev D, const double *red, int nblk, const double *qe_u,
                                                         double *out, int n_out) {
    __shared__ double lds[17 * kNFF + 8];
    double e[kNFF + 2];
    energy_epilogue(D, red, nblk, *D.qe_n, D.qe, qe_u, lds, e, true);
    if (threadIdx.x == 0)
        for (int k = 0; k < n_out; k++) out[k] = e[k];
}

template <bool MIXED, class P, bool QUAD = false>
__global__ void __launch_bounds__(kBlock) k_force(Dev D, P pos, double *contrib_) {
    auto *contrib = reinterpret_cast<typename SlotT<MIXED>::T *>(contrib_);
    const int total = D.n_items(), gs = gridDim.x * blockDim.x;
    double acc[kNFF];
    mdy_xsum xs;
    mdy_xs_zero(&xs);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gs)
        item<MIXED, QUAD>(D, i, pos, acc, xs, contrib, false, true);
}

// Quadrature force (forceN, 10^4 x 10^4 intervals) of one queued pair,
// integral component `comp`, split over the whole grid; exact partials per
// block, one record array per component.
__global__ void __launch_bounds__(kBlock) k_force_simpson_part(Dev D, const double *x, int qi, int moved, int comp,
                                                               double *part) {
    __shared__ long long lds[17 * 8];
    const int N = 10000;
    const int i = D.qf[qi];
    int4 b = D.ev_bs[i];
    PosX pos{x};
    d3 c1 = pos(b.x), c2 = pos(b.y), c3 = pos(b.z), c4 = pos(b.w);
    if (moved) {  // both beads of the first cylinder moved off the plane (:1017-1022)
        d3 n2 = moved_off_plane(c1, c2, c3, c2, 0.01), n1 = moved_off_plane(c1, c2, c3, c1, 0.01);
        c1 = n1; c2 = n2;
    }
    SimpsonGeo g = simpson_geo(c1, c2, c3, c4);
    mdy_xsum X;
    mdy_xs_zero(&X);
    if (blockIdx.x == 0 && threadIdx.x == 0) simpson_f_corners(g, comp, X);
    const long long nodes = (long long)(N + 1) * (N + 1);
    for (long long nd = (long long)blockIdx.x * blockDim.x + threadIdx.x; nd < nodes;
         nd += (long long)gridDim.x * blockDim.x)
        simpson_f_node(g, (int)(nd / (N + 1)), (int)(nd % (N + 1)), N, comp, X);
    block_reduce_xsum(X, lds);
    if (threadIdx.x == 0) store_xsum(part + (size_t)comp * gridDim.x * kRed, blockIdx.x, 0, X);
}

// flag[0] = 1 if the result was non-finite, flag[1] = pair in plane
// (the host retries with moved beads or fails like the reference)
template <bool MIXED>
__global__ void __launch_bounds__(kBlock) k_force_simpson_final(Dev D, const double *x, int qi, int moved,
                                                                const double *part, int nblk, double *contrib_,
                                                                int *flag) {
    auto *contrib = reinterpret_cast<typename SlotT<MIXED>::T *>(contrib_);
    __shared__ long long lds[17 * 8];
    const int N = 10000;
    double T[6];
    for (int q = 0; q < 6; q++) {
        mdy_xsum X;
        grid_reduce_xsum(part + (size_t)q * nblk * kRed, nblk, 0, X, lds);
        T[q] = mdy_xs_value(&X);
    }
    if (threadIdx.x != 0) return;
    const int i = D.qf[qi];
    int4 b = D.ev_bs[i];
    PosX pos{x};
    d3 c1 = pos(b.x), c2 = pos(b.y), c3 = pos(b.z), c4 = pos(b.w);
    flag[1] = are_in_plane(c1, c2, c3, c4) ? 1 : 0;
    if (moved) {
        d3 n2 = moved_off_plane(c1, c2, c3, c2, 0.01), n1 = moved_off_plane(c1, c2, c3, c1, 0.01);
        c1 = n1; c2 = n2;
    }
    SimpsonGeo g = simpson_geo(c1, c2, c3, c4);
    d3 f[4];
    simpson_f_assemble(g, T, D.ev_kr[i], N, f);
    bool ok = finite_d3(f[0]) && finite_d3(f[1]) && finite_d3(f[2]) && finite_d3(f[3]);
    flag[0] = ok ? 0 : 1;
    if (ok) {
        int s = D.slot_ev + 4 * i;
        put3(contrib, s, f[0]); put3(contrib, s + 1, f[1]); put3(contrib, s + 2, f[2]); put3(contrib, s + 3, f[3]);
    }
}

// WIDE: the wide per-bead form (fewer beads than one block per CU has
// threads: at 200k cylinders the narrow form's occupancy wins); same sums
// either way
template <bool MIXED, bool WIDE>
__global__ void __launch_bounds__(kBlock) k_gather(Dev D, const double *contrib_, double *force, int accumulate) {
    const auto *contrib = reinterpret_cast<const typename SlotT<MIXED>::T *>(contrib_);
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < D.nb; b += gridDim.x * blockDim.x) {
        d3 f = WIDE ? gather_bead_w<12, 24>(D, contrib, b) : gather_bead(D, contrib, b);
        if (accumulate) {
            force[3 * b] += f.x; force[3 * b + 1] += f.y; force[3 * b + 2] += f.z;
        } else {
            force[3 * b] = f.x; force[3 * b + 1] = f.y; force[3 * b + 2] = f.z;
        }
    }
}

__global__ void k_zero_int(int *p, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

// ---------------------------------------------------------------------------
// host launchers
// ---------------------------------------------------------------------------
static int grid_for(const Ctx &c, long long n) {
    long long g = (n + kBlock - 1) / kBlock;
    long long cap = (long long)c.num_cu * 8;
    if (g > cap) g = cap;
    if (g < 1) g = 1;
    return (int)g;
}

int launch_energy(Ctx &c, const double *d_x, double *d_out, hipStream_t st) {
    Dev D = c.dev();
    const int nblk = grid_for(c, D.n_items());
    const bool quad = quad_on(c);
    const int nqb = quad ? quad_blocks(c) : 0;   // the quadrature pairs' block records follow the others
    MDY_TRY(ensure(c.red, (size_t)kRed * (nblk + nqb)));
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, c.qn.p, 4);
    if (quad) {
        MDY_TRY(quad_classify(c, D, d_x, st));
        hipLaunchKernelGGL((k_energy<PosX, true>), dim3(nblk), dim3(kBlock), 0, st, D, PosX{d_x}, c.red.p);
        MDY_TRY(launch_quad(c, D, 0, d_x, nullptr, c.red.p, nblk, st));
    } else {
        hipLaunchKernelGGL(k_energy<PosX>, dim3(nblk), dim3(kBlock), 0, st, D, PosX{d_x}, c.red.p);
    }
    hipLaunchKernelGGL(k_energy_simpson<PosX>, dim3(64), dim3(kBlock), 0, st, D, PosX{d_x}, c.qe_u.p);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kBlock), 0, st, D, c.red.p, nblk + nqb, c.qe_u.p, d_out,
                       kNFF + 2);
    return hipGetLastError() == hipSuccess ? MDY_OK : MDY_ERR_HIP;
}

// ---------------------------------------------------------------------------
// The allocation-free energy evaluation (mdy_energy_dev_noalloc): the same
// kernels as launch_energy, with every piece of scratch memory (block
// partials, fallback queue and its counters, the per-pair quadrature
// energies, the error word) in one file-scope device array. Each device holds
// its own copy; host code reaches it only through the address that
// hipGetSymbolAddress returns for that device.
// ---------------------------------------------------------------------------
constexpr int kScratchBlocks = 2048;   // block records: 8 blocks per CU on a 256-CU device
constexpr size_t kScRed = 0;                                                   // kRed doubles per block
constexpr size_t kScQn = kScRed + sizeof(double) * kRed * kScratchBlocks;     // 16 ints: queue counters
constexpr size_t kScErr = kScQn + 256;                                         // 4 ints: error word
constexpr size_t kScQe = kScErr + 256;                                         // energy fallback queue
constexpr size_t kScQf = kScQe + sizeof(int) * kQueueCap;                      // force queue (unused here)
constexpr size_t kScQeU = kScQf + sizeof(int) * kQueueCap;                     // quadrature energies
constexpr size_t kScratchBytes = kScQeU + sizeof(double) * kQueueCap;
static_assert(kScQn % 256 == 0 && kScQe % 256 == 0 && kScQeU % 256 == 0, "scratch alignment");

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char *g_scratch_ptr[kMaxDevices];

// Called by mdy_create with `device` current: resolves this device's copy once.
int resolve_noalloc_scratch(int device) {
    if (device < 0 || device >= kMaxDevices) return MDY_ERR_ARG;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return MDY_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MDY_ERR_HIP;
    g_scratch_ptr[device] = static_cast<unsigned char *>(p);
    return MDY_OK;
}

size_t noalloc_scratch_bytes() { return kScratchBytes; }

int launch_energy_noalloc(Ctx &c, const double *d_x, double *d_out, hipStream_t st) {
    unsigned char *sc = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (c.device >= 0 && c.device < kMaxDevices) sc = g_scratch_ptr[c.device];
    }
    if (!sc) return MDY_ERR_STATE;
    Dev D = c.dev();
    const int nblk = grid_for(c, D.n_items());
    if (nblk > kScratchBlocks) return MDY_ERR_CAPACITY;
    double *red = reinterpret_cast<double *>(sc + kScRed);
    int *qn = reinterpret_cast<int *>(sc + kScQn);
    D.qe = reinterpret_cast<int *>(sc + kScQe);
    D.qe_n = qn;
    D.qf = reinterpret_cast<int *>(sc + kScQf);
    D.qf_n = qn + 1;
    D.err = reinterpret_cast<int *>(sc + kScErr);
    double *qe_u = reinterpret_cast<double *>(sc + kScQeU);
    // the counters and the error word keep the previous call's values: cleared by a kernel
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, qn, 4);
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, D.err, 4);
    hipLaunchKernelGGL(k_energy<PosX>, dim3(nblk), dim3(kBlock), 0, st, D, PosX{d_x}, red);
    hipLaunchKernelGGL(k_energy_simpson<PosX>, dim3(64), dim3(kBlock), 0, st, D, PosX{d_x}, qe_u);
    hipLaunchKernelGGL(k_energy_final, dim3(1), dim3(kBlock), 0, st, D, red, nblk, qe_u, d_out, kNFF + 1);
    return hipGetLastError() == hipSuccess ? MDY_OK : MDY_ERR_HIP;
}

int launch_forces(Ctx &c, const double *d_x, double *d_force, int accumulate, hipStream_t st) {
    Dev D = c.dev();
    const int nblk = grid_for(c, D.n_items());
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(64), 0, st, c.qn.p, 4);
    const bool mixed = c.opts.precision == MDY_PREC_MIXED;
    if (quad_on(c)) {   // FP64 only (the mode is refused for mixed-precision contexts)
        MDY_TRY(quad_classify(c, D, d_x, st));
        hipLaunchKernelGGL((k_force<false, PosX, true>), dim3(nblk), dim3(kBlock), 0, st, D, PosX{d_x},
                           c.contrib.p);
        MDY_TRY(launch_quad(c, D, 1, d_x, nullptr, c.contrib.p, 0, st));
    } else if (mixed) {
        hipLaunchKernelGGL((k_force<true, PosX>), dim3(nblk), dim3(kBlock), 0, st, D, PosX{d_x}, c.contrib.p);
    } else {
        hipLaunchKernelGGL((k_force<false, PosX>), dim3(nblk), dim3(kBlock), 0, st, D, PosX{d_x}, c.contrib.p);
    }
    if (hipGetLastError() != hipSuccess) return MDY_ERR_HIP;
    // quadrature fallbacks: rare; the count is read back to drive them
    int h_q[4];
    if (hipMemcpyAsync(h_q, c.qn.p, sizeof(h_q), hipMemcpyDeviceToHost, st) != hipSuccess) return MDY_ERR_HIP;
    if (hipStreamSynchronize(st) != hipSuccess) return MDY_ERR_HIP;
    if (h_q[1] > kQueueCap) return MDY_ERR_CAPACITY;
    const int nq = h_q[1];
    const int sblk = c.num_cu * 4;
    MDY_TRY(ensure(c.simpson_part, (size_t)6 * kRed * sblk));
    for (int q = 0; q < nq; q++) {
        for (int moved = 0; moved < 2; moved++) {
            for (int comp = 0; comp < 6; comp++)
                hipLaunchKernelGGL(k_force_simpson_part, dim3(sblk), dim3(kBlock), 0, st, D, d_x, q, moved, comp,
                                   c.simpson_part.p);
            if (mixed)
                hipLaunchKernelGGL(k_force_simpson_final<true>, dim3(1), dim3(kBlock), 0, st, D, d_x, q, moved,
                                   c.simpson_part.p, sblk, c.contrib.p, c.qn.p + 2);
            else
                hipLaunchKernelGGL(k_force_simpson_final<false>, dim3(1), dim3(kBlock), 0, st, D, d_x, q, moved,
                                   c.simpson_part.p, sblk, c.contrib.p, c.qn.p + 2);
            int fl[2];
            if (hipMemcpyAsync(fl, c.qn.p + 2, sizeof(fl), hipMemcpyDeviceToHost, st) != hipSuccess) return MDY_ERR_HIP;
            if (hipStreamSynchronize(st) != hipSuccess) return MDY_ERR_HIP;
            if (!fl[0]) break;
            if (moved || !fl[1]) return MDY_ERR_FORCE;  // the reference exits here
        }
    }
    const int gg = grid_for(c, D.nb);
    const bool wide = (long long)D.nb <= (long long)c.num_cu * kBlock;
    auto kern = mixed ? (wide ? k_gather<true, true> : k_gather<true, false>)
                      : (wide ? k_gather<false, true> : k_gather<false, false>);
    hipLaunchKernelGGL(kern, dim3(gg), dim3(kBlock), 0, st, D, c.contrib.p, d_force, accumulate);
    return hipGetLastError() == hipSuccess ? MDY_OK : MDY_ERR_HIP;
}

}  // namespace mdy
