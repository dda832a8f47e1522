// mdy_hvp.hip -- derivatives of the exact-gradient force field (no coplanar
// perturbation, no Simpson fallback; in the quadrature mode the flagged
// excluded-volume pairs come from mdy_quad.hip) for implicit differentiation
// of minimized structures:
//   * Hessian-vector products H v, by forward-mode differentiation of the
//     force code: bead positions are the dual numbers x + t v, every
//     interaction's forces are evaluated in dual arithmetic and the tangent
//     parts d/dt F(x + t v) = -H v are written (negated) to the interaction's
//     force slots, then gathered per bead in the reference order;
//   * per-force-field directional derivatives grad E_k(x) . w of the
//     energies, from the interaction forces dotted with w at their beads.
#include "mdy_eval.hpp"
#include "mdy_launch.hpp"

namespace mdy {

// Forces of one interaction on its beads in scalar type T (f[0..nbeads));
// returns the force-field index, or -1 for an empty record.
template <class T>
__device__ __forceinline__ int item_forces(const Dev &D, const ItemRec &r, const V3<T> c[4], V3<T> f[4]) {
    if (r.type == IT_EV) {
        if (D.evq && D.evq[r.j]) return -1;   // quadrature mode: evaluated by the quadrature kernels
        EvInvT<T> inv;
        ev_invariants(c[0], c[1], c[2], c[3], inv);
        ev_force_inv_t<T>(inv, c[0], c[1], c[2], c[3], r.p0, f);
        return MDY_FF_EV;
    } else if (r.type == IT_EVM) {
        const int2 cc = D.ev_cyl[r.j];
        (void)evm_eval<T, true>(c[0], c[1], c[2], c[3], D.cyl_mon[cc.x], D.cyl_mon[cc.y], r.p0, f);
        return MDY_FF_EV;
    } else if (r.type == IT_FS) {
        fs_force(c[0], c[1], r.p0, r.p1, f[0], f[1]);
        return MDY_FF_FS;
    } else if (r.type == IT_FB) {
        fb_force(c[0], c[1], c[2], r.p0, r.p1, r.p2, r.p3, f[0], f[1], f[2]);
        return MDY_FF_FB;
    } else if (r.type == IT_LM) {
        ls_force(c[0], c[1], c[2], c[3], r.p0, r.p1, r.p2, r.p3, f[0], f[1], f[2], f[3]);
        return r.j < D.n_ls ? MDY_FF_LS : MDY_FF_MS;
    } else if (r.type == IT_BS) {
        bs_force(c[0], c[1], c[2], r.p0, r.p1, r.p2, f[0], f[1], f[2]);
        return MDY_FF_BS;
    } else if (r.type == IT_BB) {
        bb_force(c[0], c[1], c[2], c[3], r.p0, r.p1, r.p2, r.p3, f[0], f[1], f[2], f[3]);
        return MDY_FF_BB;
    } else if (r.type == IT_BDH) {
        if (D.bdh_quad) bdhq_force(c[0], c[1], c[2], c[3], r.p0, r.p1, f[0], f[1], f[2], f[3]);
        else bdh_force(c[0], c[1], c[2], c[3], r.p0, r.p1, f[0], f[1], f[2], f[3]);
        return MDY_FF_BDH;
    } else if (r.type == IT_BP) {
        bp_force(c[0], c[1], c[2], r.p0, r.p1, D.cq, D.sq, f[0], f[1], f[2]);
        return MDY_FF_BP;
    } else if (r.type == IT_BD) {
        f[0] = bd_force(D.planes + 4 * r.b.y, c[0], r.p0, r.p1);
        return MDY_FF_BD;
    }
    return -1;
}

__device__ __forceinline__ int item_slot(const Dev &D, const ItemRec &r) {
    switch (r.type) {
        case IT_EV:
        case IT_EVM: return D.slot_ev + 4 * r.j;
        case IT_FS: return D.slot_fs + 2 * r.j;
        case IT_FB: return D.slot_fb + 3 * r.j;
        case IT_LM: return D.slot_lm + 4 * r.j;
        case IT_BS: return D.slot_bs + 3 * r.j;
        case IT_BB: return D.slot_bb + 4 * r.j;
        case IT_BDH: return D.slot_bdh + 4 * r.j;
        case IT_BP: return D.slot_bp + 3 * r.j;
        default: return D.slot_bd + r.j;
    }
}

__device__ __forceinline__ int bead_off(const ItemRec &r, int k) {
    return k == 0 ? r.b.x : k == 1 ? r.b.y : k == 2 ? r.b.z : r.b.w;
}

// Two waves per SIMD (the compiler fits 256 registers with ~15 words per
// lane spilled; at one wave the kernel needed 286): 13 % faster at 50k.
__global__ void __launch_bounds__(kBlock, 2) k_hvp(Dev D, const double *__restrict__ x, const double *__restrict__ v,
                                                 double *__restrict__ contrib) {
    const int total = D.n_items(), gs = gridDim.x * blockDim.x;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gs) {
        const ItemRec r = fetch_rec(D, i);
        const int nb = rec_nbeads(r.type);
        V3<Dual> c[4], f[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int o = bead_off(r, k < nb ? k : 0);
            c[k] = V3<Dual>{Dual(x[o], v[o]), Dual(x[o + 1], v[o + 1]), Dual(x[o + 2], v[o + 2])};
        }
        if (item_forces<Dual>(D, r, c, f) < 0) continue;
        const int s = item_slot(D, r);
        bool ok = true;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (k < nb) {
                const d3 h{-f[k].x.d, -f[k].y.d, -f[k].z.d};   // -dF/dt = H v
                ok = ok && finite_d3(h);
                put3(contrib, s + k, h);
            }
        }
        // a degenerate excluded-volume pair has no analytic second derivative
        if (!ok) set_err(D.err, MDY_ERR_FORCE);
    }
}

// WIDE: the wide per-bead form (fewer beads than one block per CU has
// threads); same sums either way
template <bool WIDE>
__global__ void __launch_bounds__(kBlock) k_hvp_gather(Dev D, const double *__restrict__ contrib,
                                                        double *__restrict__ hv) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < D.nb; b += gridDim.x * blockDim.x) {
        const d3 h = WIDE ? gather_bead_w<12, 24>(D, contrib, b) : gather_bead(D, contrib, b);
        hv[3 * b] = h.x; hv[3 * b + 1] = h.y; hv[3 * b + 2] = h.z;
    }
}

// per-block partial sums of grad E_k . w = -sum over an interaction's beads of F . w
__global__ void __launch_bounds__(kBlock) k_energy_jvp(Dev D, const double *__restrict__ x,
                                                        const double *__restrict__ w, double *red) {
    __shared__ double lds[17 * kNFF + 8];
    double acc[kNFF];
#pragma unroll
    for (int k = 0; k < kNFF; k++) acc[k] = 0.0;
    const int total = D.n_items(), gs = gridDim.x * blockDim.x;
    bool ok = true;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < total; i += gs) {
        const ItemRec r = fetch_rec(D, i);
        const int nb = rec_nbeads(r.type);
        d3 c[4], f[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            const int o = bead_off(r, k < nb ? k : 0);
            c[k] = d3{x[o], x[o + 1], x[o + 2]};
        }
        const int ff = item_forces<double>(D, r, c, f);
        if (ff < 0) continue;
        double s = 0.0;
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (k < nb) {
                const int o = bead_off(r, k);
                ok = ok && finite_d3(f[k]);
                s -= f[k].x * w[o] + f[k].y * w[o + 1] + f[k].z * w[o + 2];
            }
        }
#pragma unroll
        for (int k = 0; k < kNFF; k++) acc[k] += (k == ff) ? s : 0.0;
    }
    if (!ok) set_err(D.err, MDY_ERR_FORCE);
    block_reduce<kNFF>(acc, lds);
    if (threadIdx.x < kNFF) red[(size_t)blockIdx.x * kRed + threadIdx.x] = acc[threadIdx.x];
}

__global__ void __launch_bounds__(kBlock) k_energy_jvp_final(const double *red, int nblk, double *out) {
    __shared__ double lds[17 * kNFF + 8];
    double v[kNFF];
    grid_partials_reduce<kNFF>(red, nblk, v, lds);
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int k = 0; k < kNFF; k++) {
            out[k] = v[k];
            t += v[k];
        }
        out[kNFF] = t;
    }
}

__global__ void k_clear_err(int *err) { *err = 0; }

static int grid_of(const Ctx &c, long long n) {
    long long g = (n + kBlock - 1) / kBlock, cap = (long long)c.num_cu * 8;
    return (int)(g > cap ? cap : g < 1 ? 1 : g);
}

// reads and clears the error word after the stream's work
int sync_errors(Ctx &c, hipStream_t st) {
    if (hipGetLastError() != hipSuccess) return MDY_ERR_HIP;
    int e = 0;
    if (hipMemcpyAsync(&e, c.err.p, sizeof(int), hipMemcpyDeviceToHost, st) != hipSuccess) return MDY_ERR_HIP;
    if (hipStreamSynchronize(st) != hipSuccess) return MDY_ERR_HIP;
    if (e) {
        hipLaunchKernelGGL(k_clear_err, dim3(1), dim3(1), 0, st, c.err.p);
        if (hipStreamSynchronize(st) != hipSuccess) return MDY_ERR_HIP;
        return e;
    }
    return MDY_OK;
}

int launch_hvp(Ctx &c, const double *d_x, const double *d_v, double *d_hv, hipStream_t st, bool sync) {
    Dev D = c.dev();
    const bool quad = quad_on(c);
    if (quad) MDY_TRY(quad_classify(c, D, d_x, st));
    hipLaunchKernelGGL(k_hvp, dim3(grid_of(c, D.n_items())), dim3(kBlock), 0, st, D, d_x, d_v, c.contrib.p);
    if (quad) MDY_TRY(launch_quad(c, D, 2, d_x, d_v, c.contrib.p, 0, st));
    hipLaunchKernelGGL(D.nb <= c.num_cu * kBlock ? k_hvp_gather<true> : k_hvp_gather<false>, dim3(grid_of(c, D.nb)),
                       dim3(kBlock), 0, st, D, c.contrib.p, d_hv);
    if (!sync) return hipGetLastError() == hipSuccess ? MDY_OK : MDY_ERR_HIP;
    return sync_errors(c, st);
}

int launch_energy_jvp(Ctx &c, const double *d_x, const double *d_w, double *d_out, hipStream_t st) {
    Dev D = c.dev();
    const int nblk = grid_of(c, D.n_items());
    const bool quad = quad_on(c);
    const int nqb = quad ? quad_blocks(c) : 0;   // the quadrature pairs' block records follow the others
    MDY_TRY(ensure(c.red, (size_t)kRed * (nblk + nqb)));
    if (quad) MDY_TRY(quad_classify(c, D, d_x, st));
    hipLaunchKernelGGL(k_energy_jvp, dim3(nblk), dim3(kBlock), 0, st, D, d_x, d_w, c.red.p);
    if (quad) MDY_TRY(launch_quad(c, D, 3, d_x, d_w, c.red.p, nblk, st));
    hipLaunchKernelGGL(k_energy_jvp_final, dim3(1), dim3(kBlock), 0, st, c.red.p, nblk + nqb, d_out);
    return sync_errors(c, st);
}

}  // namespace mdy
