// mdy_aux.hip -- post-minimization auxiliary outputs
// (MController::afterMinimization_, Controller/MController.h:240-268):
//   * ForceFieldManager::assignallforcemags: the stretch force of every linker
//     and motor (f0 / invL, LinkerStretchingHarmonic.cpp:293,
//     MotorGhostStretchingHarmonic.cpp:339) and the 3-vector of every branch
//     point (BranchingStretchingHarmonic.cpp:273-275), at given coordinates;
//   * ForceFieldManager::computeLoadForces (ForceFieldManager.cpp:292-309) with
//     BoundaryCylinderRepulsion::computeLoadForces
//     (BoundaryCylinderRepulsion.cpp:124-189): the boundary load on every
//     monomer position ahead of each filament end in a boundary plane's
//     neighbor list.
#include <algorithm>

#include "mdy_eval.hpp"
#include "mdy_launch.hpp"

namespace mdy {

__global__ void __launch_bounds__(kBlock) k_stretch_forces(Dev D, const double *__restrict__ x, double *linker,
                                                            double *motor, double *branch3) {
    const int nbr = max(D.n_bs, max(D.n_bdh, D.n_bp));
    const int n = D.n_lm + nbr;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (i < D.n_lm) {
            const int4 b = D.lm_bs[i];
            const d3 v1 = midpt(ld3(x + b.x), ld3(x + b.y), D.lm_p1[i]);
            const d3 v2 = midpt(ld3(x + b.z), ld3(x + b.w), D.lm_p2[i]);
            const double dist = tpdist(v1, v2);
            const double invL = 1 / dist;
            const double f0 = D.lm_k[i] * (dist - D.lm_l0[i]) * invL;
            const double s = f0 / invL;
            if (i < D.n_ls) { if (linker) linker[i] = s; }
            else if (motor) motor[i - D.n_ls] = s;
        } else if (branch3) {
            // branch force: the stretching, dihedral and position contributions
            // of branch point j added in force-field order, starting from 0
            const int j = i - D.n_lm;
            d3 acc{0.0, 0.0, 0.0};
            if (j < D.n_bs) {
                const int *b = D.bs_bs + 3 * j;
                const d3 v1 = midpt(ld3(x + b[0]), ld3(x + b[1]), D.bs_p[j]);
                const d3 c3 = ld3(x + b[2]);
                const double dist = tpdist(v1, c3);
                const double invL = 1 / dist;
                const double f0 = D.bs_k[j] * (dist - D.bs_l0[j]) * invL;
                const d3 r = sub(c3, v1);
                acc = d3{acc.x + -f0 * (r.x), acc.y + -f0 * (r.y), acc.z + -f0 * (r.z)};
            }
            if (j < D.n_bdh) {
                const int4 b = D.bdh_bs[j];
                d3 f1, f2, f3, f4;
                if (D.bdh_quad)
                    bdhq_force(ld3(x + b.x), ld3(x + b.y), ld3(x + b.z), ld3(x + b.w), D.bdh_k[j], D.bdh_pos[j], f1, f2,
                               f3, f4);
                else
                    bdh_force(ld3(x + b.x), ld3(x + b.y), ld3(x + b.z), ld3(x + b.w), D.bdh_k[j], D.bdh_pos[j], f1, f2,
                              f3, f4);
                acc = d3{acc.x + f3.x, acc.y + f3.y, acc.z + f3.z};
            }
            if (j < D.n_bp) {
                const int *b = D.bp_bs + 3 * j;
                d3 f1, f2, f3;
                bp_force(ld3(x + b[0]), ld3(x + b[1]), ld3(x + b[2]), D.bp_k[j], D.bp_pos[j], D.cq, D.sq, f1, f2, f3);
                acc = d3{acc.x + f3.x, acc.y + f3.y, acc.z + f3.z};
            }
            branch3[3 * j] = acc.x;
            branch3[3 * j + 1] = acc.y;
            branch3[3 * j + 2] = acc.z;
        }
    }
}

// One thread per (cylinder, monomer position). A plus-end cylinder loads its
// second bead (direction first -> second bead), a minus-end cylinder its
// first (direction second -> first); the planes are visited in order and
// each adds proj * K exp(-r/lambda) / lambda, the reference's sum order.
__global__ void __launch_bounds__(kBlock) k_load_forces(int ncyl, int np, const int *__restrict__ b1,
                                                         const int *__restrict__ b2, const short *__restrict__ type,
                                                         const unsigned char *__restrict__ minus,
                                                         const unsigned char *__restrict__ plus,
                                                         const int *__restrict__ flag, const double *__restrict__ planes,
                                                         const double *__restrict__ pk, const double *__restrict__ psl,
                                                         const double *__restrict__ x, int ntypes,
                                                         const double *__restrict__ mon_size,
                                                         const int *__restrict__ n_mon, int max_mon, double *lp,
                                                         double *lm) {
    const long long tot = 2LL * ncyl * max_mon;
    for (long long t = blockIdx.x * (long long)blockDim.x + threadIdx.x; t < tot;
         t += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(t % max_mon);
        const int c = (int)((t / max_mon) % ncyl);
        const int end = (int)(t / ((long long)max_mon * ncyl));   // 0 plus, 1 minus
        if (end == 0 ? !plus[c] : !minus[c]) continue;
        const int ty = type[c] < ntypes ? type[c] : 0;
        if (i >= n_mon[ty]) continue;
        const int ob = end == 0 ? b1[c] : b2[c], db = end == 0 ? b2[c] : b1[c];
        const d3 bo = ld3(x + ob), bd = ld3(x + db);
        // normalizedVector(bd - bo): magnitude from the squared components in order
        const d3 v = sub(bd, bo);
        double mag2 = 0.0;
        mag2 += v.x * v.x;
        mag2 += v.y * v.y;
        mag2 += v.z * v.z;
        const double norm = sqrt(mag2);
        const d3 dir{v.x / norm, v.y / norm, v.z / norm};
        const double step = (double)i * mon_size[ty];
        const d3 nc{bd.x + step * dir.x, bd.y + step * dir.y, bd.z + step * dir.z};
        double acc = 0.0;
        for (int p = 0; p < np; p++) {
            if (!flag[(size_t)p * ncyl + c]) continue;
            const double *pl = planes + 4 * p;
            double dt = 0.0;
            dt += pl[0] * dir.x;
            dt += pl[1] * dir.y;
            dt += pl[2] * dir.z;
            const double proj = (-dt < 0.0) ? 0.0 : -dt;
            const double r = plane_dist(pl, nc);
            const double R = -r / psl[p];
            const double load = pk[p] * xexp(R) / psl[p];
            acc += proj * load;
        }
        const size_t o = (size_t)(db / 3) * max_mon + i;
        if (end == 0) lp[o] = acc;
        else lm[o] = acc;
    }
}

static int grid_n(const Ctx &c, long long n) {
    long long g = (n + kBlock - 1) / kBlock, cap = (long long)c.num_cu * 8;
    return (int)(g > cap ? cap : g < 1 ? 1 : g);
}

int launch_stretch_forces(Ctx &c, const double *d_x, double *d_linker, double *d_motor, double *d_branch,
                          hipStream_t st) {
    Dev D = c.dev();
    const long long nbr = std::max(D.n_bs, std::max(D.n_bdh, D.n_bp));
    hipLaunchKernelGGL(k_stretch_forces, dim3(grid_n(c, (long long)D.n_lm + nbr)), dim3(kBlock), 0, st, D, d_x,
                       d_linker, d_motor, d_branch);
    return hipGetLastError() == hipSuccess ? MDY_OK : MDY_ERR_HIP;
}

int launch_load_forces(Ctx &c, const double *d_x, int ntypes, const double *d_mon_size, const int *d_n_mon,
                       int max_mon, double *d_lp, double *d_lm, hipStream_t st) {
    const size_t n = (size_t)c.nb * max_mon;
    MDY_HIP(hipMemsetAsync(d_lp, 0, n * sizeof(double), st));
    MDY_HIP(hipMemsetAsync(d_lm, 0, n * sizeof(double), st));
    if (!c.bd_on || c.n_planes == 0 || c.ncyl == 0) return MDY_OK;
    hipLaunchKernelGGL(k_load_forces, dim3(grid_n(c, 2LL * c.ncyl * max_mon)), dim3(kBlock), 0, st, c.ncyl,
                       c.n_planes, c.c_b1.p, c.c_b2.p, c.c_type.p, c.c_minus.p, c.c_plus.p, c.bd_flag.p, c.planes.p,
                       c.plane_k.p, c.plane_sl.p, d_x, ntypes, d_mon_size, d_n_mon, max_mon, d_lp, d_lm);
    return hipGetLastError() == hipSuccess ? MDY_OK : MDY_ERR_HIP;
}

}  // namespace mdy
