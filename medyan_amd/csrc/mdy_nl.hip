// mdy_nl.hip -- binned cylinder neighbor-list rebuild and the vectorization of
// the excluded-volume and boundary force fields, all on the device.
//
// Pair predicate: HybridCylinderCylinderNL::updateNeighborsbin
// (src/Structure/HybridNeighborListImpl.cpp:405-441): keep (c, n) iff
// id(c) > id(n), not (same filament and |pos(c) - pos(n)| <= 2), and the
// squared midpoint distance (double) lies in [(float)rMin^2, (float)rMax^2].
// The reference's bins (edge 1.125 rMax) only prune candidates; the cells used
// here have edge >= the float cutoff, so the pair SET is the same. Rows are
// sorted by neighbour index, making the list deterministic.
// CylinderExclVolume::vectorize (CylinderExclVolume.cpp:34-179) then drops
// pairs whose neighbour is the cylinder's mother (getBranchingCylinder()).
#include <hipcub/hipcub.hpp>

#include <cmath>

#include "mdy_device.hpp"
#include "mdy_eval.hpp"
#include "mdy_launch.hpp"

namespace mdy {

__global__ void k_fill_int(int *p, int n, int v) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = v;
}

__global__ void __launch_bounds__(kBlock) k_mid_bbox(int ncyl, const int *b1, const int *b2, const double *x,
                                                     double *mid, double *red) {
    __shared__ double lds[17 * 6];
    double v[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncyl; c += gridDim.x * blockDim.x) {
        const double *p1 = x + b1[c], *p2 = x + b2[c];
        for (int k = 0; k < 3; k++) {
            double m = 0.5 * (p1[k] + p2[k]);   // Cylinder::updateCoordinate (Cylinder.cpp:45-49)
            mid[3 * c + k] = m;
            v[k] = fmin(v[k], m);
            v[3 + k] = fmax(v[3 + k], m);
        }
    }
    block_reduce<6>(v, lds, 0x38u, 0x07u);
    if (threadIdx.x < 6) red[(size_t)blockIdx.x * kRed + threadIdx.x] = v[threadIdx.x];
}

__global__ void __launch_bounds__(kBlock) k_bbox_final(const double *red, int nblk, double *bbox) {
    __shared__ double lds[17 * 6];
    double v[6];
    grid_partials_reduce<6>(red, nblk, v, lds, 0x38u, 0x07u);
    if (threadIdx.x < 6) bbox[threadIdx.x] = v[threadIdx.x];
}

struct Grid {
    double ox, oy, oz, h;
    int nx, ny, nz;
    __device__ __forceinline__ int clampi(int v, int n) const { return v < 0 ? 0 : (v >= n ? n - 1 : v); }
    __device__ __forceinline__ void cell3(const double *m, int &cx, int &cy, int &cz) const {
        cx = clampi((int)floor((m[0] - ox) / h), nx);
        cy = clampi((int)floor((m[1] - oy) / h), ny);
        cz = clampi((int)floor((m[2] - oz) / h), nz);
    }
};

__global__ void k_cell_assign(int ncyl, const double *mid, Grid g, int *cell_of, int *cell_cnt) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncyl; c += gridDim.x * blockDim.x) {
        int cx, cy, cz;
        g.cell3(mid + 3 * c, cx, cy, cz);
        int cell = (cz * g.ny + cy) * g.nx + cx;
        cell_of[c] = cell;
        atomicAdd(cell_cnt + cell, 1);
    }
}

__global__ void k_cell_fill(int ncyl, const int *cell_of, const int *cell_off, int *cell_cur, int *cell_list) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncyl; c += gridDim.x * blockDim.x) {
        int cell = cell_of[c];
        cell_list[cell_off[cell] + atomicAdd(cell_cur + cell, 1)] = c;
    }
}

struct CylView {
    const int *b1, *b2, *fil, *pos, *child;
    const long long *id;
    const short *type;
    const double *eq, *ek;
    const double *seg;   // monomer-sum form: interval * monomer size
    int mon;             // monomer-sum excluded volume
};

// Pairs the excluded-volume force field drops: the integral form skips a
// neighbour that is the cylinder's branching mother
// (cin->getBranchingCylinder() == ci, CylinderExclVolume.cpp:85, 116); the
// monomer-sum form skips mother-daughter pairs in either order
// (CylinderVolumeMon.hpp:134-136).
__device__ __forceinline__ bool ev_dropped(const CylView &C, int c, int n) {
    return C.child[n] == c || (C.mon && C.child[c] == n);
}

// the reference predicate for candidate n of cylinder c
__device__ __forceinline__ bool nl_pred(const CylView &C, const double *mid, int c, int n, double rminsq,
                                        double rmaxsq) {
    if (C.id[c] <= C.id[n]) return false;
    if (C.fil[c] == C.fil[n]) {
        int ds = C.pos[c] - C.pos[n];
        if (ds < 0) ds = -ds;
        if (ds <= 2) return false;   // ChemParams::minCylinderDistanceSameFilament (SysParams.h:315)
    }
    const double *m1 = mid + 3 * c, *m2 = mid + 3 * n;
    double d2 = (m2[0] - m1[0]) * (m2[0] - m1[0]);
    d2 += (m2[1] - m1[1]) * (m2[1] - m1[1]);
    d2 += (m2[2] - m1[2]) * (m2[2] - m1[2]);
    if (d2 < rminsq || d2 > rmaxsq) return false;
    return true;
}

template <class F>
__device__ __forceinline__ void for_candidates(const Grid &g, const double *mid, const int *cell_off,
                                               const int *cell_list, int c, F &&fn) {
    int cx, cy, cz;
    g.cell3(mid + 3 * c, cx, cy, cz);
    for (int dz = -1; dz <= 1; dz++) {
        int z = cz + dz;
        if (z < 0 || z >= g.nz) continue;
        for (int dy = -1; dy <= 1; dy++) {
            int y = cy + dy;
            if (y < 0 || y >= g.ny) continue;
            for (int dx = -1; dx <= 1; dx++) {
                int xx = cx + dx;
                if (xx < 0 || xx >= g.nx) continue;
                int cell = (z * g.ny + y) * g.nx + xx;
                for (int k = cell_off[cell], e = cell_off[cell + 1]; k < e; k++) fn(cell_list[k]);
            }
        }
    }
}

__global__ void k_pair_count(int ncyl, CylView C, const double *mid, Grid g, const int *cell_off,
                             const int *cell_list, double rminsq, double rmaxsq, int *row_cnt, int *row_keep) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncyl; c += gridDim.x * blockDim.x) {
        int cnt = 0, keep = 0;
        for_candidates(g, mid, cell_off, cell_list, c, [&](int n) {
            if (!nl_pred(C, mid, c, n, rminsq, rmaxsq)) return;
            cnt++;
            if (!ev_dropped(C, c, n)) keep++;
        });
        row_cnt[c] = cnt;
        row_keep[c] = keep;
    }
}

__global__ void k_pair_emit(int ncyl, CylView C, const double *mid, Grid g, const int *cell_off,
                            const int *cell_list, double rminsq, double rmaxsq, const int *row_off,
                            const int *row_keep_off, int2 *pairs, int4 *ev_bs, double *ev_kr, int2 *ev_cyl) {
    for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < ncyl; c += gridDim.x * blockDim.x) {
        int2 *row = pairs + row_off[c];
        int len = 0;
        for_candidates(g, mid, cell_off, cell_list, c, [&](int n) {
            if (!nl_pred(C, mid, c, n, rminsq, rmaxsq)) return;
            row[len++] = make_int2(c, n);
        });
        for (int a = 1; a < len; a++) {  // insertion sort by neighbour index
            int2 v = row[a];
            int b = a - 1;
            while (b >= 0 && row[b].y > v.y) { row[b + 1] = row[b]; b--; }
            row[b + 1] = v;
        }
        int o = row_keep_off[c];
        for (int a = 0; a < len; a++) {
            int n = row[a].y;
            if (ev_dropped(C, c, n)) continue;
            ev_bs[o] = make_int4(C.b1[c], C.b2[c], C.b1[n], C.b2[n]);
            if (C.mon) {
                // kvol = max of the two constants; coeff = kvol * segEqLen1 * segEqLen2
                // (CylinderVolumeMon.hpp:43, 160-163)
                ev_kr[o] = smax(C.ek[c], C.ek[n]) * C.seg[c] * C.seg[n];
                ev_cyl[o] = make_int2(c, n);
            } else {
                double krep = (C.type[c] != C.type[n]) ? smax(C.ek[c], C.ek[n]) : C.ek[c];
                ev_kr[o] = krep * C.eq[c] * C.eq[n];
            }
            o++;
        }
    }
}

// BoundaryCylinderNL (NeighborListImpl.cpp:738-751): plane distance of
// Cylinder::coordinate = midPointCoordinate(b1, b2, 0.5) below the float cutoff.
__global__ void k_bd_flag(int ncyl, int np, const int *b1, const int *b2, const unsigned char *minus,
                          const double *x, const double *planes, float cutoff, int *flag) {
    const long long tot = (long long)np * ncyl;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < tot;
         i += (long long)gridDim.x * blockDim.x) {
        int p = (int)(i / ncyl), c = (int)(i % ncyl);
        const double *p1 = x + b1[c], *p2 = x + b2[c];
        d3 m = d3{p1[0] * (1.0 - 0.5) + 0.5 * p2[0], p1[1] * (1.0 - 0.5) + 0.5 * p2[1],
                  p1[2] * (1.0 - 0.5) + 0.5 * p2[2]};
        double dist = plane_dist(planes + 4 * p, m);
        flag[i] = (dist < (double)cutoff) ? (minus[c] ? 2 : 1) : 0;
    }
}

// BoundaryCylinderRepulsion::vectorize (BoundaryCylinderRepulsion.cpp:33-88)
__global__ void k_bd_emit(int ncyl, int np, const int *b1, const int *b2, const int *flag, const int *off,
                          const double *pk, const double *psl, int *bd_bead, int *bd_plane, double *bd_k,
                          double *bd_sl) {
    const long long tot = (long long)np * ncyl;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < tot;
         i += (long long)gridDim.x * blockDim.x) {
        int f = flag[i];
        if (!f) continue;
        int p = (int)(i / ncyl), c = (int)(i % ncyl);
        int o = off[i];
        if (f == 2) { bd_bead[o] = b1[c]; bd_plane[o] = p; bd_k[o] = pk[p]; bd_sl[o] = psl[p]; o++; }
        bd_bead[o] = b2[c]; bd_plane[o] = p; bd_k[o] = pk[p]; bd_sl[o] = psl[p];
    }
}

// dynamic CSR: bead -> excluded-volume / boundary slots, rows sorted by slot
__device__ __forceinline__ int dyn_slot_bead(const Dev &D, int j, int &slot) {
    int nev4 = 4 * D.n_ev;
    if (j < nev4) {
        int4 b = D.ev_bs[j >> 2];
        int o = (j & 3) == 0 ? b.x : (j & 3) == 1 ? b.y : (j & 3) == 2 ? b.z : b.w;
        slot = D.slot_ev + j;
        return o / 3;
    }
    j -= nev4;
    slot = D.slot_bd + j;
    return D.bd_bead[j] / 3;
}
__global__ void k_dy_count(Dev D, int *cnt) {
    const int tot = 4 * D.n_ev + D.n_bd;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < tot; j += gridDim.x * blockDim.x) {
        int s;
        atomicAdd(cnt + dyn_slot_bead(D, j, s), 1);
    }
}
__global__ void k_dy_fill(Dev D, const int *off, int *cur, int *idx) {
    const int tot = 4 * D.n_ev + D.n_bd;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < tot; j += gridDim.x * blockDim.x) {
        int s;
        int b = dyn_slot_bead(D, j, s);
        idx[off[b] + atomicAdd(cur + b, 1)] = s;
    }
}
__global__ void k_dy_sort(int nb, const int *off, int *idx) {
    for (int b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        int s0 = off[b], e = off[b + 1];
        for (int a = s0 + 1; a < e; a++) {
            int v = idx[a], k = a - 1;
            while (k >= s0 && idx[k] > v) { idx[k + 1] = idx[k]; k--; }
            idx[k + 1] = v;
        }
    }
}

// ---------------------------------------------------------------------------
static int grid1(const Ctx &c, long long n) {
    long long g = (n + kBlock - 1) / kBlock;
    long long cap = (long long)c.num_cu * 8;
    return (int)(g < 1 ? 1 : (g > cap ? cap : g));
}

int exclusive_scan(Ctx &c, const int *in, int *out, int n, hipStream_t st) {
    // out[0..n]: exclusive prefix of in[0..n-1] followed by the total
    size_t tmp = 0;
    MDY_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp, in, out, n + 1, st));
    MDY_TRY(ensure(c.scan_tmp, tmp / sizeof(int) + 1));
    size_t cap = c.scan_tmp.n * sizeof(int);
    // the (n+1)-th input is read past `in`; callers allocate n+1 entries and zero the last
    MDY_HIP(hipcub::DeviceScan::ExclusiveSum(c.scan_tmp.p, cap, in, out, n + 1, st));
    return MDY_OK;
}

int build_neighbors(Ctx &c, const double *d_x) {
    hipStream_t st = c.stream;
    const int ncyl = c.ncyl;
    c.n_ev = 0;
    c.n_bd = 0;
    c.n_pairs = 0;
    CylView C{c.c_b1.p, c.c_b2.p, c.c_fil.p, c.c_pos.p, c.c_child.p, c.c_id.p, c.c_type.p, c.c_eq.p, c.c_ek.p,
              c.c_seg.p, c.ev_mon};
    if (c.ev_on && ncyl > 1) {
        MDY_TRY(ensure(c.mid, (size_t)3 * ncyl));
        const int g = grid1(c, ncyl);
        MDY_TRY(ensure(c.red, (size_t)kRed * c.num_cu * 64));
        hipLaunchKernelGGL(k_mid_bbox, dim3(g), dim3(kBlock), 0, st, ncyl, c.c_b1.p, c.c_b2.p, d_x, c.mid.p, c.red.p);
        hipLaunchKernelGGL(k_bbox_final, dim3(1), dim3(kBlock), 0, st, c.red.p, g, c.bbox.p);
        double bb[6];
        MDY_HIP(hipMemcpyAsync(bb, c.bbox.p, sizeof(bb), hipMemcpyDeviceToHost, st));
        MDY_HIP(hipStreamSynchronize(st));
        for (int k = 0; k < 6; k++)
            if (!std::isfinite(bb[k])) return MDY_ERR_NONFINITE;
        const float rmaxsq_f = c.ev_cut * c.ev_cut;   // float, HybridNeighborListImpl.h:78-80
        const float rminsq_f = 0.0f * 0.0f;
        Grid gr;
        gr.h = std::sqrt((double)rmaxsq_f) * (1.0 + 1e-6) + 1e-6;
        double ext[3] = {bb[3] - bb[0], bb[4] - bb[1], bb[5] - bb[2]};
        for (;;) {
            gr.nx = (int)std::floor(ext[0] / gr.h) + 1;
            gr.ny = (int)std::floor(ext[1] / gr.h) + 1;
            gr.nz = (int)std::floor(ext[2] / gr.h) + 1;
            if ((double)gr.nx * gr.ny * gr.nz <= 4.0e6) break;
            gr.h *= 1.25;   // larger cells stay correct, just prune less
        }
        gr.ox = bb[0]; gr.oy = bb[1]; gr.oz = bb[2];
        const int ncell = gr.nx * gr.ny * gr.nz;
        MDY_TRY(ensure(c.cell_of, ncyl));
        MDY_TRY(ensure(c.cell_cnt, (size_t)ncell + 1));
        MDY_TRY(ensure(c.cell_cur, (size_t)ncell + 1));
        MDY_TRY(ensure(c.cell_off, (size_t)ncell + 2));
        MDY_TRY(ensure(c.cell_list, ncyl));
        hipLaunchKernelGGL(k_fill_int, dim3(grid1(c, ncell + 1)), dim3(kBlock), 0, st, c.cell_cnt.p, ncell + 1, 0);
        hipLaunchKernelGGL(k_fill_int, dim3(grid1(c, ncell + 1)), dim3(kBlock), 0, st, c.cell_cur.p, ncell + 1, 0);
        hipLaunchKernelGGL(k_cell_assign, dim3(g), dim3(kBlock), 0, st, ncyl, c.mid.p, gr, c.cell_of.p, c.cell_cnt.p);
        MDY_TRY(exclusive_scan(c, c.cell_cnt.p, c.cell_off.p, ncell, st));
        hipLaunchKernelGGL(k_cell_fill, dim3(g), dim3(kBlock), 0, st, ncyl, c.cell_of.p, c.cell_off.p, c.cell_cur.p,
                           c.cell_list.p);
        MDY_TRY(ensure(c.row_cnt, (size_t)ncyl + 1));
        MDY_TRY(ensure(c.row_keep, (size_t)ncyl + 1));
        MDY_TRY(ensure(c.row_off, (size_t)ncyl + 2));
        MDY_TRY(ensure(c.row_keep_off, (size_t)ncyl + 2));
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, c.row_cnt.p + ncyl, 1, 0);
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, c.row_keep.p + ncyl, 1, 0);
        hipLaunchKernelGGL(k_pair_count, dim3(g), dim3(kBlock), 0, st, ncyl, C, c.mid.p, gr, c.cell_off.p,
                           c.cell_list.p, (double)rminsq_f, (double)rmaxsq_f, c.row_cnt.p, c.row_keep.p);
        MDY_TRY(exclusive_scan(c, c.row_cnt.p, c.row_off.p, ncyl, st));
        MDY_TRY(exclusive_scan(c, c.row_keep.p, c.row_keep_off.p, ncyl, st));
        int tot[2];
        MDY_HIP(hipMemcpyAsync(&tot[0], c.row_off.p + ncyl, sizeof(int), hipMemcpyDeviceToHost, st));
        MDY_HIP(hipMemcpyAsync(&tot[1], c.row_keep_off.p + ncyl, sizeof(int), hipMemcpyDeviceToHost, st));
        MDY_HIP(hipStreamSynchronize(st));
        c.n_pairs = tot[0];
        c.n_ev = tot[1];
        MDY_TRY(ensure(c.pairs, (size_t)2 * tot[0] + 2));
        MDY_TRY(ensure(c.ev_bs, (size_t)tot[1] + 1));
        MDY_TRY(ensure(c.ev_kr, (size_t)tot[1] + 1));
        if (c.ev_mon) MDY_TRY(ensure(c.ev_cyl, (size_t)tot[1] + 1));
        hipLaunchKernelGGL(k_pair_emit, dim3(g), dim3(kBlock), 0, st, ncyl, C, c.mid.p, gr, c.cell_off.p,
                           c.cell_list.p, (double)rminsq_f, (double)rmaxsq_f, c.row_off.p, c.row_keep_off.p,
                           (int2 *)c.pairs.p, c.ev_bs.p, c.ev_kr.p, c.ev_mon ? c.ev_cyl.p : nullptr);
        MDY_HIP(hipGetLastError());
    }
    if (c.bd_on && c.n_planes > 0 && ncyl > 0) {
        const long long tot = (long long)c.n_planes * ncyl;
        MDY_TRY(ensure(c.bd_flag, (size_t)tot + 1));
        MDY_TRY(ensure(c.bd_off, (size_t)tot + 2));
        hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, st, c.bd_flag.p + tot, 1, 0);
        hipLaunchKernelGGL(k_bd_flag, dim3(grid1(c, tot)), dim3(kBlock), 0, st, ncyl, c.n_planes, c.c_b1.p, c.c_b2.p,
                           c.c_minus.p, d_x, c.planes.p, c.bd_cut, c.bd_flag.p);
        MDY_TRY(exclusive_scan(c, c.bd_flag.p, c.bd_off.p, (int)tot, st));
        std::vector<int> offs(c.n_planes + 1);
        for (int p = 0; p <= c.n_planes; p++)
            MDY_HIP(hipMemcpyAsync(&offs[p], c.bd_off.p + (size_t)p * ncyl, sizeof(int), hipMemcpyDeviceToHost, st));
        MDY_HIP(hipStreamSynchronize(st));
        c.n_bd = offs[c.n_planes];
        c.h_bd_nneigh.assign(c.n_planes, 0);
        for (int p = 0; p < c.n_planes; p++) c.h_bd_nneigh[p] = offs[p + 1] - offs[p];
        MDY_TRY(ensure(c.bd_bead, (size_t)c.n_bd + 1));
        MDY_TRY(ensure(c.bd_plane, (size_t)c.n_bd + 1));
        MDY_TRY(ensure(c.bd_k, (size_t)c.n_bd + 1));
        MDY_TRY(ensure(c.bd_sl, (size_t)c.n_bd + 1));
        hipLaunchKernelGGL(k_bd_emit, dim3(grid1(c, tot)), dim3(kBlock), 0, st, ncyl, c.n_planes, c.c_b1.p, c.c_b2.p,
                           c.bd_flag.p, c.bd_off.p, c.plane_k.p, c.plane_sl.p, c.bd_bead.p, c.bd_plane.p, c.bd_k.p,
                           c.bd_sl.p);
        MDY_HIP(hipGetLastError());
    } else {
        c.h_bd_nneigh.assign(c.n_planes, 0);
    }
    // dynamic CSR over the new slots
    MDY_TRY(ensure(c.ev_bs, 1));
    MDY_TRY(ensure(c.ev_kr, 1));
    if (c.ev_mon) MDY_TRY(ensure(c.ev_cyl, 1));
    MDY_TRY(ensure(c.bd_bead, 1));
    MDY_TRY(ensure(c.bd_plane, 1));
    MDY_TRY(ensure(c.bd_k, 1));
    MDY_TRY(ensure(c.bd_sl, 1));
    c.have_nl = true;
    Dev D = c.dev();
    const int nb = c.nb;
    const int ndyn = 4 * c.n_ev + c.n_bd;
    MDY_TRY(ensure(c.dy_cnt, (size_t)nb + 1));
    MDY_TRY(ensure(c.dy_cur, (size_t)nb + 1));
    MDY_TRY(ensure(c.dy_off, (size_t)nb + 2));
    MDY_TRY(ensure(c.dy_idx, (size_t)ndyn + 1));
    hipLaunchKernelGGL(k_fill_int, dim3(grid1(c, nb + 1)), dim3(kBlock), 0, st, c.dy_cnt.p, nb + 1, 0);
    hipLaunchKernelGGL(k_fill_int, dim3(grid1(c, nb + 1)), dim3(kBlock), 0, st, c.dy_cur.p, nb + 1, 0);
    D = c.dev();
    if (ndyn > 0) hipLaunchKernelGGL(k_dy_count, dim3(grid1(c, ndyn)), dim3(kBlock), 0, st, D, c.dy_cnt.p);
    MDY_TRY(exclusive_scan(c, c.dy_cnt.p, c.dy_off.p, nb, st));
    if (ndyn > 0) {
        hipLaunchKernelGGL(k_dy_fill, dim3(grid1(c, ndyn)), dim3(kBlock), 0, st, D, c.dy_off.p, c.dy_cur.p, c.dy_idx.p);
        hipLaunchKernelGGL(k_dy_sort, dim3(grid1(c, nb)), dim3(kBlock), 0, st, nb, c.dy_off.p, c.dy_idx.p);
    }
    MDY_HIP(hipGetLastError());
    // evaluation scratch sized for the new slot count
    MDY_TRY(ensure(c.contrib, (size_t)3 * D.n_slots + 3));
    return MDY_OK;
}

}  // namespace mdy
