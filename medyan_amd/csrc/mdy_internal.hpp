// mdy_internal.hpp -- device-side tables and the library context.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "mdy.h"

namespace mdy {

constexpr int kNFF = MDY_NFF;
constexpr int kBlock = 256;          // threads per block of every kernel
constexpr int kQueueCap = 4096;      // excluded-volume quadrature fallback queue
constexpr int kRed = 32;             // 8-byte words per block in a reduction record

// Interaction tables on the device. The evaluation index space is
// [EV | FS | FB | LM | BS | BB | BD]: excluded volume first so its costly
// items are packed densely into the first waves of the grid.
// Force slots (3 doubles each) follow the reference force-field order
// FS, FB, LS, MS, BS, BB, EV, BD, so the per-bead gather adds contributions in
// the order ForceFieldManager::computeForces would.
struct Dev {
    int nb, ndof;
    int n_ev, n_fs, n_fb, n_lm, n_ls, n_bs, n_bb, n_bdh, n_bp, n_bd;
    const int4 *ev_bs; const double *ev_kr;   // monomer-sum form: ev_kr holds kvol * seg1 * seg2
    int ev_mon;                               // excluded volume as the monomer-pair sum
    int bdh_quad;                             // branching dihedral E = k theta^2 (else k (1 - cos theta))
    const int2 *ev_cyl;                       // monomer-sum form: the pair's cylinders
    const int4 *cyl_mon;                      // per cylinder: minus serial, plus serial, interval
    const int2 *fs_bs; const double *fs_k, *fs_l0;
    const int *fb_bs; const double *fb_k, *fb_th, *fb_ct, *fb_st;   // ct/st: cos/sin(theta0) from the host
    const int4 *lm_bs; const double *lm_k, *lm_l0, *lm_p1, *lm_p2;
    const int *bs_bs; const double *bs_k, *bs_l0, *bs_p;
    const int4 *bb_bs; const double *bb_k, *bb_th, *bb_ct, *bb_st;
    const int4 *bdh_bs; const double *bdh_k, *bdh_pos;           // branching dihedral
    const int *bp_bs; const double *bp_k, *bp_pos;               // branching position
    const int *bd_bead; const int *bd_plane; const double *bd_k, *bd_sl; const double *planes;
    int slot_fs, slot_fb, slot_lm, slot_bs, slot_bb, slot_bdh, slot_bp, slot_ev, slot_bd, n_slots;
    const int *st_off, *st_idx;   // static CSR: bead -> bonded/cross-link slots
    const int *dy_off, *dy_idx;   // dynamic CSR: bead -> excluded-volume/boundary slots
    int coplanar_mode;
    const unsigned char *evq;     // quadrature mode: per excluded-volume pair, 1 = evaluated by quadrature
    double cq, sq;                // cos and sin of pi/2 (C library), branching position
    // excluded-volume quadrature fallback queues (pair indices)
    int *qe, *qe_n;               // energy
    int *qf, *qf_n;               // force
    int *err;                     // sticky device error word (MDY_ERR_*)
    __host__ __device__ int n_items() const { return n_ev + n_fs + n_fb + n_lm + n_bs + n_bb + n_bdh + n_bp + n_bd; }
};

template <class T>
struct DBuf {
    T *p = nullptr;
    size_t n = 0;
};

struct Ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    mdy_opts opts{};
    std::string msg;
    int num_cu = 0;
    // system
    bool have_sys = false, have_nl = false;
    int nb = 0;
    int ncyl = 0;
    // host copies needed for rebuilds
    std::vector<int> h_static_counts;
    // device tables (static part)
    DBuf<int2> fs_bs; DBuf<double> fs_k, fs_l0;
    DBuf<int> fb_bs; DBuf<double> fb_k, fb_th, fb_ct, fb_st;
    DBuf<int4> lm_bs; DBuf<double> lm_k, lm_l0, lm_p1, lm_p2;
    DBuf<int> bs_bs; DBuf<double> bs_k, bs_l0, bs_p;
    DBuf<int4> bb_bs; DBuf<double> bb_k, bb_th, bb_ct, bb_st;
    DBuf<int4> bdh_bs; DBuf<double> bdh_k, bdh_pos;
    DBuf<int> bp_bs; DBuf<double> bp_k, bp_pos;
    DBuf<int> st_off, st_idx;
    int n_fs = 0, n_fb = 0, n_ls = 0, n_ms = 0, n_bs = 0, n_bb = 0, n_bdh = 0, n_bp = 0;
    // cylinders
    DBuf<int> c_b1, c_b2, c_fil, c_pos, c_child; DBuf<long long> c_id; DBuf<short> c_type;
    DBuf<double> c_eq, c_ek; DBuf<unsigned char> c_minus, c_plus;   // c_plus: last cylinder of its filament
    int ev_mon = 0;                // monomer-sum excluded volume
    int bdh_quad = 0;              // quadratic branching dihedral
    DBuf<int4> c_mon;              // minus serial, plus serial, interval, 0
    DBuf<double> c_seg;            // interval * monomer size
    float ev_cut = 0, bd_cut = 0;
    int ev_on = 0, bd_on = 0;
    int n_planes = 0;
    DBuf<double> planes, plane_k, plane_sl;
    // neighbor-list work
    DBuf<double> mid;             // 3 * ncyl
    DBuf<int> cell_of, cell_cnt, cell_off, cell_cur, cell_list;
    DBuf<int> row_cnt, row_keep, row_off, row_keep_off;
    DBuf<int> pairs;              // 2 * npairs (NL, before mother-daughter filtering)
    long long n_pairs = 0;
    DBuf<int4> ev_bs; DBuf<double> ev_kr; int n_ev = 0;
    DBuf<int2> ev_cyl;            // monomer-sum form: cylinders of each pair
    DBuf<int> bd_flag, bd_off, bd_bead, bd_plane; DBuf<double> bd_k, bd_sl; int n_bd = 0;
    std::vector<int> h_bd_nneigh;
    DBuf<int> dy_cnt, dy_off, dy_cur, dy_idx;
    DBuf<unsigned char> evq;      // quadrature mode: per-pair flags of the current evaluation
    DBuf<int> qpan, qcnt, qoff, qmap; // quadrature mode: panels, chunk counts, first chunks, chunk -> pair
    DBuf<int> qtot;               // quadrature mode: running chunk total (zero between calls)
    DBuf<double> qpart, qpe;      // quadrature mode: chunk records, per-pair values
    DBuf<int> scan_tmp;
    DBuf<double> bbox;            // 6
    // evaluation scratch
    DBuf<double> x, xt, f, contrib;
    DBuf<double> red;             // per-block partials
    DBuf<double> out;             // small results
    DBuf<int> qe, qf, qn, err;
    DBuf<double> qe_u, simpson_part;
    // CG state
    DBuf<double> F0, F1, S0, S1, xmin;
    DBuf<unsigned long long> bar, acc;
    DBuf<double> cg_out;          // result block
    DBuf<double> tr_lam, tr_e, tr_mf, tr_beta; DBuf<int> tr_safe;
    long long trace_cap = 0, trace_n = 0;
    DBuf<long long> prof; DBuf<int> prof_kind; long long prof_cap = 0, prof_n = 0;
    DBuf<unsigned char> batch_args;  // batched launch: per-network kernel arguments
    DBuf<int2> batch_map;            // batched launch: (network, block within its cluster)
    DBuf<int> batch_done;            // batched launch: finished minimizations (passes by completions)
    double last_ms = 0;
    // the last minimization's search direction (CGMethod::searchDir): the CG
    // buffer that holds it (0 F0, 1 F1, 2 S0, 3 S1) or, when its final update
    // s = F + beta d is still pending, the current force buffer, d and beta
    bool have_dir = false, dir_pend = false;
    int dir_buf = -1, dir_cur = 0, dir_prev = -1;
    double dir_beta = 0;
    int coop_blocks = 0;
    Dev dev() const;
};

}  // namespace mdy
