// mdy_api.hip -- the C ABI of libmdy_hip.so (declared in include/mdy.h).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <new>
#include <string>
#include <unordered_map>
#include <vector>

#include "mdy_launch.hpp"

namespace mdy {

Dev Ctx::dev() const {
    Dev D{};
    D.nb = nb;
    D.ndof = 3 * nb;
    D.n_ev = have_nl ? n_ev : 0;
    D.n_fs = n_fs; D.n_fb = n_fb; D.n_lm = n_ls + n_ms; D.n_ls = n_ls; D.n_bs = n_bs; D.n_bb = n_bb;
    D.n_bdh = n_bdh; D.n_bp = n_bp;
    D.n_bd = have_nl ? n_bd : 0;
    D.ev_bs = ev_bs.p; D.ev_kr = ev_kr.p;
    D.ev_mon = ev_mon; D.ev_cyl = ev_cyl.p; D.cyl_mon = c_mon.p;
    D.bdh_quad = bdh_quad;
    D.fs_bs = fs_bs.p; D.fs_k = fs_k.p; D.fs_l0 = fs_l0.p;
    D.fb_bs = fb_bs.p; D.fb_k = fb_k.p; D.fb_th = fb_th.p; D.fb_ct = fb_ct.p; D.fb_st = fb_st.p;
    D.lm_bs = lm_bs.p; D.lm_k = lm_k.p; D.lm_l0 = lm_l0.p; D.lm_p1 = lm_p1.p; D.lm_p2 = lm_p2.p;
    D.bs_bs = bs_bs.p; D.bs_k = bs_k.p; D.bs_l0 = bs_l0.p; D.bs_p = bs_p.p;
    D.bb_bs = bb_bs.p; D.bb_k = bb_k.p; D.bb_th = bb_th.p; D.bb_ct = bb_ct.p; D.bb_st = bb_st.p;
    D.bdh_bs = bdh_bs.p; D.bdh_k = bdh_k.p; D.bdh_pos = bdh_pos.p;
    D.bp_bs = bp_bs.p; D.bp_k = bp_k.p; D.bp_pos = bp_pos.p;
    D.bd_bead = bd_bead.p; D.bd_plane = bd_plane.p; D.bd_k = bd_k.p; D.bd_sl = bd_sl.p; D.planes = planes.p;
    D.slot_fs = 0;
    D.slot_fb = D.slot_fs + 2 * n_fs;
    D.slot_lm = D.slot_fb + 3 * n_fb;
    D.slot_bs = D.slot_lm + 4 * (n_ls + n_ms);
    D.slot_bb = D.slot_bs + 3 * n_bs;
    D.slot_bdh = D.slot_bb + 4 * n_bb;
    D.slot_bp = D.slot_bdh + 4 * n_bdh;
    D.slot_ev = D.slot_bp + 3 * n_bp;
    D.slot_bd = D.slot_ev + 4 * D.n_ev;
    D.n_slots = D.slot_bd + D.n_bd;
    D.st_off = st_off.p; D.st_idx = st_idx.p; D.dy_off = dy_off.p; D.dy_idx = dy_idx.p;
    D.coplanar_mode = opts.coplanar_mode;
    D.evq = nullptr;   // set by the one-shot launchers in quadrature mode
    D.cq = std::cos(0.5 * M_PI);   // BranchingPositionCosine.cpp:235, 372 (posheta = pi/2)
    D.sq = std::sin(0.5 * M_PI);
    D.qe = qe.p; D.qe_n = qn.p; D.qf = qf.p; D.qf_n = qn.p + 1; D.err = err.p;
    return D;
}

template <class T>
static int upload(DBuf<T> &b, const T *h, size_t n, hipStream_t st) {
    MDY_TRY(ensure(b, n));
    if (n && h) MDY_HIP(hipMemcpyAsync(b.p, h, n * sizeof(T), hipMemcpyHostToDevice, st));
    return MDY_OK;
}

static bool bad_offsets(const int32_t *bs, size_t n, int nb) {
    if (n && !bs) return true;
    for (size_t i = 0; i < n; i++)
        if (bs[i] < 0 || bs[i] % 3 != 0 || bs[i] / 3 >= nb) return true;
    return false;
}

// a force field with interactions needs every one of its parameter arrays
static bool missing_params(const mdy_ff_tables &t) {
    if (t.fs_n > 0 && (!t.fs_k || !t.fs_l0)) return true;
    if (t.fb_n > 0 && (!t.fb_k || !t.fb_theta0)) return true;
    if (t.ls_n > 0 && (!t.ls_k || !t.ls_l0 || !t.ls_pos1 || !t.ls_pos2)) return true;
    if (t.ms_n > 0 && (!t.ms_k || !t.ms_l0 || !t.ms_pos1 || !t.ms_pos2)) return true;
    if (t.bs_n > 0 && (!t.bs_k || !t.bs_l0 || !t.bs_pos)) return true;
    if (t.bb_n > 0 && (!t.bb_k || !t.bb_theta0)) return true;
    if (t.bdh_n > 0 && (!t.bdh_k || !t.bdh_pos)) return true;
    if (t.bp_n > 0 && (!t.bp_k || !t.bp_pos)) return true;
    return false;
}

}  // namespace mdy

using namespace mdy;

struct mdy_ctx {
    Ctx c;
};

static int fail(mdy_ctx *ctx, int code, const std::string &msg) {
    if (ctx) ctx->c.msg = msg;
    return code;
}
static int hipfail(mdy_ctx *ctx, const char *what) {
    hipError_t e = hipGetLastError();
    return fail(ctx, MDY_ERR_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

extern "C" {

int mdy_abi_version(void) { return MDY_ABI_VERSION; }

int mdy_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

void mdy_default_opts(mdy_opts *o) {
    std::memset(o, 0, sizeof(*o));
    o->precision = MDY_PREC_FP64;
    o->coplanar_mode = MDY_COPLANAR_REFERENCE;
}

static const char *bad_opts(const mdy_opts *o) {
    if (o->precision != MDY_PREC_FP64 && o->precision != MDY_PREC_MIXED) return "precision";
    if (o->coplanar_mode < MDY_COPLANAR_REFERENCE || o->coplanar_mode > MDY_COPLANAR_QUADRATURE)
        return "coplanar_mode";
    if (o->coplanar_mode == MDY_COPLANAR_QUADRATURE && o->precision != MDY_PREC_FP64)
        return "the quadrature mode needs an FP64 context";
    return nullptr;
}

int mdy_create(mdy_ctx **out, int device, const mdy_opts *opts) {
    if (!out) return MDY_ERR_ARG;
    *out = nullptr;
    int n = mdy_device_count();
    if (n <= 0) return MDY_ERR_NOGPU;
    if (device < 0 || device >= n) return MDY_ERR_ARG;
    mdy_ctx *ctx = new (std::nothrow) mdy_ctx();
    if (!ctx) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    c.device = device;
    if (opts) c.opts = *opts;
    else mdy_default_opts(&c.opts);
    if (bad_opts(&c.opts)) { delete ctx; return MDY_ERR_ARG; }
    if (hipSetDevice(device) != hipSuccess) { delete ctx; return MDY_ERR_HIP; }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) { delete ctx; return MDY_ERR_HIP; }
    c.num_cu = prop.multiProcessorCount;
    if (!prop.cooperativeLaunch) { delete ctx; return MDY_ERR_LAUNCH; }
    if (hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking) != hipSuccess) { delete ctx; return MDY_ERR_HIP; }
    if (ensure(c.qn, 16) || ensure(c.err, 4) || ensure(c.out, 16) || ensure(c.bbox, 8) || ensure(c.qe, 3 * kQueueCap) ||
        ensure(c.qf, 3 * kQueueCap) || ensure(c.qe_u, 3 * kQueueCap) || ensure(c.red, (size_t)kRed * c.num_cu * 64) ||
        ensure(c.simpson_part, (size_t)kRed * c.num_cu * 8)) {
        mdy_destroy(ctx);
        return MDY_ERR_HIP;
    }
    (void)hipMemsetAsync(c.err.p, 0, 4 * sizeof(int), c.stream);
    (void)hipMemsetAsync(c.qn.p, 0, 16 * sizeof(int), c.stream);
    if (resolve_noalloc_scratch(device) != MDY_OK) {
        mdy_destroy(ctx);
        return MDY_ERR_HIP;
    }
    *out = ctx;
    return MDY_OK;
}

void mdy_destroy(mdy_ctx *ctx) {
    if (!ctx) return;
    Ctx &c = ctx->c;
    (void)hipSetDevice(c.device);
    if (c.stream) (void)hipStreamSynchronize(c.stream);
    release(c.fs_bs); release(c.fs_k); release(c.fs_l0); release(c.fb_bs); release(c.fb_k); release(c.fb_th); release(c.fb_ct); release(c.fb_st); release(c.bb_ct); release(c.bb_st);
    release(c.lm_bs); release(c.lm_k); release(c.lm_l0); release(c.lm_p1); release(c.lm_p2);
    release(c.bs_bs); release(c.bs_k); release(c.bs_l0); release(c.bs_p); release(c.bb_bs); release(c.bb_k);
    release(c.bb_th); release(c.st_off); release(c.st_idx);
    release(c.bdh_bs); release(c.bdh_k); release(c.bdh_pos); release(c.bp_bs); release(c.bp_k); release(c.bp_pos);
    release(c.c_b1); release(c.c_b2); release(c.c_fil); release(c.c_pos); release(c.c_child); release(c.c_id);
    release(c.c_type); release(c.c_eq); release(c.c_ek); release(c.c_minus); release(c.c_plus);
    release(c.c_mon); release(c.c_seg); release(c.ev_cyl);
    release(c.planes); release(c.plane_k); release(c.plane_sl);
    release(c.mid); release(c.cell_of); release(c.cell_cnt); release(c.cell_off); release(c.cell_cur);
    release(c.cell_list); release(c.row_cnt); release(c.row_keep); release(c.row_off); release(c.row_keep_off);
    release(c.pairs); release(c.ev_bs); release(c.ev_kr); release(c.bd_flag); release(c.bd_off); release(c.bd_bead);
    release(c.bd_plane); release(c.bd_k); release(c.bd_sl); release(c.dy_cnt); release(c.dy_off); release(c.dy_cur);
    release(c.dy_idx); release(c.scan_tmp); release(c.bbox); release(c.x); release(c.xt); release(c.f);
    release(c.batch_args); release(c.batch_map); release(c.batch_done); release(c.evq); release(c.qpan); release(c.qcnt); release(c.qoff);
    release(c.qmap); release(c.qpart); release(c.qpe); release(c.qtot);
    release(c.contrib); release(c.red); release(c.out); release(c.qe); release(c.qf); release(c.qn); release(c.err);
    release(c.qe_u); release(c.simpson_part); release(c.F0); release(c.F1); release(c.S0); release(c.S1);
    release(c.xmin); release(c.bar); release(c.cg_out); release(c.tr_lam); release(c.tr_e); release(c.tr_mf);
    release(c.tr_beta); release(c.tr_safe); release(c.prof); release(c.prof_kind); release(c.acc);
    if (c.stream) (void)hipStreamDestroy(c.stream);
    delete ctx;
}

const char *mdy_last_error(const mdy_ctx *ctx) { return ctx ? ctx->c.msg.c_str() : "null context"; }

int mdy_set_opts(mdy_ctx *ctx, const mdy_opts *o) {
    if (!ctx || !o) return MDY_ERR_ARG;
    if (const char *m = bad_opts(o)) return fail(ctx, MDY_ERR_ARG, m);
    ctx->c.opts = *o;
    return MDY_OK;
}

int mdy_set_network(mdy_ctx *ctx, const mdy_system *s) {
    if (!ctx || !s) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    (void)hipSetDevice(c.device);
    // the previous network is dropped before validation: a rejected call
    // leaves the context with no network, never with the old tables
    c.have_sys = false;
    c.have_nl = false;
    c.have_dir = false;
    const int nb = s->n_beads;
    const mdy_ff_tables &t = s->ff;
    const mdy_cylinders &cy = s->cyl;
    if (nb <= 0) return fail(ctx, MDY_ERR_ARG, "n_beads must be positive");
    if (t.fs_n < 0 || t.fb_n < 0 || t.ls_n < 0 || t.ms_n < 0 || t.bs_n < 0 || t.bb_n < 0 || t.bdh_n < 0 ||
        t.bp_n < 0 || cy.n < 0 || s->bnd.n_planes < 0)
        return fail(ctx, MDY_ERR_ARG, "negative count");
    if (bad_offsets(t.fs_beadset, 2 * (size_t)t.fs_n, nb) || bad_offsets(t.fb_beadset, 3 * (size_t)t.fb_n, nb) ||
        bad_offsets(t.ls_beadset, 4 * (size_t)t.ls_n, nb) || bad_offsets(t.ms_beadset, 4 * (size_t)t.ms_n, nb) ||
        bad_offsets(t.bs_beadset, 3 * (size_t)t.bs_n, nb) || bad_offsets(t.bb_beadset, 4 * (size_t)t.bb_n, nb) ||
        bad_offsets(t.bdh_beadset, 4 * (size_t)t.bdh_n, nb) || bad_offsets(t.bp_beadset, 3 * (size_t)t.bp_n, nb) ||
        bad_offsets(cy.bead1, cy.n, nb) || bad_offsets(cy.bead2, cy.n, nb))
        return fail(ctx, MDY_ERR_ARG, "bead offset out of range (must be 3*index < 3*n_beads)");
    if (missing_params(t))
        return fail(ctx, MDY_ERR_ARG, "null parameter array for a force field with interactions");
    for (int i = 0; i < cy.n; i++)
        if (cy.branch_child && (cy.branch_child[i] < -1 || cy.branch_child[i] >= cy.n))
            return fail(ctx, MDY_ERR_ARG, "branch_child out of range");
    if (cy.n > 0 && (!cy.id || !cy.filament || !cy.position || !cy.eq_length || !cy.exvol_k || !cy.minus_end))
        return fail(ctx, MDY_ERR_ARG, "incomplete cylinder table");
    if (s->bnd.n_planes > 0 && (!s->bnd.plane || !s->bnd.k || !s->bnd.screen))
        return fail(ctx, MDY_ERR_ARG, "incomplete boundary table");
    if (s->exvol_type != MDY_EXVOL_INTEGRAL && s->exvol_type != MDY_EXVOL_MONOMER)
        return fail(ctx, MDY_ERR_ARG, "exvol_type");
    if (s->br_dihedral_type != MDY_BRDIH_COSINE && s->br_dihedral_type != MDY_BRDIH_QUADRATIC)
        return fail(ctx, MDY_ERR_ARG, "br_dihedral_type");
    const bool mon = s->exvol_enabled && s->exvol_type == MDY_EXVOL_MONOMER && cy.n > 0;
    if (mon) {
        if (!cy.mon_minus || !cy.mon_plus || !cy.mon_size || !cy.mon_interval)
            return fail(ctx, MDY_ERR_ARG, "monomer-sum excluded volume needs the monomer tables");
        for (int i = 0; i < cy.n; i++)
            if (cy.mon_interval[i] < 1 || cy.mon_plus[i] < cy.mon_minus[i] || !std::isfinite(cy.mon_size[i]))
                return fail(ctx, MDY_ERR_ARG, "monomer table: interval >= 1 and plus >= minus serial required");
    }
    hipStream_t st = c.stream;
    c.nb = nb;
    c.n_fs = t.fs_n; c.n_fb = t.fb_n; c.n_ls = t.ls_n; c.n_ms = t.ms_n; c.n_bs = t.bs_n; c.n_bb = t.bb_n;
    c.n_bdh = t.bdh_n; c.n_bp = t.bp_n;
    int rc = MDY_OK;
    rc |= upload(c.fs_bs, reinterpret_cast<const int2 *>(t.fs_beadset), t.fs_n, st);
    rc |= upload(c.fs_k, t.fs_k, t.fs_n, st);
    rc |= upload(c.fs_l0, t.fs_l0, t.fs_n, st);
    rc |= upload(c.fb_bs, t.fb_beadset, 3 * (size_t)t.fb_n, st);
    rc |= upload(c.fb_k, t.fb_k, t.fb_n, st);
    rc |= upload(c.fb_th, t.fb_theta0, t.fb_n, st);
    // cos/sin of the equilibrium angles from the host C library, the values the
    // CPU path computes inside its loops (FilamentBendingCosine.cpp:226, 299)
    std::vector<double> fct(t.fb_n), fst(t.fb_n), bct(t.bb_n), bst(t.bb_n);
    for (int i = 0; i < t.fb_n; i++) { fct[i] = std::cos(t.fb_theta0[i]); fst[i] = std::sin(t.fb_theta0[i]); }
    for (int i = 0; i < t.bb_n; i++) { bct[i] = std::cos(t.bb_theta0[i]); bst[i] = std::sin(t.bb_theta0[i]); }
    rc |= upload(c.fb_ct, fct.data(), t.fb_n, st);
    rc |= upload(c.fb_st, fst.data(), t.fb_n, st);
    rc |= upload(c.bb_ct, bct.data(), t.bb_n, st);
    rc |= upload(c.bb_st, bst.data(), t.bb_n, st);
    // linkers then motors in one spring table
    {
        const size_t nl = (size_t)t.ls_n + t.ms_n;
        std::vector<int4> lb(nl);
        std::vector<double> k(nl), l0(nl), p1(nl), p2(nl);
        for (int i = 0; i < t.ls_n; i++) {
            lb[i] = make_int4(t.ls_beadset[4 * i], t.ls_beadset[4 * i + 1], t.ls_beadset[4 * i + 2], t.ls_beadset[4 * i + 3]);
            k[i] = t.ls_k[i]; l0[i] = t.ls_l0[i]; p1[i] = t.ls_pos1[i]; p2[i] = t.ls_pos2[i];
        }
        for (int i = 0; i < t.ms_n; i++) {
            size_t j = t.ls_n + i;
            lb[j] = make_int4(t.ms_beadset[4 * i], t.ms_beadset[4 * i + 1], t.ms_beadset[4 * i + 2], t.ms_beadset[4 * i + 3]);
            k[j] = t.ms_k[i]; l0[j] = t.ms_l0[i]; p1[j] = t.ms_pos1[i]; p2[j] = t.ms_pos2[i];
        }
        rc |= upload(c.lm_bs, lb.data(), nl, st);
        rc |= upload(c.lm_k, k.data(), nl, st);
        rc |= upload(c.lm_l0, l0.data(), nl, st);
        rc |= upload(c.lm_p1, p1.data(), nl, st);
        rc |= upload(c.lm_p2, p2.data(), nl, st);
        if (hipStreamSynchronize(st) != hipSuccess) return hipfail(ctx, "upload");
    }
    rc |= upload(c.bs_bs, t.bs_beadset, 3 * (size_t)t.bs_n, st);
    rc |= upload(c.bs_k, t.bs_k, t.bs_n, st);
    rc |= upload(c.bs_l0, t.bs_l0, t.bs_n, st);
    rc |= upload(c.bs_p, t.bs_pos, t.bs_n, st);
    rc |= upload(c.bb_bs, reinterpret_cast<const int4 *>(t.bb_beadset), t.bb_n, st);
    rc |= upload(c.bb_k, t.bb_k, t.bb_n, st);
    rc |= upload(c.bb_th, t.bb_theta0, t.bb_n, st);
    rc |= upload(c.bdh_bs, reinterpret_cast<const int4 *>(t.bdh_beadset), t.bdh_n, st);
    rc |= upload(c.bdh_k, t.bdh_k, t.bdh_n, st);
    rc |= upload(c.bdh_pos, t.bdh_pos, t.bdh_n, st);
    rc |= upload(c.bp_bs, t.bp_beadset, 3 * (size_t)t.bp_n, st);
    rc |= upload(c.bp_k, t.bp_k, t.bp_n, st);
    rc |= upload(c.bp_pos, t.bp_pos, t.bp_n, st);
    // static CSR (bead -> bonded / cross-link slots, rows in slot order)
    {
        std::vector<int> owner;
        owner.reserve(2 * (size_t)t.fs_n + 3 * (size_t)t.fb_n + 4 * ((size_t)t.ls_n + t.ms_n) + 3 * (size_t)t.bs_n +
                      4 * (size_t)t.bb_n);
        for (size_t i = 0; i < 2 * (size_t)t.fs_n; i++) owner.push_back(t.fs_beadset[i] / 3);
        for (size_t i = 0; i < 3 * (size_t)t.fb_n; i++) owner.push_back(t.fb_beadset[i] / 3);
        for (size_t i = 0; i < 4 * (size_t)t.ls_n; i++) owner.push_back(t.ls_beadset[i] / 3);
        for (size_t i = 0; i < 4 * (size_t)t.ms_n; i++) owner.push_back(t.ms_beadset[i] / 3);
        for (size_t i = 0; i < 3 * (size_t)t.bs_n; i++) owner.push_back(t.bs_beadset[i] / 3);
        for (size_t i = 0; i < 4 * (size_t)t.bb_n; i++) owner.push_back(t.bb_beadset[i] / 3);
        for (size_t i = 0; i < 4 * (size_t)t.bdh_n; i++) owner.push_back(t.bdh_beadset[i] / 3);
        for (size_t i = 0; i < 3 * (size_t)t.bp_n; i++) owner.push_back(t.bp_beadset[i] / 3);
        std::vector<int> off(nb + 1, 0), idx(owner.size()), cur;
        for (int o : owner) off[o + 1]++;
        for (int b = 0; b < nb; b++) off[b + 1] += off[b];
        cur.assign(off.begin(), off.end() - 1);
        for (size_t s2 = 0; s2 < owner.size(); s2++) idx[cur[owner[s2]]++] = (int)s2;
        rc |= upload(c.st_off, off.data(), off.size(), st);
        rc |= upload(c.st_idx, idx.data(), idx.size(), st);
        if (hipStreamSynchronize(st) != hipSuccess) return hipfail(ctx, "upload");
    }
    // cylinders
    c.ncyl = cy.n;
    {
        std::vector<int> child(cy.n, -1);
        std::vector<short> type(cy.n, 0);
        for (int i = 0; i < cy.n; i++) {
            if (cy.branch_child) child[i] = cy.branch_child[i];
            if (cy.type) type[i] = cy.type[i];
        }
        std::vector<long long> id(cy.n);
        for (int i = 0; i < cy.n; i++) id[i] = cy.id[i];
        rc |= upload(c.c_b1, cy.bead1, cy.n, st);
        rc |= upload(c.c_b2, cy.bead2, cy.n, st);
        rc |= upload(c.c_fil, cy.filament, cy.n, st);
        rc |= upload(c.c_pos, cy.position, cy.n, st);
        rc |= upload(c.c_child, child.data(), cy.n, st);
        rc |= upload(c.c_id, id.data(), cy.n, st);
        rc |= upload(c.c_type, type.data(), cy.n, st);
        rc |= upload(c.c_eq, cy.eq_length, cy.n, st);
        rc |= upload(c.c_ek, cy.exvol_k, cy.n, st);
        rc |= upload(c.c_minus, cy.minus_end, cy.n, st);
        // isPlusEnd(): the highest position on its filament
        std::unordered_map<int, int> last;
        for (int i = 0; i < cy.n; i++) {
            auto it = last.find(cy.filament[i]);
            if (it == last.end() || cy.position[i] > cy.position[it->second]) last[cy.filament[i]] = i;
        }
        std::vector<unsigned char> plus(cy.n, 0);
        for (const auto &kv : last) plus[kv.second] = 1;
        rc |= upload(c.c_plus, plus.data(), cy.n, st);
        c.ev_mon = mon ? 1 : 0;
        c.bdh_quad = s->br_dihedral_type == MDY_BRDIH_QUADRATIC ? 1 : 0;
        if (mon) {
            // segEqLen = monInterval * monEqLen (CylinderVolumeMon.hpp:41-43)
            std::vector<int4> m(cy.n);
            std::vector<double> seg(cy.n);
            for (int i = 0; i < cy.n; i++) {
                m[i] = make_int4(cy.mon_minus[i], cy.mon_plus[i], cy.mon_interval[i], 0);
                seg[i] = (double)cy.mon_interval[i] * cy.mon_size[i];
            }
            rc |= upload(c.c_mon, m.data(), cy.n, st);
            rc |= upload(c.c_seg, seg.data(), cy.n, st);
        }
        if (hipStreamSynchronize(st) != hipSuccess) return hipfail(ctx, "upload");
    }
    c.ev_cut = s->exvol_cutoff;
    c.ev_on = s->exvol_enabled;
    c.bd_on = s->boundary_enabled;
    c.n_planes = s->bnd.n_planes;
    c.bd_cut = s->bnd.cutoff;
    rc |= upload(c.planes, s->bnd.plane, 4 * (size_t)c.n_planes, st);
    rc |= upload(c.plane_k, s->bnd.k, c.n_planes, st);
    rc |= upload(c.plane_sl, s->bnd.screen, c.n_planes, st);
    rc |= ensure(c.x, 3 * (size_t)nb);
    rc |= ensure(c.xt, 3 * (size_t)nb);
    rc |= ensure(c.f, 3 * (size_t)nb);
    if (rc != MDY_OK) return fail(ctx, MDY_ERR_HIP, "device allocation failed");
    if (hipStreamSynchronize(st) != hipSuccess) return hipfail(ctx, "upload");
    c.have_sys = true;
    return MDY_OK;
}

int mdy_build_neighbors_dev(mdy_ctx *ctx, const double *d_coord) {
    if (!ctx || !d_coord) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.have_sys) return fail(ctx, MDY_ERR_STATE, "no network set");
    (void)hipSetDevice(c.device);
    int rc = build_neighbors(c, d_coord);
    if (rc != MDY_OK) return fail(ctx, rc, "neighbor-list build failed");
    if (hipStreamSynchronize(c.stream) != hipSuccess) return hipfail(ctx, "neighbor-list build");
    return MDY_OK;
}

int mdy_build_neighbors(mdy_ctx *ctx, const double *coord) {
    if (!ctx || !coord) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.have_sys) return fail(ctx, MDY_ERR_STATE, "no network set");
    (void)hipSetDevice(c.device);
    if (hipMemcpyAsync(c.x.p, coord, sizeof(double) * 3 * c.nb, hipMemcpyHostToDevice, c.stream) != hipSuccess)
        return hipfail(ctx, "coord upload");
    return mdy_build_neighbors_dev(ctx, c.x.p);
}

int mdy_get_search_dir(mdy_ctx *ctx, double *dir_out) {
    if (!ctx || !dir_out) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    (void)hipSetDevice(c.device);
    const int rc = final_search_dir(c, dir_out);
    if (rc == MDY_ERR_STATE) return fail(ctx, rc, "no completed minimization on this network");
    if (rc != MDY_OK) return fail(ctx, rc, "search direction readback failed");
    return MDY_OK;
}

int mdy_get_pairs(mdy_ctx *ctx, int32_t *pairs, int64_t cap, int64_t *n) {
    if (!ctx || !n) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.have_nl) return fail(ctx, MDY_ERR_STATE, "no neighbor list");
    *n = c.n_pairs;
    if (!pairs) return MDY_OK;
    if (cap < c.n_pairs) return fail(ctx, MDY_ERR_CAPACITY, "pairs buffer too small");
    if (c.n_pairs &&
        hipMemcpy(pairs, c.pairs.p, sizeof(int32_t) * 2 * c.n_pairs, hipMemcpyDeviceToHost) != hipSuccess)
        return hipfail(ctx, "pairs copy");
    return MDY_OK;
}

int mdy_get_exvol_table(mdy_ctx *ctx, int32_t *beadset, double *krep, int64_t cap, int64_t *n) {
    if (!ctx || !n) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.have_nl) return fail(ctx, MDY_ERR_STATE, "no neighbor list");
    *n = c.n_ev;
    if (!beadset && !krep) return MDY_OK;
    if (cap < c.n_ev) return fail(ctx, MDY_ERR_CAPACITY, "exvol buffer too small");
    if (c.n_ev && beadset &&
        hipMemcpy(beadset, c.ev_bs.p, sizeof(int4) * c.n_ev, hipMemcpyDeviceToHost) != hipSuccess)
        return hipfail(ctx, "exvol copy");
    if (c.n_ev && krep && hipMemcpy(krep, c.ev_kr.p, sizeof(double) * c.n_ev, hipMemcpyDeviceToHost) != hipSuccess)
        return hipfail(ctx, "exvol copy");
    return MDY_OK;
}

int mdy_get_boundary_table(mdy_ctx *ctx, int32_t *nneighbors, int32_t *beadset, int64_t cap, int64_t *n) {
    if (!ctx || !n) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.have_nl) return fail(ctx, MDY_ERR_STATE, "no neighbor list");
    *n = c.n_bd;
    if (nneighbors)
        for (int p = 0; p < c.n_planes; p++) nneighbors[p] = c.h_bd_nneigh[p];
    if (!beadset) return MDY_OK;
    if (cap < c.n_bd) return fail(ctx, MDY_ERR_CAPACITY, "beadset buffer too small");
    if (c.n_bd && hipMemcpy(beadset, c.bd_bead.p, sizeof(int32_t) * c.n_bd, hipMemcpyDeviceToHost) != hipSuccess)
        return hipfail(ctx, "boundary copy");
    return MDY_OK;
}

int mdy_energy_dev(mdy_ctx *ctx, const double *d_coord, double *d_out, void *stream) {
    if (!ctx || !d_coord || !d_out) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.have_sys || !c.have_nl) return fail(ctx, MDY_ERR_STATE, "network / neighbor list missing");
    (void)hipSetDevice(c.device);
    hipStream_t st = stream ? (hipStream_t)stream : c.stream;
    int rc = launch_energy(c, d_coord, c.out.p, st);
    if (rc != MDY_OK) return hipfail(ctx, "energy");
    // d_out: MDY_NFF energies then the total (sentinel form)
    if (hipMemcpyAsync(d_out, c.out.p, sizeof(double) * (MDY_NFF + 1), hipMemcpyDeviceToDevice, st) != hipSuccess)
        return hipfail(ctx, "energy copy");
    return MDY_OK;
}

int mdy_energy_dev_noalloc(mdy_ctx *ctx, const double *d_coord, double *d_out, void *stream) {
    if (!ctx || !d_coord || !d_out) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.have_sys || !c.have_nl) return fail(ctx, MDY_ERR_STATE, "network / neighbor list missing");
    if (c.opts.coplanar_mode == MDY_COPLANAR_QUADRATURE)
        return fail(ctx, MDY_ERR_ARG, "the quadrature mode allocates its per-pair flags: use mdy_energy_dev");
    (void)hipSetDevice(c.device);
    hipStream_t st = stream ? (hipStream_t)stream : c.stream;
    const int rc = launch_energy_noalloc(c, d_coord, d_out, st);
    if (rc == MDY_ERR_CAPACITY)
        return fail(ctx, rc, "the grid needs more block records than the static scratch holds (2048)");
    if (rc == MDY_ERR_STATE) return fail(ctx, rc, "static scratch not resolved for this device");
    if (rc != MDY_OK) return hipfail(ctx, "energy (no allocation)");
    return MDY_OK;
}

int mdy_energy(mdy_ctx *ctx, const double *coord, double *e_total, double *e_per_ff) {
    if (!ctx || !coord) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.have_sys || !c.have_nl) return fail(ctx, MDY_ERR_STATE, "network / neighbor list missing");
    (void)hipSetDevice(c.device);
    if (hipMemcpyAsync(c.xt.p, coord, sizeof(double) * 3 * c.nb, hipMemcpyHostToDevice, c.stream) != hipSuccess)
        return hipfail(ctx, "coord upload");
    int rc = launch_energy(c, c.xt.p, c.out.p, c.stream);
    if (rc != MDY_OK) return hipfail(ctx, "energy");
    double h[MDY_NFF + 2];
    if (hipMemcpyAsync(h, c.out.p, sizeof(h), hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess)
        return hipfail(ctx, "energy readback");
    int e = 0;
    (void)hipMemcpy(&e, c.err.p, sizeof(int), hipMemcpyDeviceToHost);
    if (e) return fail(ctx, e, "device error during energy");
    if (e_total) *e_total = h[MDY_NFF];
    if (e_per_ff)
        for (int k = 0; k < MDY_NFF; k++) e_per_ff[k] = h[k];
    return MDY_OK;
}

int mdy_forces_dev(mdy_ctx *ctx, const double *d_coord, double *d_force, int accumulate, void *stream) {
    if (!ctx || !d_coord || !d_force) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.have_sys || !c.have_nl) return fail(ctx, MDY_ERR_STATE, "network / neighbor list missing");
    (void)hipSetDevice(c.device);
    hipStream_t st = stream ? (hipStream_t)stream : c.stream;
    int rc = launch_forces(c, d_coord, d_force, accumulate, st);
    if (rc == MDY_ERR_FORCE) return fail(ctx, rc, "excluded-volume force non-finite after quadrature fallback");
    if (rc != MDY_OK) return hipfail(ctx, "forces");
    return MDY_OK;
}

int mdy_forces(mdy_ctx *ctx, const double *coord, double *force_accum) {
    if (!ctx || !coord || !force_accum) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.have_sys || !c.have_nl) return fail(ctx, MDY_ERR_STATE, "network / neighbor list missing");
    (void)hipSetDevice(c.device);
    const size_t n = 3 * (size_t)c.nb;
    if (hipMemcpyAsync(c.xt.p, coord, sizeof(double) * n, hipMemcpyHostToDevice, c.stream) != hipSuccess ||
        hipMemcpyAsync(c.f.p, force_accum, sizeof(double) * n, hipMemcpyHostToDevice, c.stream) != hipSuccess)
        return hipfail(ctx, "upload");
    int rc = mdy_forces_dev(ctx, c.xt.p, c.f.p, 1, c.stream);
    if (rc != MDY_OK) return rc;
    if (hipMemcpyAsync(force_accum, c.f.p, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
        hipStreamSynchronize(c.stream) != hipSuccess)
        return hipfail(ctx, "force readback");
    return MDY_OK;
}

int mdy_quadrature_pair_count(mdy_ctx *ctx, const double *coord, int32_t *count) {
    if (!ctx || !coord || !count) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.have_sys || !c.have_nl) return fail(ctx, MDY_ERR_STATE, "network / neighbor list missing");
    (void)hipSetDevice(c.device);
    if (hipMemcpyAsync(c.xt.p, coord, sizeof(double) * 3 * c.nb, hipMemcpyHostToDevice, c.stream) != hipSuccess)
        return hipfail(ctx, "coord upload");
    int n = 0;
    if (quad_pair_count(c, c.xt.p, c.stream, &n) != MDY_OK) return hipfail(ctx, "quadrature pair count");
    *count = n;
    return MDY_OK;
}

int mdy_hvp_dev(mdy_ctx *ctx, const double *d_coord, const double *d_v, double *d_hv, void *stream) {
    if (!ctx || !d_coord || !d_v || !d_hv) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.have_sys || !c.have_nl) return fail(ctx, MDY_ERR_STATE, "network / neighbor list missing");
    (void)hipSetDevice(c.device);
    hipStream_t st = stream ? (hipStream_t)stream : c.stream;
    int rc = launch_hvp(c, d_coord, d_v, d_hv, st);
    if (rc != MDY_OK) return fail(ctx, rc, "hvp failed");
    return MDY_OK;
}

int mdy_hvp_dev_async(mdy_ctx *ctx, const double *d_coord, const double *d_v, double *d_hv, void *stream) {
    if (!ctx || !d_coord || !d_v || !d_hv) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.have_sys || !c.have_nl) return fail(ctx, MDY_ERR_STATE, "network / neighbor list missing");
    (void)hipSetDevice(c.device);
    hipStream_t st = stream ? (hipStream_t)stream : c.stream;
    int rc = launch_hvp(c, d_coord, d_v, d_hv, st, false);
    if (rc != MDY_OK) return fail(ctx, rc, "hvp launch failed");
    return MDY_OK;
}

int mdy_sync_errors(mdy_ctx *ctx, void *stream) {
    if (!ctx) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    (void)hipSetDevice(c.device);
    hipStream_t st = stream ? (hipStream_t)stream : c.stream;
    if (!c.err.p) return MDY_OK;
    int rc = sync_errors(c, st);
    if (rc != MDY_OK) return fail(ctx, rc, "a device call enqueued earlier failed");
    return MDY_OK;
}

int mdy_energy_jvp_dev(mdy_ctx *ctx, const double *d_coord, const double *d_w, double *d_out, void *stream) {
    if (!ctx || !d_coord || !d_w || !d_out) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.have_sys || !c.have_nl) return fail(ctx, MDY_ERR_STATE, "network / neighbor list missing");
    (void)hipSetDevice(c.device);
    hipStream_t st = stream ? (hipStream_t)stream : c.stream;
    int rc = launch_energy_jvp(c, d_coord, d_w, d_out, st);
    if (rc != MDY_OK) return fail(ctx, rc, "energy jvp failed");
    return MDY_OK;
}

int mdy_minimize_dev(mdy_ctx *ctx, const mdy_cg_params *p, double *d_coord, double *d_force_out,
                     mdy_min_result *result) {
    if (!ctx || !p || !d_coord || !result) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.have_sys || !c.have_nl) return fail(ctx, MDY_ERR_STATE, "network / neighbor list missing");
    if (p->method < MDY_CG_STEEPEST || p->method > MDY_CG_POLAK_RIBIERE) return fail(ctx, MDY_ERR_ARG, "method");
    if (p->line_search != MDY_LS_BACKTRACKING && p->line_search != MDY_LS_QUADRATIC)
        return fail(ctx, MDY_ERR_ARG, "line_search");
    if (c.opts.coplanar_mode == MDY_COPLANAR_QUADRATURE)
        return fail(ctx, MDY_ERR_ARG, "the quadrature mode serves one-shot evaluations; minimize in another mode");
    (void)hipSetDevice(c.device);
    int rc = run_minimize(c, p, d_coord, d_force_out, result);
    if (rc == MDY_ERR_LAUNCH) return fail(ctx, rc, "cooperative launch failed");
    if (rc == MDY_ERR_TIMEOUT) return fail(ctx, rc, "grid barrier watchdog fired");
    if (rc != MDY_OK) return hipfail(ctx, "minimize");
    int e = 0;
    (void)hipMemcpy(&e, c.err.p, sizeof(int), hipMemcpyDeviceToHost);
    if (e) return fail(ctx, e, "device error during minimization");
    if (result->status != MDY_OK) {
        const char *m = result->status == MDY_ERR_NONFINITE ? "maxF is not finite"
                        : result->status == MDY_ERR_FORCE   ? "excluded-volume force non-finite"
                        : result->status == MDY_ERR_RECOVERY ? "recovery step is infinite"
                                                             : "minimization error";
        return fail(ctx, result->status, m);
    }
    return MDY_OK;
}

int mdy_minimize_batch(mdy_ctx *const *ctxs, int n, const mdy_cg_params *params, double *const *d_coord,
                       double *const *d_force_out, mdy_min_result *results, double *net_ms,
                       int threads_per_block, const double *weights, int first_pass_calls, void *stream,
                       double *total_ms) {
    if (!ctxs || n <= 0 || !params || !d_coord || !results) return MDY_ERR_ARG;
    if (threads_per_block == 0) threads_per_block = 64;
    if (threads_per_block != 64 && threads_per_block != 128 && threads_per_block != 256) {
        if (ctxs[0]) ctxs[0]->c.msg = "threads_per_block must be 64, 128 or 256";
        return MDY_ERR_ARG;
    }
    std::vector<Ctx *> cs(n);
    for (int i = 0; i < n; i++) {
        if (!ctxs[i] || !d_coord[i]) return MDY_ERR_ARG;
        Ctx &c = ctxs[i]->c;
        if (!c.have_sys || !c.have_nl) return fail(ctxs[0], MDY_ERR_STATE, "network / neighbor list missing");
        if (c.device != ctxs[0]->c.device || c.opts.precision != ctxs[0]->c.opts.precision)
            return fail(ctxs[0], MDY_ERR_ARG, "batched contexts must share device and precision");
        if (c.opts.coplanar_mode == MDY_COPLANAR_QUADRATURE)
            return fail(ctxs[0], MDY_ERR_ARG, "the quadrature mode serves one-shot evaluations; minimize in another mode");
        if (params[i].method < MDY_CG_STEEPEST || params[i].method > MDY_CG_POLAK_RIBIERE)
            return fail(ctxs[0], MDY_ERR_ARG, "method");
        if (params[i].line_search != MDY_LS_BACKTRACKING && params[i].line_search != MDY_LS_QUADRATIC)
            return fail(ctxs[0], MDY_ERR_ARG, "line_search");
        for (int j = 0; j < i; j++)
            if (ctxs[j] == ctxs[i]) return fail(ctxs[0], MDY_ERR_ARG, "a context appears twice in the batch");
        cs[i] = &c;
    }
    Ctx &c0 = ctxs[0]->c;
    (void)hipSetDevice(c0.device);
    hipStream_t st = stream ? (hipStream_t)stream : c0.stream;
    if (first_pass_calls < -1024) return fail(ctxs[0], MDY_ERR_ARG, "first_pass_calls must be >= -1024");
    int rc = run_minimize_batch(cs.data(), n, params, d_coord, d_force_out, results, net_ms, threads_per_block,
                                weights, first_pass_calls, st, total_ms);
    if (rc == MDY_ERR_LAUNCH) return fail(ctxs[0], rc, "cooperative launch failed");
    if (rc != MDY_OK) return hipfail(ctxs[0], "minimize batch");
    int worst = MDY_OK;
    for (int i = 0; i < n; i++) {
        int e = 0;
        (void)hipMemcpy(&e, cs[i]->err.p, sizeof(int), hipMemcpyDeviceToHost);
        if (e && results[i].status == MDY_OK) results[i].status = e;
        if (results[i].status != MDY_OK) {
            ctxs[i]->c.msg = "minimization error in batch";
            if (worst == MDY_OK) worst = results[i].status;
        }
    }
    if (worst != MDY_OK) return fail(ctxs[0], worst, "a minimization of the batch failed (see results[].status)");
    return MDY_OK;
}

int mdy_stretch_forces(mdy_ctx *ctx, const double *coord, double *linker, double *motor, double *branch3) {
    if (!ctx || !coord) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.have_sys) return fail(ctx, MDY_ERR_STATE, "no network set");
    (void)hipSetDevice(c.device);
    const Dev D = c.dev();
    const int nbr = std::max(D.n_bs, std::max(D.n_bdh, D.n_bp));
    const size_t n = (size_t)D.n_ls + (D.n_lm - D.n_ls) + 3 * (size_t)nbr;
    if (ensure(c.out, n + 16) != MDY_OK) return hipfail(ctx, "allocation");
    double *dl = c.out.p, *dm = dl + D.n_ls, *db = dm + (D.n_lm - D.n_ls);
    if (hipMemcpyAsync(c.xt.p, coord, sizeof(double) * 3 * c.nb, hipMemcpyHostToDevice, c.stream) != hipSuccess)
        return hipfail(ctx, "coord upload");
    if (launch_stretch_forces(c, c.xt.p, dl, dm, db, c.stream) != MDY_OK) return hipfail(ctx, "stretch forces");
    if ((linker && D.n_ls &&
         hipMemcpyAsync(linker, dl, sizeof(double) * D.n_ls, hipMemcpyDeviceToHost, c.stream) != hipSuccess) ||
        (motor && D.n_lm > D.n_ls &&
         hipMemcpyAsync(motor, dm, sizeof(double) * (D.n_lm - D.n_ls), hipMemcpyDeviceToHost, c.stream) !=
             hipSuccess) ||
        (branch3 && nbr &&
         hipMemcpyAsync(branch3, db, sizeof(double) * 3 * nbr, hipMemcpyDeviceToHost, c.stream) != hipSuccess) ||
        hipStreamSynchronize(c.stream) != hipSuccess)
        return hipfail(ctx, "stretch force readback");
    return MDY_OK;
}

int mdy_load_forces(mdy_ctx *ctx, const double *coord, int n_types, const double *mon_size, const int32_t *n_mon,
                    int max_mon, double *load_plus, double *load_minus) {
    if (!ctx || !coord || n_types <= 0 || !mon_size || !n_mon || max_mon <= 0 || !load_plus || !load_minus)
        return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.have_sys || !c.have_nl) return fail(ctx, MDY_ERR_STATE, "network / neighbor list missing");
    for (int t = 0; t < n_types; t++)
        if (n_mon[t] < 0 || n_mon[t] > max_mon) return fail(ctx, MDY_ERR_ARG, "n_mon[type] must lie in [0, max_mon]");
    (void)hipSetDevice(c.device);
    const size_t n = (size_t)c.nb * max_mon;
    double *d = nullptr;
    if (hipMalloc((void **)&d, sizeof(double) * (2 * n + n_types) + sizeof(int) * n_types) != hipSuccess)
        return hipfail(ctx, "allocation");
    double *lp = d, *lm = d + n, *ms = d + 2 * n;
    int *nm = reinterpret_cast<int *>(ms + n_types);
    int rc = MDY_OK;
    if (hipMemcpyAsync(c.xt.p, coord, sizeof(double) * 3 * c.nb, hipMemcpyHostToDevice, c.stream) != hipSuccess ||
        hipMemcpyAsync(ms, mon_size, sizeof(double) * n_types, hipMemcpyHostToDevice, c.stream) != hipSuccess ||
        hipMemcpyAsync(nm, n_mon, sizeof(int) * n_types, hipMemcpyHostToDevice, c.stream) != hipSuccess)
        rc = MDY_ERR_HIP;
    if (rc == MDY_OK) rc = launch_load_forces(c, c.xt.p, n_types, ms, nm, max_mon, lp, lm, c.stream);
    if (rc == MDY_OK &&
        (hipMemcpyAsync(load_plus, lp, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
         hipMemcpyAsync(load_minus, lm, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream) != hipSuccess ||
         hipStreamSynchronize(c.stream) != hipSuccess))
        rc = MDY_ERR_HIP;
    (void)hipFree(d);
    if (rc != MDY_OK) return hipfail(ctx, "load forces");
    return MDY_OK;
}

int mdy_minimize(mdy_ctx *ctx, const mdy_cg_params *p, double *coord_inout, double *force_out,
                 mdy_min_result *result) {
    if (!ctx || !p || !coord_inout || !result) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    if (!c.have_sys) return fail(ctx, MDY_ERR_STATE, "no network set");
    (void)hipSetDevice(c.device);
    const size_t n = 3 * (size_t)c.nb;
    if (hipMemcpyAsync(c.x.p, coord_inout, sizeof(double) * n, hipMemcpyHostToDevice, c.stream) != hipSuccess)
        return hipfail(ctx, "coord upload");
    if (!c.have_nl) {
        int rc = build_neighbors(c, c.x.p);
        if (rc != MDY_OK) return fail(ctx, rc, "neighbor-list build failed");
    }
    int rc = mdy_minimize_dev(ctx, p, c.x.p, force_out ? c.f.p : nullptr, result);
    if (hipMemcpyAsync(coord_inout, c.x.p, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream) != hipSuccess)
        return hipfail(ctx, "coord readback");
    if (force_out && rc == MDY_OK &&
        hipMemcpyAsync(force_out, c.f.p, sizeof(double) * n, hipMemcpyDeviceToHost, c.stream) != hipSuccess)
        return hipfail(ctx, "force readback");
    if (hipStreamSynchronize(c.stream) != hipSuccess) return hipfail(ctx, "sync");
    return rc;
}

int mdy_set_trace(mdy_ctx *ctx, int64_t cap) {
    if (!ctx || cap < 0) return MDY_ERR_ARG;
    ctx->c.trace_cap = cap;
    return MDY_OK;
}

int mdy_get_trace(mdy_ctx *ctx, double *lambda, double *energy, double *maxf, double *beta, int32_t *safe,
                  int64_t cap, int64_t *n) {
    if (!ctx || !n) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    long long m = c.trace_n < c.trace_cap ? c.trace_n : c.trace_cap;
    *n = c.trace_n;
    if (m > cap) m = cap;
    if (m <= 0) return MDY_OK;
    if ((lambda && hipMemcpy(lambda, c.tr_lam.p, m * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) ||
        (energy && hipMemcpy(energy, c.tr_e.p, m * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) ||
        (maxf && hipMemcpy(maxf, c.tr_mf.p, m * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) ||
        (beta && hipMemcpy(beta, c.tr_beta.p, m * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) ||
        (safe && hipMemcpy(safe, c.tr_safe.p, m * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess))
        return hipfail(ctx, "trace copy");
    return MDY_OK;
}

int mdy_set_profile(mdy_ctx *ctx, int64_t cap) {
    if (!ctx || cap < 0) return MDY_ERR_ARG;
    ctx->c.prof_cap = cap;
    return MDY_OK;
}

int mdy_get_profile(mdy_ctx *ctx, int32_t *kind, int64_t *ticks, int64_t cap, int64_t *n) {
    if (!ctx || !n) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    long long m = c.prof_n < c.prof_cap ? c.prof_n : c.prof_cap;
    *n = c.prof_n;
    if (cap < m) m = cap;
    if (m <= 0) return MDY_OK;
    std::vector<long long> rec(5 * (size_t)m);
    if ((kind && hipMemcpy(kind, c.prof_kind.p, m * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess) ||
        (ticks && hipMemcpy(rec.data(), c.prof.p, rec.size() * sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess))
        return hipfail(ctx, "profile copy");
    if (ticks)
        for (long long i = 0; i < m; i++) ticks[i] = rec[5 * i];
    return MDY_OK;
}

int mdy_get_profile_detail(mdy_ctx *ctx, int64_t *rec, int64_t cap, int64_t *n) {
    if (!ctx || !n) return MDY_ERR_ARG;
    Ctx &c = ctx->c;
    long long m = c.prof_n < c.prof_cap ? c.prof_n : c.prof_cap;
    *n = m;
    if (!rec) return MDY_OK;
    if (cap < m) m = cap;
    if (m > 0 && hipMemcpy(rec, c.prof.p, 5 * m * sizeof(int64_t), hipMemcpyDeviceToHost) != hipSuccess)
        return hipfail(ctx, "profile copy");
    return MDY_OK;
}

int mdy_last_timing(mdy_ctx *ctx, double *ms_total) {
    if (!ctx || !ms_total) return MDY_ERR_ARG;
    *ms_total = ctx->c.last_ms;
    return MDY_OK;
}

}  // extern "C"
