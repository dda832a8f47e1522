// netgen.cpp -- seeded synthetic actin networks for the benchmark workloads.
//
// Geometry rules (the same network for CPU and GPU runs of a seed):
//   * a cubic box of edge L with MEDYAN's six boundary planes inset 25 nm
//     (BoundaryCubic, src/Structure/BoundaryImpl.cpp:25-52);
//   * straight filaments of `cyl_per_filament` full-length cylinders; start
//     point uniform in the box, direction uniform on the sphere; both end
//     points must lie at least BoundaryCutoff/4 inside every plane, the rule of
//     FilamentInitializer (src/FilamentInitializer.cpp:112-131), and, a rule of
//     this generator, no segment may come closer than `min_sep` to an already
//     placed segment (avoids numerically singular initial overlaps);
//   * branched daughters (the last n_branches filaments) start BRSTRETCHINGL
//     from a point on a mother cylinder, at the branching angle to its axis;
//   * every bead is then jittered by N(0, jitter) per component;
//   * linkers and motors join binding points (one of 4 sites per cylinder) on
//     random cylinder pairs of different filaments closer than `pair_cut`;
//     rest length = initial distance + N(0, rest_sigma), redrawn until > 0.
// Random numbers: xoshiro256** seeded through SplitMix64, uniform doubles from
// the top 53 bits, normals by the Box-Muller transform.
#include "mdy.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <new>
#include <unordered_map>
#include <vector>

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t z = seed;
        for (int i = 0; i < 4; i++) {
            z += 0x9E3779B97F4A7C15ull;
            uint64_t t = z;
            t = (t ^ (t >> 30)) * 0xBF58476D1CE4E5B9ull;
            t = (t ^ (t >> 27)) * 0x94D049BB133111EBull;
            s[i] = t ^ (t >> 31);
        }
    }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    int below(int n) { return (int)(uniform() * n); }
    bool have_spare = false;
    double spare = 0;
    double normal() {
        if (have_spare) { have_spare = false; return spare; }
        double u1;
        do { u1 = uniform(); } while (u1 <= 0.0);
        double u2 = uniform();
        double r = std::sqrt(-2.0 * std::log(u1));
        spare = r * std::sin(2.0 * M_PI * u2);
        have_spare = true;
        return r * std::cos(2.0 * M_PI * u2);
    }
    void unit_vector(double *v) {
        double z = 2.0 * uniform() - 1.0;
        double phi = 2.0 * M_PI * uniform();
        double r = std::sqrt(1.0 - z * z);
        v[0] = r * std::cos(phi); v[1] = r * std::sin(phi); v[2] = z;
    }
};

struct V3 { double x, y, z; };
inline V3 operator+(V3 a, V3 b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
inline V3 operator-(V3 a, V3 b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
inline V3 operator*(double s, V3 a) { return {s * a.x, s * a.y, s * a.z}; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 crossv(V3 a, V3 b) { return {a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x}; }
inline double norm(V3 a) { return std::sqrt(dot(a, a)); }

// Closest distance between segments [p1,q1] and [p2,q2] (standard clamped
// parametric solution, e.g. Ericson, Real-Time Collision Detection, 5.1.9).
double seg_dist(V3 p1, V3 q1, V3 p2, V3 q2) {
    V3 d1 = q1 - p1, d2 = q2 - p2, r = p1 - p2;
    double a = dot(d1, d1), e = dot(d2, d2), f = dot(d2, r);
    double s, t;
    const double eps = 1e-12;
    if (a <= eps && e <= eps) return norm(r);
    if (a <= eps) { s = 0; t = std::fmin(std::fmax(f / e, 0.0), 1.0); }
    else {
        double c = dot(d1, r);
        if (e <= eps) { t = 0; s = std::fmin(std::fmax(-c / a, 0.0), 1.0); }
        else {
            double b = dot(d1, d2), den = a * e - b * b;
            s = den != 0 ? std::fmin(std::fmax((b * f - c * e) / den, 0.0), 1.0) : 0.0;
            t = (b * s + f) / e;
            if (t < 0) { t = 0; s = std::fmin(std::fmax(-c / a, 0.0), 1.0); }
            else if (t > 1) { t = 1; s = std::fmin(std::fmax((b - c) / a, 0.0), 1.0); }
        }
    }
    V3 c1 = p1 + s * d1, c2 = p2 + t * d2;
    return norm(c1 - c2);
}

// Uniform hash grid of segments (by midpoint) for the placement test.
struct SegGrid {
    double cell;
    std::unordered_map<int64_t, std::vector<int>> cells;
    std::vector<V3> a, b;
    std::vector<int> owner;
    explicit SegGrid(double c) : cell(c) {}
    static int64_t key(int64_t i, int64_t j, int64_t k) { return (i * 73856093LL) ^ (j * 19349663LL) ^ (k * 83492791LL); }
    void cidx(V3 m, int64_t *c) const {
        c[0] = (int64_t)std::floor(m.x / cell); c[1] = (int64_t)std::floor(m.y / cell); c[2] = (int64_t)std::floor(m.z / cell);
    }
    void add(V3 p, V3 q, int own) {
        int id = (int)a.size();
        a.push_back(p); b.push_back(q); owner.push_back(own);
        int64_t c[3];
        cidx(0.5 * (p + q), c);
        cells[key(c[0], c[1], c[2])].push_back(id);
    }
    // min distance of [p,q] to stored segments not owned by `skip`
    double min_dist(V3 p, V3 q, int skip) const {
        int64_t c[3];
        cidx(0.5 * (p + q), c);
        double best = 1e300;
        for (int64_t i = -1; i <= 1; i++)
            for (int64_t j = -1; j <= 1; j++)
                for (int64_t k = -1; k <= 1; k++) {
                    auto it = cells.find(key(c[0] + i, c[1] + j, c[2] + k));
                    if (it == cells.end()) continue;
                    for (int id : it->second) {
                        if (owner[id] == skip) continue;
                        double d = seg_dist(p, q, a[id], b[id]);
                        if (d < best) best = d;
                    }
                }
        return best;
    }
};

struct Storage {
    std::vector<double> coord;
    std::vector<int32_t> fil_first, fil_ncyl;
    // cylinders
    std::vector<int32_t> cb1, cb2, cfil, cpos, cchild;
    std::vector<int64_t> cid;
    std::vector<int16_t> ctype;
    std::vector<double> ceq, cek;
    std::vector<uint8_t> cminus;
    // tables
    std::vector<int32_t> fs_bs, fb_bs, ls_bs, ms_bs, bs_bs, bb_bs;
    std::vector<double> fs_k, fs_l0, fb_k, fb_t0, ls_k, ls_l0, ls_p1, ls_p2, ms_k, ms_l0, ms_p1, ms_p2;
    std::vector<double> bs_k, bs_l0, bs_p, bb_k, bb_t0;
    std::vector<int32_t> dh_bs, bp_bs;
    std::vector<double> dh_k, dh_p, bp_k, bp_p;
    std::vector<double> plane, pk, ps;
};

const double kSites[4] = {0.125, 0.375, 0.625, 0.875};

}  // namespace

extern "C" void mdy_gen_defaults(mdy_gen_params *p) {
    std::memset(p, 0, sizeof(*p));
    p->seed = 0x4D454459ull;
    p->n_filaments = 50;
    p->cyl_per_filament = 10;
    p->box = 1000.0;
    p->n_linkers = 200;
    p->n_motors = 20;
    p->n_branches = 0;
    p->cyl_size = 108.0;
    p->jitter = 2.0;
    p->min_sep = 8.0;
    p->pair_cut = 175.0;
    p->rest_sigma = 5.0;
    p->k_stretch = 100.0; p->k_bend = 672.0; p->theta_bend = 0.0;
    p->k_exvol = 1e5; p->exvol_cutoff = 108.0f;
    p->k_bd = 41.0; p->screen_bd = 2.7; p->bd_cutoff = 125.0f;
    p->k_linker = 8.0; p->k_motor = 2.5;
    p->k_br_stretch = 100.0; p->l_br_stretch = 6.0; p->k_br_bend = 10.0;
    p->theta_br_bend = 70.0 * M_PI / 180.0;
    p->k_br_dihedral = 0.0; p->k_br_position = 0.0;
}

extern "C" int mdy_generate(const mdy_gen_params *p, mdy_network **out) {
    if (!p || !out) return MDY_ERR_ARG;
    *out = nullptr;
    const int nfil = p->n_filaments, ncyl = p->cyl_per_filament, nbr = p->n_branches;
    if (nfil <= 0 || ncyl <= 0 || nbr < 0 || nbr >= nfil || p->box <= 100.0) return MDY_ERR_ARG;
    const int nmother = nfil - nbr;
    const double L = p->box, cs = p->cyl_size;
    const double lo = 25.0 + p->bd_cutoff / 4.0, hi = (L - 25.0) - p->bd_cutoff / 4.0;
    if (hi - lo <= 0) return MDY_ERR_ARG;
    Rng rng(p->seed);
    Storage *st = new (std::nothrow) Storage();
    if (!st) return MDY_ERR_ARG;

    std::vector<V3> beads;                      // ideal bead positions
    beads.reserve((size_t)nfil * (ncyl + 1));
    std::vector<V3> fdir(nfil);
    SegGrid grid(cs + p->min_sep + 1.0);
    auto inside = [&](V3 v) { return v.x >= lo && v.x <= hi && v.y >= lo && v.y <= hi && v.z >= lo && v.z <= hi; };
    const long max_tries = 2000000;

    // mothers: uniform start and direction
    for (int f = 0; f < nmother; f++) {
        long tries = 0;
        for (;;) {
            if (++tries > max_tries) { delete st; return MDY_ERR_ARG; }
            V3 p0 = {rng.uniform() * L, rng.uniform() * L, rng.uniform() * L};
            double u[3];
            rng.unit_vector(u);
            V3 d = {u[0], u[1], u[2]};
            V3 end = p0 + (ncyl * cs) * d;
            if (!inside(p0) || !inside(end)) continue;
            bool ok = true;
            if (p->min_sep > 0)
                for (int c = 0; c < ncyl && ok; c++)
                    ok = grid.min_dist(p0 + (c * cs) * d, p0 + ((c + 1) * cs) * d, -1) >= p->min_sep;
            if (!ok) continue;
            st->fil_first.push_back((int32_t)beads.size());
            st->fil_ncyl.push_back(ncyl);
            for (int b = 0; b <= ncyl; b++) beads.push_back(p0 + (b * cs) * d);
            for (int c = 0; c < ncyl; c++) grid.add(p0 + (c * cs) * d, p0 + ((c + 1) * cs) * d, f);
            fdir[f] = d;
            break;
        }
    }
    // daughters: from a point on a mother cylinder, at the branching angle
    std::vector<int> mother_cyl(nbr), branch_of_cyl((size_t)nmother * ncyl, -1);
    std::vector<double> branch_pos(nbr);
    const double ct = std::cos(p->theta_br_bend), stt = std::sin(p->theta_br_bend);
    for (int q = 0; q < nbr; q++) {
        long tries = 0;
        for (;;) {
            if (++tries > max_tries) { delete st; return MDY_ERR_ARG; }
            int mc = rng.below(nmother * ncyl);
            if (branch_of_cyl[mc] >= 0) continue;
            int mf = mc / ncyl, mpos = mc % ncyl;
            double pos = kSites[rng.below(4)];
            V3 b1 = beads[st->fil_first[mf] + mpos], b2 = beads[st->fil_first[mf] + mpos + 1];
            V3 v = (1.0 - pos) * b1 + pos * b2;
            V3 m = fdir[mf];
            V3 helper = std::fabs(m.x) < 0.9 ? V3{1, 0, 0} : V3{0, 1, 0};
            V3 u = crossv(m, helper);
            u = (1.0 / norm(u)) * u;
            V3 w = crossv(m, u);
            double phi = 2.0 * M_PI * rng.uniform();
            V3 d = ct * m + stt * (std::cos(phi) * u + std::sin(phi) * w);
            V3 p0 = v + p->l_br_stretch * d;
            V3 end = p0 + (ncyl * cs) * d;
            if (!inside(p0) || !inside(end)) continue;
            bool ok = true;
            if (p->min_sep > 0)
                for (int c = 0; c < ncyl && ok; c++)
                    ok = grid.min_dist(p0 + (c * cs) * d, p0 + ((c + 1) * cs) * d, mf) >= p->min_sep;
            if (!ok) continue;
            int f = nmother + q;
            st->fil_first.push_back((int32_t)beads.size());
            st->fil_ncyl.push_back(ncyl);
            for (int b = 0; b <= ncyl; b++) beads.push_back(p0 + (b * cs) * d);
            for (int c = 0; c < ncyl; c++) grid.add(p0 + (c * cs) * d, p0 + ((c + 1) * cs) * d, f);
            fdir[f] = d;
            mother_cyl[q] = mc;
            branch_pos[q] = pos;
            branch_of_cyl[mc] = q;
            break;
        }
    }
    // jitter
    const int nbeads = (int)beads.size();
    st->coord.resize((size_t)3 * nbeads);
    for (int b = 0; b < nbeads; b++) {
        st->coord[3 * b] = beads[b].x + p->jitter * rng.normal();
        st->coord[3 * b + 1] = beads[b].y + p->jitter * rng.normal();
        st->coord[3 * b + 2] = beads[b].z + p->jitter * rng.normal();
    }
    // cylinders (filament-major; daughters last, so their ids are larger)
    const int ncyl_tot = nfil * ncyl;
    st->cb1.resize(ncyl_tot); st->cb2.resize(ncyl_tot); st->cfil.resize(ncyl_tot); st->cpos.resize(ncyl_tot);
    st->cchild.assign(ncyl_tot, -1); st->cid.resize(ncyl_tot); st->ctype.assign(ncyl_tot, 0);
    st->ceq.assign(ncyl_tot, cs); st->cek.assign(ncyl_tot, p->k_exvol); st->cminus.resize(ncyl_tot);
    for (int f = 0; f < nfil; f++)
        for (int c = 0; c < ncyl; c++) {
            int i = f * ncyl + c;
            st->cb1[i] = 3 * (st->fil_first[f] + c);
            st->cb2[i] = 3 * (st->fil_first[f] + c + 1);
            st->cfil[i] = f; st->cpos[i] = c; st->cid[i] = i;
            st->cminus[i] = (c == 0) ? 1 : 0;
        }
    for (int q = 0; q < nbr; q++) st->cchild[mother_cyl[q]] = (nmother + q) * ncyl;
    // FilamentStretching: one interaction per cylinder, k scaled by cylinderSize/eqLength (MCylinder.cpp:35-43)
    for (int i = 0; i < ncyl_tot; i++) {
        st->fs_bs.push_back(st->cb1[i]); st->fs_bs.push_back(st->cb2[i]);
        st->fs_k.push_back(p->k_stretch * (cs / st->ceq[i]));
        st->fs_l0.push_back(st->ceq[i]);
    }
    // FilamentBending: consecutive cylinder pairs (FilamentBending.cpp:40-56)
    for (int f = 0; f < nfil; f++)
        for (int c = 1; c < ncyl; c++) {
            int prev = f * ncyl + c - 1, cur = f * ncyl + c;
            st->fb_bs.push_back(st->cb1[prev]); st->fb_bs.push_back(st->cb1[cur]); st->fb_bs.push_back(st->cb2[cur]);
            st->fb_k.push_back(p->k_bend * (cs / st->ceq[cur]));
            st->fb_t0.push_back(p->theta_bend);
        }
    // linkers and motors
    std::vector<double> mid((size_t)3 * ncyl_tot);
    for (int i = 0; i < ncyl_tot; i++)
        for (int k = 0; k < 3; k++) mid[3 * i + k] = 0.5 * (st->coord[st->cb1[i] + k] + st->coord[st->cb2[i] + k]);
    const double mcell = p->pair_cut + cs;
    std::unordered_map<int64_t, std::vector<int>> mgrid;
    auto mkey = [&](int i, int64_t *c) {
        for (int k = 0; k < 3; k++) c[k] = (int64_t)std::floor(mid[3 * i + k] / mcell);
    };
    for (int i = 0; i < ncyl_tot; i++) {
        int64_t c[3];
        mkey(i, c);
        mgrid[SegGrid::key(c[0], c[1], c[2])].push_back(i);
    }
    // the cylinders of the 27 cells around each occupied cell, in scan order,
    // built once per cell (the candidates of a first cylinder are this list
    // without its own filament's cylinders)
    std::unordered_map<int64_t, int> uid;
    std::vector<int> u_off(1, 0), u_list, cyl_u(ncyl_tot);
    for (int i = 0; i < ncyl_tot; i++) {
        int64_t c[3];
        mkey(i, c);
        const int64_t exact = ((c[0] + (1 << 20)) << 42) | ((c[1] + (1 << 20)) << 21) | (c[2] + (1 << 20));
        auto found = uid.find(exact);
        if (found != uid.end()) { cyl_u[i] = found->second; continue; }
        for (int64_t a = -1; a <= 1; a++)
            for (int64_t b = -1; b <= 1; b++)
                for (int64_t d = -1; d <= 1; d++) {
                    auto it = mgrid.find(SegGrid::key(c[0] + a, c[1] + b, c[2] + d));
                    if (it != mgrid.end()) u_list.insert(u_list.end(), it->second.begin(), it->second.end());
                }
        u_off.push_back((int)u_list.size());
        cyl_u[i] = uid[exact] = (int)u_off.size() - 2;
    }
    auto make_links = [&](int count, double kval, std::vector<int32_t> &bs, std::vector<double> &kk,
                          std::vector<double> &l0, std::vector<double> &p1v, std::vector<double> &p2v) -> bool {
        for (int l = 0; l < count; l++) {
            long tries = 0;
            for (;;) {
                if (++tries > max_tries) return false;
                int c1 = rng.below(ncyl_tot);
                const int *U = u_list.data() + u_off[cyl_u[c1]];
                const int nU = u_off[cyl_u[c1] + 1] - u_off[cyl_u[c1]];
                const int f1 = st->cfil[c1];
                int nvalid = 0;
                for (int q = 0; q < nU; q++) nvalid += st->cfil[U[q]] != f1;
                if (nvalid == 0) continue;
                int r = rng.below(nvalid), c2 = -1;
                for (int q = 0; q < nU; q++)
                    if (st->cfil[U[q]] != f1 && r-- == 0) { c2 = U[q]; break; }
                double q1 = kSites[rng.below(4)], q2 = kSites[rng.below(4)];
                double v1[3], v2[3];
                for (int k = 0; k < 3; k++) {
                    v1[k] = st->coord[st->cb1[c1] + k] * (1 - q1) + q1 * st->coord[st->cb2[c1] + k];
                    v2[k] = st->coord[st->cb1[c2] + k] * (1 - q2) + q2 * st->coord[st->cb2[c2] + k];
                }
                double dd = std::sqrt((v2[0] - v1[0]) * (v2[0] - v1[0]) + (v2[1] - v1[1]) * (v2[1] - v1[1]) +
                                      (v2[2] - v1[2]) * (v2[2] - v1[2]));
                if (!(dd < p->pair_cut)) continue;
                double rest;
                do { rest = dd + p->rest_sigma * rng.normal(); } while (rest <= 0.0);
                bs.push_back(st->cb1[c1]); bs.push_back(st->cb2[c1]); bs.push_back(st->cb1[c2]); bs.push_back(st->cb2[c2]);
                kk.push_back(kval); l0.push_back(rest); p1v.push_back(q1); p2v.push_back(q2);
                break;
            }
        }
        return true;
    };
    if (!make_links(p->n_linkers, p->k_linker, st->ls_bs, st->ls_k, st->ls_l0, st->ls_p1, st->ls_p2) ||
        !make_links(p->n_motors, p->k_motor, st->ms_bs, st->ms_k, st->ms_l0, st->ms_p1, st->ms_p2)) {
        delete st;
        return MDY_ERR_ARG;
    }
    // branching stretching / bending (BranchingStretching.cpp:38-46, BranchingBending.cpp:37-46)
    for (int q = 0; q < nbr; q++) {
        int mc = mother_cyl[q], dc = (nmother + q) * ncyl;
        st->bs_bs.push_back(st->cb1[mc]); st->bs_bs.push_back(st->cb2[mc]); st->bs_bs.push_back(st->cb1[dc]);
        st->bs_k.push_back(p->k_br_stretch); st->bs_l0.push_back(p->l_br_stretch); st->bs_p.push_back(branch_pos[q]);
        st->bb_bs.push_back(st->cb1[mc]); st->bb_bs.push_back(st->cb2[mc]);
        st->bb_bs.push_back(st->cb1[dc]); st->bb_bs.push_back(st->cb2[dc]);
        st->bb_k.push_back(p->k_br_bend); st->bb_t0.push_back(p->theta_br_bend);
        // branching dihedral / position (BranchingDihedral.cpp:41-50, BranchingPosition.cpp:37-45)
        if (p->k_br_dihedral != 0.0) {
            st->dh_bs.push_back(st->cb1[mc]); st->dh_bs.push_back(st->cb2[mc]);
            st->dh_bs.push_back(st->cb1[dc]); st->dh_bs.push_back(st->cb2[dc]);
            st->dh_k.push_back(p->k_br_dihedral); st->dh_p.push_back(branch_pos[q]);
        }
        if (p->k_br_position != 0.0) {
            st->bp_bs.push_back(st->cb1[mc]); st->bp_bs.push_back(st->cb2[mc]); st->bp_bs.push_back(st->cb1[dc]);
            st->bp_k.push_back(p->k_br_position); st->bp_p.push_back(branch_pos[q]);
        }
    }
    // boundary planes, BoundaryCubic (BoundaryImpl.cpp:25-52) and the plane
    // equation of PlaneBoundaryElement (BoundaryElementImpl.h:44-49)
    {
        const double z0 = 25, sys = L - 25;
        const double pts[6][3] = {{z0, sys / 2, sys / 2}, {sys, sys / 2, sys / 2}, {sys / 2, z0, sys / 2},
                                  {sys / 2, sys, sys / 2}, {sys / 2, sys / 2, z0}, {sys / 2, sys / 2, sys}};
        const double nrm[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
        for (int i = 0; i < 6; i++) {
            double a = nrm[i][0], b = nrm[i][1], c = nrm[i][2];
            double d = -a * pts[i][0] - b * pts[i][1] - c * pts[i][2];
            st->plane.push_back(a); st->plane.push_back(b); st->plane.push_back(c); st->plane.push_back(d);
            st->pk.push_back(p->k_bd); st->ps.push_back(p->screen_bd);
        }
    }

    mdy_network *net = new (std::nothrow) mdy_network();
    if (!net) { delete st; return MDY_ERR_ARG; }
    std::memset(net, 0, sizeof(*net));
    net->storage = st;
    net->coord = st->coord.data();
    net->box = L;
    net->n_filaments = nfil;
    net->fil_first_bead = st->fil_first.data();
    net->fil_ncyl = st->fil_ncyl.data();
    mdy_system &s = net->sys;
    s.n_beads = nbeads;
    s.cyl.n = ncyl_tot;
    s.cyl.bead1 = st->cb1.data(); s.cyl.bead2 = st->cb2.data(); s.cyl.id = st->cid.data();
    s.cyl.filament = st->cfil.data(); s.cyl.position = st->cpos.data(); s.cyl.type = st->ctype.data();
    s.cyl.eq_length = st->ceq.data(); s.cyl.exvol_k = st->cek.data(); s.cyl.branch_child = st->cchild.data();
    s.cyl.minus_end = st->cminus.data();
    mdy_ff_tables &t = s.ff;
    t.fs_n = (int32_t)st->fs_k.size(); t.fs_beadset = st->fs_bs.data(); t.fs_k = st->fs_k.data(); t.fs_l0 = st->fs_l0.data();
    t.fb_n = (int32_t)st->fb_k.size(); t.fb_beadset = st->fb_bs.data(); t.fb_k = st->fb_k.data(); t.fb_theta0 = st->fb_t0.data();
    t.ls_n = (int32_t)st->ls_k.size(); t.ls_beadset = st->ls_bs.data(); t.ls_k = st->ls_k.data(); t.ls_l0 = st->ls_l0.data();
    t.ls_pos1 = st->ls_p1.data(); t.ls_pos2 = st->ls_p2.data();
    t.ms_n = (int32_t)st->ms_k.size(); t.ms_beadset = st->ms_bs.data(); t.ms_k = st->ms_k.data(); t.ms_l0 = st->ms_l0.data();
    t.ms_pos1 = st->ms_p1.data(); t.ms_pos2 = st->ms_p2.data();
    t.bs_n = (int32_t)st->bs_k.size(); t.bs_beadset = st->bs_bs.data(); t.bs_k = st->bs_k.data(); t.bs_l0 = st->bs_l0.data();
    t.bs_pos = st->bs_p.data();
    t.bb_n = (int32_t)st->bb_k.size(); t.bb_beadset = st->bb_bs.data(); t.bb_k = st->bb_k.data(); t.bb_theta0 = st->bb_t0.data();
    t.bdh_n = (int32_t)st->dh_k.size(); t.bdh_beadset = st->dh_bs.data(); t.bdh_k = st->dh_k.data();
    t.bdh_pos = st->dh_p.data();
    t.bp_n = (int32_t)st->bp_k.size(); t.bp_beadset = st->bp_bs.data(); t.bp_k = st->bp_k.data();
    t.bp_pos = st->bp_p.data();
    s.bnd.n_planes = 6; s.bnd.plane = st->plane.data(); s.bnd.k = st->pk.data(); s.bnd.screen = st->ps.data();
    s.bnd.cutoff = p->bd_cutoff;
    s.exvol_cutoff = p->exvol_cutoff;
    s.exvol_enabled = 1;
    s.boundary_enabled = 1;
    *out = net;
    return MDY_OK;
}

extern "C" void mdy_network_free(mdy_network *net) {
    if (!net) return;
    delete static_cast<Storage *>(net->storage);
    delete net;
}
