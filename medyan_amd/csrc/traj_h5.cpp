// traj_h5.cpp -- MEDYAN's snapshot trajectory (traj.h5) through the HDF5 C
// library; the layout and types are described in include/mdy_traj.h.
// Host-only code (g++), built into libmdy_traj.so when the HDF5 headers are
// available; the minimizer library does not depend on it.
#include "mdy_traj.h"

#include <hdf5.h>

#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}

// RAII handle for one HDF5 identifier and its close function
struct H {
    hid_t id = -1;
    herr_t (*close)(hid_t) = nullptr;
    H() = default;
    H(hid_t i, herr_t (*c)(hid_t)) : id(i), close(c) {}
    H(const H &) = delete;
    H &operator=(const H &) = delete;
    ~H() {
        if (id >= 0 && close) close(id);
    }
    bool ok() const { return id >= 0; }
    operator hid_t() const { return id; }
};

struct Silence {   // the library's own error stack printing, off for one call
    H5E_auto2_t fn = nullptr;
    void *data = nullptr;
    Silence() {
        H5Eget_auto2(H5E_DEFAULT, &fn, &data);
        H5Eset_auto2(H5E_DEFAULT, nullptr, nullptr);
    }
    ~Silence() { H5Eset_auto2(H5E_DEFAULT, fn, data); }
};

// thrown inside this file only, turned into its code at the boundary
struct Err {
    int code;
};
[[noreturn]] void raise(const std::string &msg, int code = MDY_ERR_IO) {
    g_err = msg;
    throw Err{code};
}

hid_t check(hid_t id, const char *what) {
    if (id < 0) raise(std::string("HDF5: ") + what);
    return id;
}
void check(herr_t rc, const char *what, int) {
    if (rc < 0) raise(std::string("HDF5: ") + what);
}

H vlen_string_type() {
    H t(check(H5Tcopy(H5T_C_S1), "string type"), H5Tclose);
    check(H5Tset_size(t, H5T_VARIABLE), "string size", 0);
    check(H5Tset_cset(t, H5T_CSET_UTF8), "string charset", 0);
    return H(std::exchange(t.id, -1), H5Tclose);
}

H bool_type() {   // enum {FALSE = 0, TRUE = 1} over int8
    H t(check(H5Tenum_create(H5T_NATIVE_INT8), "bool type"), H5Tclose);
    int8_t v = 0;
    check(H5Tenum_insert(t, "FALSE", &v), "bool FALSE", 0);
    v = 1;
    check(H5Tenum_insert(t, "TRUE", &v), "bool TRUE", 0);
    return H(std::exchange(t.id, -1), H5Tclose);
}

bool exists(hid_t loc, const char *name) { return H5Lexists(loc, name, H5P_DEFAULT) > 0; }

// create a dataset, or open it when it exists (the reference's writeDataSet
// reuses an existing dataset of the same name)
H dataset(hid_t loc, const char *name, hid_t type, hid_t space) {
    hid_t d = exists(loc, name) ? H5Dopen2(loc, name, H5P_DEFAULT)
                                : H5Dcreate2(loc, name, type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    return H(check(d, name), H5Dclose);
}

H group(hid_t loc, const char *name) {
    hid_t g = exists(loc, name) ? H5Gopen2(loc, name, H5P_DEFAULT)
                                : H5Gcreate2(loc, name, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    return H(check(g, name), H5Gclose);
}

void write_i64(hid_t loc, const char *name, int64_t v) {
    H sp(check(H5Screate(H5S_SCALAR), "scalar space"), H5Sclose);
    H d = dataset(loc, name, H5T_STD_I64LE, sp);
    check(H5Dwrite(d, H5T_NATIVE_INT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v), name, 0);
}

void write_f64(hid_t loc, const char *name, double v) {
    H sp(check(H5Screate(H5S_SCALAR), "scalar space"), H5Sclose);
    H d = dataset(loc, name, H5T_IEEE_F64LE, sp);
    check(H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v), name, 0);
}

void write_bool(hid_t loc, const char *name, bool b) {
    H t = bool_type();
    H sp(check(H5Screate(H5S_SCALAR), "scalar space"), H5Sclose);
    H d = dataset(loc, name, t, sp);
    const int8_t v = b ? 1 : 0;
    check(H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v), name, 0);
}

void write_str(hid_t loc, const char *name, const char *s) {
    H t = vlen_string_type();
    H sp(check(H5Screate(H5S_SCALAR), "scalar space"), H5Sclose);
    H d = dataset(loc, name, t, sp);
    const char *p = s ? s : "";
    check(H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, &p), name, 0);
}

void write_strs(hid_t loc, const char *name, const std::vector<const char *> &v) {
    H t = vlen_string_type();
    const hsize_t n = v.size();
    H sp(check(H5Screate_simple(1, &n, nullptr), "string list space"), H5Sclose);
    H d = dataset(loc, name, t, sp);
    if (n) check(H5Dwrite(d, t, H5S_ALL, H5S_ALL, H5P_DEFAULT, v.data()), name, 0);
}

// rank-1 or rank-2 double array, row-major dims
void write_f64s(hid_t loc, const char *name, int rank, const hsize_t *dims, const double *p) {
    H sp(check(H5Screate_simple(rank, dims, nullptr), "array space"), H5Sclose);
    H d = dataset(loc, name, H5T_IEEE_F64LE, sp);
    hsize_t n = 1;
    for (int i = 0; i < rank; i++) n *= dims[i];
    if (n) check(H5Dwrite(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, p), name, 0);
}

void write_u64s_empty(hid_t loc, const char *name, int rank, const hsize_t *dims) {
    H sp(check(H5Screate_simple(rank, dims, nullptr), "array space"), H5Sclose);
    H d = dataset(loc, name, H5T_STD_U64LE, sp);
}

H open_rw(const char *path) {
    if (!path) raise("path is null", -1);
    return H(check(H5Fopen(path, H5F_ACC_RDWR, H5P_DEFAULT), "cannot open file"), H5Fclose);
}

int64_t read_i64(hid_t loc, const char *name) {
    H d(check(H5Dopen2(loc, name, H5P_DEFAULT), name), H5Dclose);
    int64_t v = 0;
    check(H5Dread(d, H5T_NATIVE_INT64, H5S_ALL, H5S_ALL, H5P_DEFAULT, &v), name, 0);
    return v;
}

template <class F> int guarded(F &&f) {
    Silence quiet;
    try {
        return f();
    } catch (const Err &e) {
        return e.code;
    } catch (...) {
        g_err = "unexpected exception";
        return MDY_ERR_IO;
    }
}

}   // namespace

extern "C" {

const char *mdy_traj_last_error(void) { return g_err.c_str(); }

int mdy_traj_create(const char *path, const char *version, int32_t n_energy_names, const char *const *energy_names,
                    const char *system_input, const char *chem_input) {
    if (!path || n_energy_names < 0 || (n_energy_names > 0 && !energy_names)) return fail(-1, "invalid argument");
    return guarded([&] {
        H f(check(H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT), "cannot create file"), H5Fclose);
        H hdr = group(f, "header");
        write_i64(hdr, "count", 0);
        {   // attribute "version"
            H t = vlen_string_type();
            H sp(check(H5Screate(H5S_SCALAR), "scalar space"), H5Sclose);
            H a(check(H5Acreate2(hdr, "version", t, sp, H5P_DEFAULT, H5P_DEFAULT), "version attribute"), H5Aclose);
            const char *v = version ? version : "";
            check(H5Awrite(a, t, &v), "version", 0);
        }
        write_bool(hdr, "finished", false);
        H meta = group(hdr, "meta");
        write_strs(meta, "vertexColumnNamesFloat64", {});
        write_strs(meta, "vertexColumnNamesInt64", {});
        write_str(meta, "globalFilamentModel", "bead-cylinder");
        write_str(meta, "system", system_input);
        write_str(meta, "chem", chem_input);
        std::vector<const char *> names;
        for (int32_t i = 0; i < n_energy_names; i++) {
            if (!energy_names[i]) raise("energy name is null", -1);
            names.push_back(energy_names[i]);
        }
        write_strs(meta, "energyNames", names);
        write_strs(meta, "diffusingSpeciesNames", {});
        write_strs(meta, "globalSpeciesNames", {});
        return 0;
    });
}

int mdy_traj_append(const char *path, const mdy_traj_snapshot *s) {
    if (!s || s->n_filaments < 0 || s->n_linkers < 0 || s->n_energies < 0) return fail(-1, "invalid argument");
    if ((s->n_filaments > 0 && (!s->fil_n_beads || !s->fil_coord)) ||
        (s->n_linkers > 0 && (!s->lk_id || !s->lk_kind || !s->lk_subtype || !s->lk_coord)) ||
        (s->n_energies > 0 && !s->energies))
        return fail(-1, "missing snapshot array");
    for (int32_t i = 0; i < s->n_filaments; i++)
        if (s->fil_n_beads[i] < 2) return fail(-1, "a filament has fewer than 2 beads");
    for (int32_t i = 0; i < s->n_linkers; i++)
        if (s->lk_kind[i] < MDY_TRAJ_LINKER || s->lk_kind[i] > MDY_TRAJ_BRANCHER) return fail(-1, "linker kind");
    return guarded([&] {
        H f = open_rw(path);
        H hdr(check(H5Gopen2(f, "header", H5P_DEFAULT), "header"), H5Gclose);
        const int64_t count = read_i64(hdr, "count");
        H snaps = group(f, "snapshots");
        const std::string key = std::to_string(count);
        if (exists(snaps, key.c_str())) raise("snapshot " + key + " exists");
        H snap = group(snaps, key.c_str());
        write_f64(snap, "time", s->time);
        {
            H fils = group(snap, "filaments");
            write_i64(fils, "count", s->n_filaments);
            const double *c = s->fil_coord;
            for (int32_t i = 0; i < s->n_filaments; i++) {
                const hsize_t dims[2] = {(hsize_t)s->fil_n_beads[i], 3};
                write_f64s(fils, std::to_string(i).c_str(), 2, dims, c);
                c += 3 * (size_t)s->fil_n_beads[i];
            }
        }
        {
            static const char *kind[3] = {"linker", "motor", "brancher"};
            H lks = group(snap, "linkers");
            write_i64(lks, "count", s->n_linkers);
            for (int32_t i = 0; i < s->n_linkers; i++) {
                H lk = group(lks, std::to_string(i).c_str());
                write_i64(lk, "id", s->lk_id[i]);
                write_str(lk, "type", kind[s->lk_kind[i]]);
                write_i64(lk, "subtype", s->lk_subtype[i]);
                const hsize_t dims[2] = {2, 3};
                write_f64s(lk, "coords", 2, dims, s->lk_coord + 6 * (size_t)i);
            }
        }
        {
            H b = group(snap, "bubbles");
            write_i64(b, "count", 0);
            H m = group(snap, "membranes");
            write_i64(m, "count", 0);
        }
        {   // no chemistry on this path: zero diffusing species, no global species
            const hsize_t d4[4] = {1, 1, 1, 0};
            write_u64s_empty(snap, "diffusingSpeciesCopyNumbers", 4, d4);
            const hsize_t d1[1] = {0};
            write_u64s_empty(snap, "globalSpeciesCopyNumbers", 1, d1);
        }
        const hsize_t ne = (hsize_t)s->n_energies;
        write_f64s(snap, "energies", 1, &ne, s->energies);
        write_i64(hdr, "count", count + 1);
        return 0;
    });
}

int mdy_traj_finish(const char *path) {
    return guarded([&] {
        H f = open_rw(path);
        H hdr(check(H5Gopen2(f, "header", H5P_DEFAULT), "header"), H5Gclose);
        write_bool(hdr, "finished", true);
        return 0;
    });
}

int64_t mdy_traj_count(const char *path) {
    if (!path) return fail(-1, "path is null");
    int64_t out = 0;
    int rc = guarded([&] {
        H f(check(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), "cannot open file"), H5Fclose);
        H hdr(check(H5Gopen2(f, "header", H5P_DEFAULT), "header"), H5Gclose);
        out = read_i64(hdr, "count");
        return 0;
    });
    return rc ? rc : out;
}

int64_t mdy_traj_read_filaments(const char *path, int64_t k, double *coord, int64_t cap) {
    if (!path || k < 0 || cap < 0 || (cap > 0 && !coord)) return fail(-1, "invalid argument");
    int64_t total = 0;
    int rc = guarded([&] {
        H f(check(H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT), "cannot open file"), H5Fclose);
        const std::string g = "snapshots/" + std::to_string(k) + "/filaments";
        H fils(check(H5Gopen2(f, g.c_str(), H5P_DEFAULT), "no such snapshot"), H5Gclose);
        const int64_t n = read_i64(fils, "count");
        for (int64_t i = 0; i < n; i++) {
            H d(check(H5Dopen2(fils, std::to_string(i).c_str(), H5P_DEFAULT), "filament"), H5Dclose);
            H sp(check(H5Dget_space(d), "filament space"), H5Sclose);
            hsize_t dims[2] = {0, 0};
            if (H5Sget_simple_extent_ndims(sp) != 2) raise("filament rank");
            H5Sget_simple_extent_dims(sp, dims, nullptr);
            const int64_t m = (int64_t)(dims[0] * dims[1]);
            if (total + m > cap) raise("capacity", -5);   // MDY_ERR_CAPACITY
            check(H5Dread(d, H5T_NATIVE_DOUBLE, H5S_ALL, H5S_ALL, H5P_DEFAULT, coord + total), "filament read", 0);
            total += m;
        }
        return 0;
    });
    return rc ? rc : total;
}

}   // extern "C"
