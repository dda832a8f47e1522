/*
 * mdy_traj.h -- C ABI of libmdy_traj.so: MEDYAN's snapshot trajectory file
 * (traj.h5) written with the HDF5 C library, in the layout the reference's
 * SnapshotOutput writes (src/Output.h:504-575, src/Structure/Output/O*.hpp):
 *
 *   /header               count (int64), finished (bool), attribute "version"
 *   /header/meta          vertexColumnNamesFloat64, vertexColumnNamesInt64,
 *                         globalFilamentModel, system, chem, energyNames,
 *                         diffusingSpeciesNames, globalSpeciesNames
 *   /snapshots/<k>        time, energies [total, per force field...]
 *   /snapshots/<k>/filaments   count, <i> = bead coordinates (numBeads x 3)
 *   /snapshots/<k>/linkers     count, <i>/{id, type, subtype, coords (2 x 3)}
 *   /snapshots/<k>/bubbles     count (0)
 *   /snapshots/<k>/membranes   count (0)
 *   /snapshots/<k>             diffusingSpeciesCopyNumbers (0 species),
 *                              globalSpeciesCopyNumbers (empty)
 *
 * Types: int64 little-endian, IEEE double, variable-length UTF-8 strings,
 * bool as an int8 enum {FALSE = 0, TRUE = 1}. Matrices are stored
 * transposed from the reference's column-major Eigen storage, so a
 * filament's dataset is (numBeads, 3) with bead b's x, y, z contiguous
 * (src/Util/Io/H5.hpp:106-118).
 *
 * Each call opens and closes the file (as SnapshotOutput::append does), so a
 * file is complete after every call. Calls come from one thread at a time
 * (the HDF5 library may be built without thread safety). Returns MDY_OK (0) or MDY_ERR_ARG (-1,
 * bad argument), MDY_ERR_CAPACITY (-5, readback buffer too small) or
 * MDY_ERR_IO (the HDF5 library reported an error, or the file's contents
 * are not what this layout expects).
 */
#ifndef MDY_TRAJ_H
#define MDY_TRAJ_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { MDY_ERR_IO = -12 };

/* one snapshot: OutputStructSnapshot (src/Structure/OutputStruct.hpp:157-186) */
typedef struct mdy_traj_snapshot {
    double time;                    /* SubSystem::tau() */
    int32_t n_filaments;
    const int32_t *fil_n_beads;     /* numBeads per filament */
    const double *fil_coord;        /* 3 * sum(numBeads): beads in filament order */
    int32_t n_linkers;              /* linkers, then motors, then branching points */
    const int64_t *lk_id;
    const int32_t *lk_kind;         /* MDY_TRAJ_LINKER / _MOTOR / _BRANCHER -> "linker" / "motor" / "brancher" */
    const int64_t *lk_subtype;
    const double *lk_coord;         /* 6 per linker: the two binding points */
    int32_t n_energies;             /* energiesAfter: total, then one per force field */
    const double *energies;
} mdy_traj_snapshot;
enum { MDY_TRAJ_LINKER = 0, MDY_TRAJ_MOTOR = 1, MDY_TRAJ_BRANCHER = 2 };

/* SnapshotOutput::init + createSnapshotHeader (src/Output.h:504-538): creates
 * (truncates) `path`; count 0, finished false; energy_names = "total"
 * followed by the force fields' names; system_input / chem_input are the
 * simulation configuration as text (may be empty). */
int mdy_traj_create(const char *path, const char *version, int32_t n_energy_names, const char *const *energy_names,
                    const char *system_input, const char *chem_input);

/* SnapshotOutput::append (src/Output.h:540-563): writes snapshot <count> and
 * increments the header's count. */
int mdy_traj_append(const char *path, const mdy_traj_snapshot *snap);

/* SnapshotOutput::finish (src/Output.h:565-569): finished = true. */
int mdy_traj_finish(const char *path);

/* Readback for tests and tools: the number of snapshots, and snapshot k's
 * filament coordinates (3 * total beads, in filament order) into `coord`
 * (capacity `cap` doubles). Returns the number of doubles, or an error. */
int64_t mdy_traj_count(const char *path);
int64_t mdy_traj_read_filaments(const char *path, int64_t k, double *coord, int64_t cap);

const char *mdy_traj_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* MDY_TRAJ_H */
