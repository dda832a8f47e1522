"""traj.h5 snapshot output (SnapshotOutput, src/Output.h:527-571) through
libmdy_traj.so, which writes MEDYAN's snapshot layout with the HDF5 C
library (layout in include/mdy_traj.h).

    out = SnapshotOutput("traj.h5", energy_names=["total", *FF_NAMES])
    out.append(snapshot_of(net, x, energies_after, time=0.0))
    out.finish()
"""
import ctypes as C
import os

import numpy as np

from . import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libmdy_traj.so")
MDY_ERR_IO = -12
MDY_TRAJ_LINKER, MDY_TRAJ_MOTOR, MDY_TRAJ_BRANCHER = 0, 1, 2

_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)


class Snapshot(C.Structure):
    _fields_ = [("time", C.c_double),
                ("n_filaments", C.c_int32), ("fil_n_beads", _i32p), ("fil_coord", _f64p),
                ("n_linkers", C.c_int32), ("lk_id", _i64p), ("lk_kind", _i32p), ("lk_subtype", _i64p),
                ("lk_coord", _f64p),
                ("n_energies", C.c_int32), ("energies", _f64p)]


_lib = None


def load_library():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} is not built (needs the HDF5 C library; see medyan_amd/build.py)")
        L = C.CDLL(LIB)
        L.mdy_traj_create.argtypes = [C.c_char_p, C.c_char_p, C.c_int32, C.POINTER(C.c_char_p), C.c_char_p,
                                      C.c_char_p]
        L.mdy_traj_append.argtypes = [C.c_char_p, C.POINTER(Snapshot)]
        L.mdy_traj_finish.argtypes = [C.c_char_p]
        L.mdy_traj_count.argtypes = [C.c_char_p]
        L.mdy_traj_count.restype = C.c_int64
        L.mdy_traj_read_filaments.argtypes = [C.c_char_p, C.c_int64, _f64p, C.c_int64]
        L.mdy_traj_read_filaments.restype = C.c_int64
        L.mdy_traj_last_error.restype = C.c_char_p
        _lib = L
    return _lib


def _check(rc):
    if rc < 0:
        msg = load_library().mdy_traj_last_error().decode()
        raise RuntimeError(f"{'MDY_ERR_IO' if rc == MDY_ERR_IO else _abi.ERRORS.get(rc, rc)}: {msg}")
    return rc


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


def filament_beads(net):
    """Bead indices of each filament, minus end first, as
    OutputStructFilament stores them (each cylinder's first bead, then the
    last cylinder's second bead; src/Structure/Output/OFilament.hpp:22-40)."""
    out = []
    for f in np.unique(net.cyl_filament):
        cs = np.nonzero(net.cyl_filament == f)[0]
        cs = cs[np.argsort(net.cyl_position[cs], kind="stable")]
        out.append(np.concatenate([net.cyl_bead1[cs] // 3, [net.cyl_bead2[cs[-1]] // 3]]).astype(np.int64))
    return out


def snapshot_of(net, coord, energies, time=0.0):
    """The snapshot of a network at coordinates `coord`: filament bead
    coordinates; linkers and motors with their two binding points
    (midpoints at pos1/pos2 of the bound cylinders, OLinker.hpp:10-54);
    branching points with the branch point on the mother cylinder and the
    daughter's first bead (OLinker.hpp:56-70); `energies` = total followed by
    one energy per force field (energiesAfter, OSnapshot.hpp:21-29)."""
    x = np.asarray(coord, np.float64).reshape(-1, 3)
    fils = filament_beads(net)
    lk_id, lk_kind, lk_coord = [], [], []

    def mid(b1, b2, p):
        return (1 - p) * x[b1 // 3] + p * x[b2 // 3]

    for kind, bs, p1, p2 in ((MDY_TRAJ_LINKER, net.ls_beadset, net.ls_pos1, net.ls_pos2),
                             (MDY_TRAJ_MOTOR, net.ms_beadset, net.ms_pos1, net.ms_pos2)):
        b = np.asarray(bs).reshape(-1, 4)
        for i in range(len(b)):
            lk_id.append(i)
            lk_kind.append(kind)
            lk_coord.append(np.concatenate([mid(b[i, 0], b[i, 1], p1[i]), mid(b[i, 2], b[i, 3], p2[i])]))
    b = np.asarray(net.bs_beadset).reshape(-1, 3)
    for i in range(len(b)):
        lk_id.append(i)
        lk_kind.append(MDY_TRAJ_BRANCHER)
        lk_coord.append(np.concatenate([mid(b[i, 0], b[i, 1], net.bs_pos[i]), x[b[i, 2] // 3]]))
    return dict(time=float(time), filaments=[x[f] for f in fils], lk_id=np.array(lk_id, np.int64),
                lk_kind=np.array(lk_kind, np.int32), lk_subtype=np.zeros(len(lk_id), np.int64),
                lk_coord=np.array(lk_coord, np.float64).reshape(-1, 6), energies=np.asarray(energies, np.float64))


class SnapshotOutput:
    """SnapshotOutput (src/Output.h:527-571): init creates the file and its
    header, append writes the next snapshot, finish marks the file finished."""

    def __init__(self, filename, energy_names=None, version="medyan_amd", system_input="", chem_input=""):
        self.lib = load_library()
        self.filename = os.fsencode(filename)
        names = list(energy_names if energy_names is not None else ["total", *_abi.FF_NAMES])
        arr = (C.c_char_p * max(len(names), 1))(*[n.encode() for n in names])
        _check(self.lib.mdy_traj_create(self.filename, version.encode(), len(names), arr, system_input.encode(),
                                        chem_input.encode()))

    def append(self, snap):
        fil_n = np.array([len(f) for f in snap["filaments"]], np.int32)
        fil_c = (np.ascontiguousarray(np.concatenate(snap["filaments"]), np.float64) if len(fil_n)
                 else np.zeros(0))
        lk_id = np.ascontiguousarray(snap["lk_id"], np.int64)
        lk_kind = np.ascontiguousarray(snap["lk_kind"], np.int32)
        lk_sub = np.ascontiguousarray(snap["lk_subtype"], np.int64)
        lk_c = np.ascontiguousarray(snap["lk_coord"], np.float64)
        en = np.ascontiguousarray(snap["energies"], np.float64)
        s = Snapshot(snap["time"], len(fil_n), _p(fil_n, C.c_int32), _p(fil_c), len(lk_id), _p(lk_id, C.c_int64),
                     _p(lk_kind, C.c_int32), _p(lk_sub, C.c_int64), _p(lk_c), len(en), _p(en))
        _check(self.lib.mdy_traj_append(self.filename, C.byref(s)))

    def finish(self):
        _check(self.lib.mdy_traj_finish(self.filename))

    def count(self):
        return _check(self.lib.mdy_traj_count(self.filename))

    def read_filaments(self, k, total_beads):
        out = np.zeros(3 * total_beads)
        n = _check(self.lib.mdy_traj_read_filaments(self.filename, k, _p(out), out.size))
        return out[:n]
