"""Networks as numpy arrays, and their conversion to the C ABI's mdy_system.

A `Network` owns numpy copies of every table of a system (coordinates,
cylinder database, vectorized bonded / cross-link tables, boundary planes), so
tests can perturb them and hand the same system to the HIP library and to the
CPU oracle.
"""
import ctypes as C
from dataclasses import dataclass, field, fields
from typing import Optional

import numpy as np

from . import _abi

# Benchmark workloads. Seeds are fixed per workload; geometry rules live in
# medyan_amd/csrc/netgen.cpp.
WORKLOADS = {
    # 50 filaments x 10 cylinders, 1 um box, 200 linkers, 20 motors
    "net500": dict(seed=0x4D454459 + 0, n_filaments=50, cyl_per_filament=10, box=1000.0,
                   n_linkers=200, n_motors=20, n_branches=0),
    # 4,000 filaments x 10 cylinders = 40k cylinders, 4 um box, 20k linkers, 2k motors, 2k branches
    "net40k": dict(seed=0x4D454459 + 1, n_filaments=4000, cyl_per_filament=10, box=4000.0,
                   n_linkers=20000, n_motors=2000, n_branches=2000),
    # 5,000 filaments x 10 cylinders, 5 um box
    "net50k": dict(seed=0x4D454459 + 3, n_filaments=5000, cyl_per_filament=10, box=5000.0,
                   n_linkers=0, n_motors=0, n_branches=0),
    # 25,000 filaments x 8 cylinders = 200k cylinders, 10 um box
    "net200k": dict(seed=0x4D454459 + 4, n_filaments=25000, cyl_per_filament=8, box=10000.0,
                    n_linkers=400000, n_motors=20000, n_branches=10000),
}


@dataclass
class Network:
    coord: np.ndarray
    n_beads: int
    # cylinders
    cyl_bead1: np.ndarray
    cyl_bead2: np.ndarray
    cyl_id: np.ndarray
    cyl_filament: np.ndarray
    cyl_position: np.ndarray
    cyl_type: np.ndarray
    cyl_eq_length: np.ndarray
    cyl_exvol_k: np.ndarray
    cyl_branch_child: np.ndarray
    cyl_minus_end: np.ndarray
    # tables
    fs_beadset: np.ndarray
    fs_k: np.ndarray
    fs_l0: np.ndarray
    fb_beadset: np.ndarray
    fb_k: np.ndarray
    fb_theta0: np.ndarray
    ls_beadset: np.ndarray
    ls_k: np.ndarray
    ls_l0: np.ndarray
    ls_pos1: np.ndarray
    ls_pos2: np.ndarray
    ms_beadset: np.ndarray
    ms_k: np.ndarray
    ms_l0: np.ndarray
    ms_pos1: np.ndarray
    ms_pos2: np.ndarray
    bs_beadset: np.ndarray
    bs_k: np.ndarray
    bs_l0: np.ndarray
    bs_pos: np.ndarray
    bb_beadset: np.ndarray
    bb_k: np.ndarray
    bb_theta0: np.ndarray
    plane: np.ndarray
    plane_k: np.ndarray
    plane_screen: np.ndarray
    bd_cutoff: float = 125.0
    exvol_cutoff: float = 108.0
    exvol_enabled: int = 1
    boundary_enabled: int = 1
    box: float = 0.0
    _keep: list = field(default_factory=list, repr=False)

    @property
    def n_cyl(self):
        return int(self.cyl_bead1.shape[0])

    def copy(self):
        kw = {}
        for f in fields(self):
            v = getattr(self, f.name)
            if f.name == "_keep":
                continue
            kw[f.name] = v.copy() if isinstance(v, np.ndarray) else v
        return Network(**kw)

    # -- C ABI view -------------------------------------------------------
    def system(self) -> _abi.System:
        """Build an mdy_system pointing into this object's arrays (kept alive
        as long as this Network lives)."""
        keep = []

        def ptr(a, ctype, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            keep.append(a)
            if a.size == 0:
                return C.cast(None, C.POINTER(ctype))
            return a.ctypes.data_as(C.POINTER(ctype))

        i32 = lambda a: ptr(a, C.c_int32, np.int32)
        f64 = lambda a: ptr(a, C.c_double, np.float64)
        s = _abi.System()
        s.n_beads = int(self.n_beads)
        c = s.cyl
        c.n = self.n_cyl
        c.bead1, c.bead2 = i32(self.cyl_bead1), i32(self.cyl_bead2)
        c.id = ptr(self.cyl_id, C.c_int64, np.int64)
        c.filament, c.position = i32(self.cyl_filament), i32(self.cyl_position)
        c.type = ptr(self.cyl_type, C.c_int16, np.int16)
        c.eq_length, c.exvol_k = f64(self.cyl_eq_length), f64(self.cyl_exvol_k)
        c.branch_child = i32(self.cyl_branch_child)
        c.minus_end = ptr(self.cyl_minus_end, C.c_uint8, np.uint8)
        t = s.ff
        t.fs_n = len(self.fs_k); t.fs_beadset = i32(self.fs_beadset); t.fs_k = f64(self.fs_k); t.fs_l0 = f64(self.fs_l0)
        t.fb_n = len(self.fb_k); t.fb_beadset = i32(self.fb_beadset); t.fb_k = f64(self.fb_k)
        t.fb_theta0 = f64(self.fb_theta0)
        t.ls_n = len(self.ls_k); t.ls_beadset = i32(self.ls_beadset); t.ls_k = f64(self.ls_k)
        t.ls_l0 = f64(self.ls_l0); t.ls_pos1 = f64(self.ls_pos1); t.ls_pos2 = f64(self.ls_pos2)
        t.ms_n = len(self.ms_k); t.ms_beadset = i32(self.ms_beadset); t.ms_k = f64(self.ms_k)
        t.ms_l0 = f64(self.ms_l0); t.ms_pos1 = f64(self.ms_pos1); t.ms_pos2 = f64(self.ms_pos2)
        t.bs_n = len(self.bs_k); t.bs_beadset = i32(self.bs_beadset); t.bs_k = f64(self.bs_k)
        t.bs_l0 = f64(self.bs_l0); t.bs_pos = f64(self.bs_pos)
        t.bb_n = len(self.bb_k); t.bb_beadset = i32(self.bb_beadset); t.bb_k = f64(self.bb_k)
        t.bb_theta0 = f64(self.bb_theta0)
        b = s.bnd
        b.n_planes = len(self.plane_k)
        b.plane, b.k, b.screen = f64(self.plane), f64(self.plane_k), f64(self.plane_screen)
        b.cutoff = float(self.bd_cutoff)
        s.exvol_cutoff = float(self.exvol_cutoff)
        s.exvol_enabled = int(self.exvol_enabled)
        s.boundary_enabled = int(self.boundary_enabled)
        self._keep = keep
        return s


def _arr(p, n, dtype):
    if n == 0 or not p:
        return np.zeros(0, dtype=dtype)
    return np.ctypeslib.as_array(p, shape=(n,)).astype(dtype, copy=True)


def from_c(net: _abi.Network) -> Network:
    s = net.sys
    nc, t = s.cyl.n, s.ff
    return Network(
        coord=_arr(net.coord, 3 * s.n_beads, np.float64), n_beads=s.n_beads,
        cyl_bead1=_arr(s.cyl.bead1, nc, np.int32), cyl_bead2=_arr(s.cyl.bead2, nc, np.int32),
        cyl_id=_arr(s.cyl.id, nc, np.int64), cyl_filament=_arr(s.cyl.filament, nc, np.int32),
        cyl_position=_arr(s.cyl.position, nc, np.int32), cyl_type=_arr(s.cyl.type, nc, np.int16),
        cyl_eq_length=_arr(s.cyl.eq_length, nc, np.float64), cyl_exvol_k=_arr(s.cyl.exvol_k, nc, np.float64),
        cyl_branch_child=_arr(s.cyl.branch_child, nc, np.int32),
        cyl_minus_end=_arr(s.cyl.minus_end, nc, np.uint8),
        fs_beadset=_arr(t.fs_beadset, 2 * t.fs_n, np.int32), fs_k=_arr(t.fs_k, t.fs_n, np.float64),
        fs_l0=_arr(t.fs_l0, t.fs_n, np.float64),
        fb_beadset=_arr(t.fb_beadset, 3 * t.fb_n, np.int32), fb_k=_arr(t.fb_k, t.fb_n, np.float64),
        fb_theta0=_arr(t.fb_theta0, t.fb_n, np.float64),
        ls_beadset=_arr(t.ls_beadset, 4 * t.ls_n, np.int32), ls_k=_arr(t.ls_k, t.ls_n, np.float64),
        ls_l0=_arr(t.ls_l0, t.ls_n, np.float64), ls_pos1=_arr(t.ls_pos1, t.ls_n, np.float64),
        ls_pos2=_arr(t.ls_pos2, t.ls_n, np.float64),
        ms_beadset=_arr(t.ms_beadset, 4 * t.ms_n, np.int32), ms_k=_arr(t.ms_k, t.ms_n, np.float64),
        ms_l0=_arr(t.ms_l0, t.ms_n, np.float64), ms_pos1=_arr(t.ms_pos1, t.ms_n, np.float64),
        ms_pos2=_arr(t.ms_pos2, t.ms_n, np.float64),
        bs_beadset=_arr(t.bs_beadset, 3 * t.bs_n, np.int32), bs_k=_arr(t.bs_k, t.bs_n, np.float64),
        bs_l0=_arr(t.bs_l0, t.bs_n, np.float64), bs_pos=_arr(t.bs_pos, t.bs_n, np.float64),
        bb_beadset=_arr(t.bb_beadset, 4 * t.bb_n, np.int32), bb_k=_arr(t.bb_k, t.bb_n, np.float64),
        bb_theta0=_arr(t.bb_theta0, t.bb_n, np.float64),
        plane=_arr(s.bnd.plane, 4 * s.bnd.n_planes, np.float64),
        plane_k=_arr(s.bnd.k, s.bnd.n_planes, np.float64),
        plane_screen=_arr(s.bnd.screen, s.bnd.n_planes, np.float64),
        bd_cutoff=float(s.bnd.cutoff), exvol_cutoff=float(s.exvol_cutoff),
        exvol_enabled=int(s.exvol_enabled), boundary_enabled=int(s.boundary_enabled), box=float(net.box),
    )


def gen_params(lib, **overrides) -> _abi.GenParams:
    p = _abi.GenParams()
    lib.mdy_gen_defaults(C.byref(p))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


def generate(lib, workload: Optional[str] = None, **overrides) -> Network:
    """Generate a seeded network with the library's generator (host code; no
    GPU needed). `workload` names one of WORKLOADS; keyword arguments override
    generator fields."""
    kw = dict(WORKLOADS[workload]) if workload else {}
    kw.update(overrides)
    p = gen_params(lib, **kw)
    out = C.POINTER(_abi.Network)()
    rc = lib.mdy_generate(C.byref(p), C.byref(out))
    if rc != 0:
        raise RuntimeError(f"mdy_generate failed: {_abi.ERRORS.get(rc, rc)}")
    try:
        return from_c(out.contents)
    finally:
        lib.mdy_network_free(out)
