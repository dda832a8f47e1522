"""ctypes binding of the CPU oracle (oracle/libmdy_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product package.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from medyan_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PARITY = os.path.join(HERE, "libmdy_oracle.so")       # -O2 -ffp-contract=off (parity)
LIB_FAST = os.path.join(HERE, "libmdy_oracle_fast.so")    # reference release flags (timing)

_f64p = C.POINTER(C.c_double)
_i32p = C.POINTER(C.c_int32)


def build(quiet=True):
    subprocess.run(["make", "-C", HERE], check=True,
                   stdout=subprocess.DEVNULL if quiet else None)


class OrcTrace(C.Structure):
    _fields_ = [("cap", C.c_int64), ("n", C.c_int64), ("lambda_", _f64p), ("energy", _f64p),
                ("maxf", _f64p), ("beta", _f64p), ("safe", _i32p)]


class OrcSystem(C.Structure):
    _fields_ = [("sys", _abi.System),
                ("ev_n", C.c_int64), ("ev_bs", _i32p), ("ev_krep", _f64p), ("ev_eql", _f64p), ("ev_mon", _i32p),
                ("bd_n", C.c_int64), ("bd_nneigh", _i32p), ("bd_bs", _i32p), ("bd_krep", _f64p),
                ("bd_slen", _f64p), ("coplanar_mode", C.c_int)]


ENERGY_FN = C.CFUNCTYPE(C.c_double, C.c_void_p, _f64p, _f64p, C.c_int)
FORCE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, _f64p, _f64p, C.c_int32)
RECOVER_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, _f64p, _f64p, C.c_int32, C.c_double)


def _p(a, t=C.c_double):
    return a.ctypes.data_as(C.POINTER(t))


class Oracle:
    def __init__(self, fast=False):
        path = LIB_FAST if fast else LIB_PARITY
        if not os.path.exists(path):
            build()
        self.lib = C.CDLL(path)
        L = self.lib
        L.orc_ffm_energy.restype = C.c_double
        L.orc_ffm_energy_individual.restype = C.c_double
        L.orc_ev_energy.restype = C.c_double
        L.orc_evm_energy.restype = C.c_double
        L.orc_evm_vectorize.restype = C.c_int64
        L.orc_last_search_dir.restype = C.c_int32
        L.orc_ev_pair_energy_analytic.restype = C.c_double
        L.orc_ev_pair_energy_simpson.restype = C.c_double
        for fn in ("orc_fs_energy", "orc_fb_energy", "orc_ls_energy", "orc_bs_energy", "orc_bb_energy",
                   "orc_bd_energy", "orc_bdh_energy", "orc_bdhq_energy", "orc_bp_energy"):
            getattr(L, fn).restype = C.c_double
        L.orc_nl_pairs.restype = C.c_int64
        L.orc_nl_pairs_grid.restype = C.c_int64
        L.orc_ev_vectorize.restype = C.c_int64
        L.orc_bd_vectorize.restype = C.c_int64
        L.orc_minimize.argtypes = [C.POINTER(_abi.CGParams), _f64p, C.c_int32, C.c_int32, ENERGY_FN, FORCE_FN,
                                   RECOVER_FN, C.c_void_p, _f64p, C.POINTER(_abi.MinResult), C.POINTER(OrcTrace),
                                   C.c_int32]
        L.orc_set_exact.argtypes = [C.c_int]
        L.orc_get_exact.restype = C.c_int
        L.orc_xsum.restype = C.c_double
        L.orc_xsum.argtypes = [_f64p, C.c_int64, C.c_int64]
        L.orc_xmath_atan.argtypes = [_f64p, _f64p, C.c_int64]
        L.orc_xmath_exp.argtypes = [_f64p, _f64p, C.c_int64]
        L.orc_xmath_acos.argtypes = [_f64p, _f64p, C.c_int64]

    # -- numerics mode ------------------------------------------------------
    def set_exact(self, on):
        """1: the GPU's shared atan/exp and exact sums; 0: libm and sequential sums."""
        self.lib.orc_set_exact(1 if on else 0)

    @property
    def exact(self):
        return bool(self.lib.orc_get_exact())

    def xmath_atan(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self.lib.orc_xmath_atan(_p(x), _p(y), x.size)
        return y

    def xmath_exp(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self.lib.orc_xmath_exp(_p(x), _p(y), x.size)
        return y

    def xmath_acos(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        self.lib.orc_xmath_acos(_p(x), _p(y), x.size)
        return y

    def xsum(self, x, stride=1):
        x = np.ascontiguousarray(x, dtype=np.float64)
        return self.lib.orc_xsum(_p(x), x.size, int(stride))

    # -- whole system -----------------------------------------------------
    def system(self, net, coord=None, coplanar_mode=_abi.MDY_COPLANAR_REFERENCE):
        return OracleSystem(self, net, net.coord if coord is None else coord, coplanar_mode)

    # -- single-pair helpers ----------------------------------------------
    def ev_pair_energy(self, c1, c2, c3, c4, kr):
        a = [np.ascontiguousarray(c, dtype=np.float64) for c in (c1, c2, c3, c4)]
        return self.lib.orc_ev_pair_energy_analytic(*[_p(x) for x in a], C.c_double(kr))

    def ev_pair_energy_simpson(self, c1, c2, c3, c4, kr, n1=400, n2=400):
        a = [np.ascontiguousarray(c, dtype=np.float64) for c in (c1, c2, c3, c4)]
        return self.lib.orc_ev_pair_energy_simpson(*[_p(x) for x in a], C.c_double(kr), n1, n2)

    def ev_pair_force_simpson(self, c1, c2, c3, c4, kr, n1, n2):
        a = [np.ascontiguousarray(c, dtype=np.float64) for c in (c1, c2, c3, c4)]
        out = np.zeros(12)
        self.lib.orc_ev_pair_force_simpson(*[_p(x) for x in a], C.c_double(kr), n1, n2, _p(out))
        return out.reshape(4, 3)

    def are_in_plane(self, c1, c2, c3, c4):
        a = [np.ascontiguousarray(c, dtype=np.float64) for c in (c1, c2, c3, c4)]
        return bool(self.lib.orc_are_in_plane(*[_p(x) for x in a]))

    # -- per-FF on explicit tables ----------------------------------------
    # bonded / cross-link force fields: name in fs, fb, ls, bs, bb; params in
    # the C order (fs: k, l0; fb: k, theta0; ls: k, l0, pos1, pos2;
    # bs: k, l0, pos; bb: k, theta0); bead offsets already x3
    _NPAR = {"fs": 2, "fb": 2, "ls": 4, "bs": 3, "bb": 2, "bdh": 2, "bdhq": 2, "bp": 2}
    _FSTRETCH = {"ls": 1, "bs": 3, "bb": 3, "bdh": 3, "bdhq": 3, "bp": 3}

    def _ff_args(self, name, x, bs, params):
        if len(params) != self._NPAR[name]:
            raise ValueError(f"{name}: {self._NPAR[name]} parameter arrays expected")
        x = np.ascontiguousarray(x, np.float64)
        bs = np.ascontiguousarray(bs, np.int32)
        ps = [np.ascontiguousarray(p, np.float64) for p in params]
        return x, bs, ps, len(ps[0])

    def ff_energy(self, name, x, bs, *params):
        x, bs, ps, n = self._ff_args(name, x, bs, params)
        return getattr(self.lib, f"orc_{name}_energy")(_p(x), n, _p(bs, C.c_int32), *[_p(p) for p in ps])

    def ff_forces(self, name, x, bs, *params):
        x, bs, ps, n = self._ff_args(name, x, bs, params)
        f = np.zeros_like(x)
        extra = []
        if name in self._FSTRETCH:
            extra = [None]
        getattr(self.lib, f"orc_{name}_forces")(_p(x), _p(f), n, _p(bs, C.c_int32), *[_p(p) for p in ps], *extra)
        return f

    def ev_energy(self, x, bs, krep, eql):
        x, bs, krep, eql = (np.ascontiguousarray(x, np.float64), np.ascontiguousarray(bs, np.int32),
                            np.ascontiguousarray(krep, np.float64), np.ascontiguousarray(eql, np.float64))
        return self.lib.orc_ev_energy(_p(x), len(krep), _p(bs, C.c_int32), _p(krep), _p(eql))

    def ev_forces(self, x, bs, krep, eql, coplanar_mode=0):
        x, bs, krep, eql = (np.ascontiguousarray(x, np.float64), np.ascontiguousarray(bs, np.int32),
                            np.ascontiguousarray(krep, np.float64), np.ascontiguousarray(eql, np.float64))
        f = np.zeros_like(x)
        rc = self.lib.orc_ev_forces(_p(x), _p(f), len(krep), _p(bs, C.c_int32), _p(krep), _p(eql), coplanar_mode)
        if rc:
            raise RuntimeError(_abi.ERRORS.get(rc, rc))
        return f

    def last_search_dir(self):
        """CGMethod::searchDir as the last minimization left it."""
        n = self.lib.orc_last_search_dir(None, 0)
        out = np.zeros(n)
        self.lib.orc_last_search_dir(_p(out), n)
        return out

    def evm_energy(self, x, bs, mon, meq, kvol):
        """Monomer-sum excluded volume: bs (n, 4) coordinate offsets, mon (n, 6)
        minus/plus/interval of both cylinders, meq (n, 2) monomer sizes, kvol (n)."""
        x, bs, mon, meq, kvol = (np.ascontiguousarray(x, np.float64), np.ascontiguousarray(bs, np.int32),
                                 np.ascontiguousarray(mon, np.int32), np.ascontiguousarray(meq, np.float64),
                                 np.ascontiguousarray(kvol, np.float64))
        return self.lib.orc_evm_energy(_p(x), len(kvol), _p(bs, C.c_int32), _p(mon, C.c_int32), _p(meq), _p(kvol))

    def evm_forces(self, x, bs, mon, meq, kvol):
        x, bs, mon, meq, kvol = (np.ascontiguousarray(x, np.float64), np.ascontiguousarray(bs, np.int32),
                                 np.ascontiguousarray(mon, np.int32), np.ascontiguousarray(meq, np.float64),
                                 np.ascontiguousarray(kvol, np.float64))
        f = np.zeros_like(x)
        self.lib.orc_evm_forces(_p(x), _p(f), len(kvol), _p(bs, C.c_int32), _p(mon, C.c_int32), _p(meq), _p(kvol))
        return f

    def bd_energy(self, x, plane, nneigh, bs, krep, slen):
        arrs = [np.ascontiguousarray(a, t) for a, t in ((x, np.float64), (plane, np.float64), (nneigh, np.int32),
                                                        (bs, np.int32), (krep, np.float64), (slen, np.float64))]
        x, plane, nneigh, bs, krep, slen = arrs
        return self.lib.orc_bd_energy(_p(x), len(nneigh), _p(plane), _p(nneigh, C.c_int32), _p(bs, C.c_int32),
                                      _p(krep), _p(slen))

    def bd_forces(self, x, plane, nneigh, bs, krep, slen):
        arrs = [np.ascontiguousarray(a, t) for a, t in ((x, np.float64), (plane, np.float64), (nneigh, np.int32),
                                                        (bs, np.int32), (krep, np.float64), (slen, np.float64))]
        x, plane, nneigh, bs, krep, slen = arrs
        f = np.zeros_like(x)
        self.lib.orc_bd_forces(_p(x), _p(f), len(nneigh), _p(plane), _p(nneigh, C.c_int32), _p(bs, C.c_int32),
                               _p(krep), _p(slen))
        return f

    def nl_pairs(self, net, coord=None, rmax=None, rmin=0.0, brute=False):
        """Neighbor pairs; brute=True runs the O(n^2) loop, otherwise the
        grid-pruned form (same pairs, same order)."""
        coord = net.coord if coord is None else np.ascontiguousarray(coord, np.float64)
        sysc = net.system()
        rmax = net.exvol_cutoff if rmax is None else rmax
        cap = 4 * net.n_cyl + 16
        fn = self.lib.orc_nl_pairs if brute else self.lib.orc_nl_pairs_grid
        while True:
            out = np.zeros(2 * cap, dtype=np.int32)
            n = fn(C.byref(sysc.cyl), _p(coord), C.c_float(rmax), C.c_float(rmin), _p(out, C.c_int32),
                   C.c_int64(cap))
            if n <= cap:
                return out[: 2 * n].reshape(-1, 2)
            cap = n

    # -- generic minimizer over Python callables ----------------------------
    def minimize(self, params, x, ndof, energy, force, recover=None, iter_cap=0, trace_cap=0):
        """Run the oracle CGMethod on Python callables (used for the reference's
        quadratic test problem). energy(x)->float, force(x, f) fills f."""
        nvar = len(x)

        @ENERGY_FN
        def efn(user, xp, per_ff, individual):
            xv = np.ctypeslib.as_array(xp, shape=(nvar,))
            return float(energy(xv))

        @FORCE_FN
        def ffn(user, xp, fp, n):
            xv = np.ctypeslib.as_array(xp, shape=(nvar,))
            fv = np.ctypeslib.as_array(fp, shape=(n,))
            force(xv, fv)
            return 0

        @RECOVER_FN
        def rfn(user, xp, dp, nd, md):
            if recover is not None:
                recover(np.ctypeslib.as_array(xp, shape=(nvar,)), np.ctypeslib.as_array(dp, shape=(nd,)), md)
            return 0

        xa = np.ascontiguousarray(x, dtype=np.float64).copy()
        fout = np.zeros(nvar)
        res = _abi.MinResult()
        tr, keep = _make_trace(trace_cap)
        self.lib.orc_minimize(C.byref(params), _p(xa), nvar, ndof, efn, ffn, rfn, None, _p(fout), C.byref(res),
                              C.byref(tr) if tr is not None else None, iter_cap)
        return xa, fout, res, _read_trace(tr, keep)


def _make_trace(cap):
    if not cap:
        return None, None
    keep = dict(lam=np.zeros(cap), e=np.zeros(cap), m=np.zeros(cap), b=np.zeros(cap),
                s=np.zeros(cap, dtype=np.int32))
    tr = OrcTrace(cap, 0, _p(keep["lam"]), _p(keep["e"]), _p(keep["m"]), _p(keep["b"]), _p(keep["s"], C.c_int32))
    return tr, keep


def _read_trace(tr, keep):
    if tr is None:
        return None
    n = tr.n
    return dict(lambda_=keep["lam"][:n].copy(), energy=keep["e"][:n].copy(), maxf=keep["m"][:n].copy(),
                beta=keep["b"][:n].copy(), safe=keep["s"][:n].copy())


class OracleSystem:
    """orc_system: the reference ForceFieldManager over one network (neighbor
    lists built from `coord` at construction, like MController::run)."""

    def __init__(self, oracle, net, coord, coplanar_mode):
        self.o = oracle
        self.net = net
        self.sysc = net.system()
        self.coord0 = np.ascontiguousarray(coord, dtype=np.float64).copy()
        self.s = OrcSystem()
        oracle.lib.orc_system_setup(C.byref(self.s), C.byref(self.sysc), _p(self.coord0))
        self.s.coplanar_mode = coplanar_mode

    def __del__(self):
        try:
            self.o.lib.orc_system_free(C.byref(self.s))
        except Exception:
            pass

    def stretch_forces(self, x=None):
        """assignforcemags outputs at x: linker and motor stretch forces,
        branch-point 3-vectors (LinkerStretchingHarmonic.cpp:293,
        MotorGhostStretchingHarmonic.cpp:339, BranchingStretchingHarmonic.cpp:273-275)."""
        x = self.coord0 if x is None else np.ascontiguousarray(x, np.float64)
        n = self.net
        f = np.zeros_like(x)
        out = []
        for bs, k, l0, p1, p2 in ((n.ls_beadset, n.ls_k, n.ls_l0, n.ls_pos1, n.ls_pos2),
                                  (n.ms_beadset, n.ms_k, n.ms_l0, n.ms_pos1, n.ms_pos2)):
            st = np.zeros(len(k))
            if len(k):
                arr = [np.ascontiguousarray(a, t) for a, t in ((bs, np.int32), (k, np.float64), (l0, np.float64),
                                                               (p1, np.float64), (p2, np.float64))]
                self.o.lib.orc_ls_forces(_p(x), _p(f), len(k), _p(arr[0], C.c_int32), _p(arr[1]), _p(arr[2]),
                                         _p(arr[3]), _p(arr[4]), _p(st))
            out.append(st)
        # branch force: stretching, dihedral and position contributions added
        # in force-field order (assignforcemags of each, BranchingStretching.cpp:86-97,
        # BranchingDihedral.cpp:89-98, BranchingPosition.cpp)
        nbr = max(len(n.bs_k), len(n.bdh_k), len(n.bp_k))
        br = np.zeros(3 * nbr)
        if len(n.bs_k):
            part = np.zeros(3 * len(n.bs_k))
            arr = [np.ascontiguousarray(a, t) for a, t in ((n.bs_beadset, np.int32), (n.bs_k, np.float64),
                                                           (n.bs_l0, np.float64), (n.bs_pos, np.float64))]
            self.o.lib.orc_bs_forces(_p(x), _p(f), len(n.bs_k), _p(arr[0], C.c_int32), _p(arr[1]), _p(arr[2]),
                                     _p(arr[3]), _p(part))
            br[: part.size] += part
        bdh_fn = (self.o.lib.orc_bdhq_forces if n.br_dihedral_type == _abi.MDY_BRDIH_QUADRATIC
                  else self.o.lib.orc_bdh_forces)
        for fn, bs, k, pos in ((bdh_fn, n.bdh_beadset, n.bdh_k, n.bdh_pos),
                               (self.o.lib.orc_bp_forces, n.bp_beadset, n.bp_k, n.bp_pos)):
            if len(k):
                part = np.zeros(3 * len(k))
                arr = [np.ascontiguousarray(a, t) for a, t in ((bs, np.int32), (k, np.float64), (pos, np.float64))]
                fn(_p(x), _p(f), len(k), _p(arr[0], C.c_int32), _p(arr[1]), _p(arr[2]), _p(part))
                br[: part.size] += part
        out.append(br)
        return tuple(out)

    def load_forces(self, x, mon_size, n_mon, max_mon):
        """computeLoadForces at x with the boundary neighbor list of the
        construction coordinates: (plus, minus), each n_beads x max_mon."""
        x = np.ascontiguousarray(x, np.float64)
        nb = self.net.n_beads
        lp = np.zeros(nb * max_mon)
        lm = np.zeros(nb * max_mon)
        ms = np.ascontiguousarray(mon_size, np.float64)
        nm = np.ascontiguousarray(n_mon, np.int32)
        self.o.lib.orc_load_forces(C.byref(self.sysc.cyl), nb, _p(self.coord0), _p(x), C.byref(self.sysc.bnd),
                                   len(ms), _p(ms), _p(nm, C.c_int32), int(max_mon), _p(lp), _p(lm))
        return lp.reshape(nb, max_mon), lm.reshape(nb, max_mon)

    @property
    def ev_table(self):
        n = self.s.ev_n
        if n == 0:
            return np.zeros((0, 4), np.int32), np.zeros(0), np.zeros((0, 2))
        bs = np.ctypeslib.as_array(self.s.ev_bs, shape=(4 * n,)).reshape(-1, 4).copy()
        kr = np.ctypeslib.as_array(self.s.ev_krep, shape=(n,)).copy()
        eq = np.ctypeslib.as_array(self.s.ev_eql, shape=(2 * n,)).reshape(-1, 2).copy()
        return bs, kr, eq

    @property
    def bd_table(self):
        n = self.s.bd_n
        npl = self.sysc.bnd.n_planes
        nn = np.ctypeslib.as_array(self.s.bd_nneigh, shape=(npl,)).copy() if npl else np.zeros(0, np.int32)
        bs = np.ctypeslib.as_array(self.s.bd_bs, shape=(n,)).copy() if n else np.zeros(0, np.int32)
        return nn, bs

    def energy(self, x=None, individual=False):
        x = self.coord0 if x is None else np.ascontiguousarray(x, np.float64)
        per = np.zeros(_abi.MDY_NFF)
        fn = self.o.lib.orc_ffm_energy_individual if individual else self.o.lib.orc_ffm_energy
        tot = fn(C.byref(self.s), _p(x), _p(per))
        return tot, per

    def forces(self, x=None):
        x = self.coord0 if x is None else np.ascontiguousarray(x, np.float64)
        f = np.zeros_like(x)
        rc = self.o.lib.orc_ffm_forces(C.byref(self.s), _p(x), _p(f), len(x))
        if rc:
            raise RuntimeError(_abi.ERRORS.get(rc, rc))
        return f

    def minimize(self, params, x=None, iter_cap=0, trace_cap=0):
        x = (self.coord0 if x is None else np.ascontiguousarray(x, np.float64)).copy()
        fout = np.zeros_like(x)
        res = _abi.MinResult()
        tr, keep = _make_trace(trace_cap)
        self.o.lib.orc_minimize_system(C.byref(self.s), C.byref(params), _p(x), _p(fout), C.byref(res),
                                       C.byref(tr) if tr is not None else None, iter_cap)
        return x, fout, res, _read_trace(tr, keep)
