"""Python host binding of libmdy_hip.so.

Mirrors MEDYAN's mechanics interface for the minimization path:

* ``ForceFieldManager``  -- vectorizeAllForceFields / computeEnergy /
  computeIndividualEnergies / computeForces
  (src/Mechanics/ForceField/ForceFieldManager.h:48-216)
* ``CGMethod.minimize``  -- CGMethod::minimize<polakRibiere> with backtracking
  (src/Mechanics/Minimizer/CGMethod.hpp:826-1759), run on the device
* ``ConjugateGradientParams`` / ``MinimizationResult``

Error behaviour follows the reference: an energy that a force field reports
as invalid makes computeEnergy return +inf; a non-finite max force raises
RuntimeError("maxF error") like CGMethod::maxF; a line-search failure without
recovery sets MinimizationResult.errorFlags.

The library is required: there is no CPU fallback. Loading fails loudly if
libmdy_hip.so is missing, and every GPU entry point fails if no HIP device is
present.
"""
import atexit
import ctypes as C
import os
import weakref
from dataclasses import dataclass

import numpy as np

from . import _abi
from .network import Network

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdy_hip.so")

_lib = None
# every open context, closed at interpreter exit before the HIP runtime and
# torch tear down (a context freed from a finalizer after that would call
# into a runtime that is gone)
_live = weakref.WeakSet()


@atexit.register
def close_all():
    """Destroys every context still open (also run at interpreter exit)."""
    for ctx in list(_live):
        ctx.close()


def load_library(path: str = LIB_PATH):
    """Load libmdy_hip.so (built in-tree by medyan_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: build it with `python -m medyan_amd.build` "
                           "(or __graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(path)
    f64p = C.POINTER(C.c_double)
    i32p = C.POINTER(C.c_int32)
    vp = C.c_void_p
    sig = {
        "mdy_abi_version": (C.c_int, []),
        "mdy_device_count": (C.c_int, []),
        "mdy_default_opts": (None, [C.POINTER(_abi.Opts)]),
        "mdy_create": (C.c_int, [C.POINTER(vp), C.c_int, C.POINTER(_abi.Opts)]),
        "mdy_destroy": (None, [vp]),
        "mdy_last_error": (C.c_char_p, [vp]),
        "mdy_set_opts": (C.c_int, [vp, C.POINTER(_abi.Opts)]),
        "mdy_set_network": (C.c_int, [vp, C.POINTER(_abi.System)]),
        "mdy_build_neighbors": (C.c_int, [vp, f64p]),
        "mdy_build_neighbors_dev": (C.c_int, [vp, vp]),
        "mdy_get_pairs": (C.c_int, [vp, i32p, C.c_int64, C.POINTER(C.c_int64)]),
        "mdy_get_boundary_table": (C.c_int, [vp, i32p, i32p, C.c_int64, C.POINTER(C.c_int64)]),
        "mdy_get_exvol_table": (C.c_int, [vp, i32p, f64p, C.c_int64, C.POINTER(C.c_int64)]),
        "mdy_energy": (C.c_int, [vp, f64p, f64p, f64p]),
        "mdy_forces": (C.c_int, [vp, f64p, f64p]),
        "mdy_energy_dev": (C.c_int, [vp, vp, vp, vp]),
        "mdy_forces_dev": (C.c_int, [vp, vp, vp, C.c_int, vp]),
        "mdy_energy_dev_noalloc": (C.c_int, [vp, vp, vp, vp]),
        "mdy_hvp_dev": (C.c_int, [vp, vp, vp, vp, vp]),
        "mdy_hvp_dev_async": (C.c_int, [vp, vp, vp, vp, vp]),
        "mdy_sync_errors": (C.c_int, [vp, vp]),
        "mdy_energy_jvp_dev": (C.c_int, [vp, vp, vp, vp, vp]),
        "mdy_quadrature_pair_count": (C.c_int, [vp, f64p, i32p]),
        "mdy_minimize": (C.c_int, [vp, C.POINTER(_abi.CGParams), f64p, f64p, C.POINTER(_abi.MinResult)]),
        "mdy_minimize_dev": (C.c_int, [vp, C.POINTER(_abi.CGParams), vp, vp, C.POINTER(_abi.MinResult)]),
        "mdy_minimize_batch": (C.c_int, [C.POINTER(vp), C.c_int, C.POINTER(_abi.CGParams), C.POINTER(vp),
                                         C.POINTER(vp), C.POINTER(_abi.MinResult), f64p, C.c_int, f64p, C.c_int,
                                         vp, f64p]),
        "mdy_stretch_forces": (C.c_int, [vp, f64p, f64p, f64p, f64p]),
        "mdy_load_forces": (C.c_int, [vp, f64p, C.c_int, f64p, i32p, C.c_int, f64p, f64p]),
        "mdy_get_trace": (C.c_int, [vp, f64p, f64p, f64p, f64p, i32p, C.c_int64, C.POINTER(C.c_int64)]),
        "mdy_set_trace": (C.c_int, [vp, C.c_int64]),
        "mdy_last_timing": (C.c_int, [vp, f64p]),
        "mdy_set_profile": (C.c_int, [vp, C.c_int64]),
        "mdy_get_profile": (C.c_int, [vp, i32p, C.POINTER(C.c_int64), C.c_int64, C.POINTER(C.c_int64)]),
        "mdy_get_profile_detail": (C.c_int, [vp, C.POINTER(C.c_int64), C.c_int64, C.POINTER(C.c_int64)]),
        "mdy_gen_defaults": (None, [C.POINTER(_abi.GenParams)]),
        "mdy_generate": (C.c_int, [C.POINTER(_abi.GenParams), C.POINTER(C.POINTER(_abi.Network))]),
        "mdy_network_free": (None, [C.POINTER(_abi.Network)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if lib.mdy_abi_version() != _abi.MDY_ABI_VERSION:
        raise RuntimeError("libmdy_hip.so ABI version mismatch")
    _lib = lib
    return lib


def _f64(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class MdyError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{_abi.ERRORS.get(code, code)}: {msg}")
        self.code = code


@dataclass
class ConjugateGradientParams:
    """ConjugateGradientParams (CGMethod.hpp:44-61) with the benchmark defaults
    (GRADIENTTOLERANCE 1, MAXDISTANCE 1, LAMBDAMAX 0.01)."""
    descentSearchMethod: int = _abi.MDY_CG_POLAK_RIBIERE
    maxDist: float = 1.0
    lambdaMax: float = 0.01
    gradTol: float = 1.0
    energyRelTol: float = 1e-9
    energyCallLimit: int = 0
    forceCallLimit: int = 10000
    tryToRecoverInLineSearchError: bool = True
    # "BACKTRACKING" or "QUADRATIC"; CGMethod::minimize (CGMethod.hpp:1319-1331)
    # takes any other name as QUADRATIC
    lineSearchAlgorithm: str = "BACKTRACKING"
    # restore the lowest-energy backup when the call limit ends the run
    # (CGMethod.hpp:59, 1589-1608)
    reportOnCallLimit: bool = True

    def to_c(self, n_beads) -> _abi.CGParams:
        p = _abi.CGParams()
        p.method = int(self.descentSearchMethod)
        p.max_dist = self.maxDist
        p.lambda_max = self.lambdaMax
        p.grad_tol = self.gradTol
        p.energy_rel_tol = self.energyRelTol
        p.energy_call_limit = int(self.energyCallLimit)
        p.force_call_limit = int(self.forceCallLimit)
        p.n_beads_limit = int(n_beads)
        p.try_to_recover = 1 if self.tryToRecoverInLineSearchError else 0
        p.line_search = (_abi.MDY_LS_BACKTRACKING if self.lineSearchAlgorithm == "BACKTRACKING"
                         else _abi.MDY_LS_QUADRATIC)
        p.report_on_call_limit = 1 if self.reportOnCallLimit else 0
        return p


class Context:
    """One mdy_ctx: a HIP device, its stream and the device-resident system."""

    def __init__(self, device=0, precision=_abi.MDY_PREC_FP64, coplanar_mode=_abi.MDY_COPLANAR_REFERENCE,
                 blocks_per_cu=0):
        self.lib = load_library()
        o = _abi.Opts()
        self.lib.mdy_default_opts(C.byref(o))
        o.precision = precision
        o.coplanar_mode = coplanar_mode
        o.blocks_per_cu = blocks_per_cu
        self.opts = o
        h = C.c_void_p()
        rc = self.lib.mdy_create(C.byref(h), device, C.byref(o))
        if rc != 0:
            raise MdyError(rc, "mdy_create failed (no HIP device?)")
        self.h = h
        self.net = None
        self._sys = None
        _live.add(self)

    def close(self):
        """mdy_destroy: frees the device memory; the context is unusable after."""
        if getattr(self, "h", None):
            self.lib.mdy_destroy(self.h)
            self.h = None
        _live.discard(self)

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        if rc != 0:
            raise MdyError(rc, self.lib.mdy_last_error(self.h).decode())

    def set_opts(self, **kw):
        """Changes context options (mdy_set_opts); on refusal the options stay as they were."""
        old = {k: getattr(self.opts, k) for k in kw}
        for k, v in kw.items():
            setattr(self.opts, k, v)
        rc = self.lib.mdy_set_opts(self.h, C.byref(self.opts))
        if rc != 0:
            for k, v in old.items():
                setattr(self.opts, k, v)
        self._check(rc)

    # -- system ----------------------------------------------------------------
    def set_network(self, net: Network):
        """Uploads a network. On failure the context holds no network (the
        library drops the previous one before validating)."""
        self.net = None
        self._sys = None
        sys_ = net.system()
        self._check(self.lib.mdy_set_network(self.h, C.byref(sys_)))
        self.net = net
        self._sys = sys_

    def _coord(self, coord, copy=False):
        """A contiguous float64 view (or copy) of coord, checked against the
        network's 3 * n_beads: the library reads and writes that many."""
        if self.net is None:
            raise MdyError(_abi.MDY_ERR_STATE, "no network set")
        x = np.ascontiguousarray(self.net.coord if coord is None else coord, dtype=np.float64)
        if x.size != 3 * self.net.n_beads:
            raise ValueError(f"coordinates have {x.size} entries, the network needs 3 * {self.net.n_beads}")
        return x.copy() if copy else x

    def build_neighbors(self, coord=None):
        x = self._coord(coord)
        self._check(self.lib.mdy_build_neighbors(self.h, _f64(x)))

    def pairs(self):
        n = C.c_int64()
        self._check(self.lib.mdy_get_pairs(self.h, None, 0, C.byref(n)))
        out = np.zeros(2 * max(n.value, 1), dtype=np.int32)
        self._check(self.lib.mdy_get_pairs(self.h, out.ctypes.data_as(C.POINTER(C.c_int32)), n.value, C.byref(n)))
        return out[: 2 * n.value].reshape(-1, 2)

    def exvol_table(self):
        """(beadset [n x 4], kR [n]) of the excluded-volume interactions."""
        n = C.c_int64()
        self._check(self.lib.mdy_get_exvol_table(self.h, None, None, 0, C.byref(n)))
        bs = np.zeros(4 * max(n.value, 1), dtype=np.int32)
        kr = np.zeros(max(n.value, 1))
        self._check(self.lib.mdy_get_exvol_table(self.h, bs.ctypes.data_as(C.POINTER(C.c_int32)), _f64(kr), n.value,
                                                 C.byref(n)))
        return bs[: 4 * n.value].reshape(-1, 4), kr[: n.value]

    def boundary_table(self):
        n = C.c_int64()
        npl = len(self.net.plane_k)
        nn = np.zeros(max(npl, 1), dtype=np.int32)
        self._check(self.lib.mdy_get_boundary_table(self.h, nn.ctypes.data_as(C.POINTER(C.c_int32)), None, 0,
                                                    C.byref(n)))
        bs = np.zeros(max(n.value, 1), dtype=np.int32)
        self._check(self.lib.mdy_get_boundary_table(self.h, nn.ctypes.data_as(C.POINTER(C.c_int32)),
                                                    bs.ctypes.data_as(C.POINTER(C.c_int32)), n.value, C.byref(n)))
        return nn[:npl], bs[: n.value]

    # -- ForceFieldManager -------------------------------------------------------
    def energy(self, coord):
        x = self._coord(coord)
        tot = C.c_double()
        per = np.zeros(_abi.MDY_NFF)
        self._check(self.lib.mdy_energy(self.h, _f64(x), C.byref(tot), _f64(per)))
        return tot.value, per

    def forces(self, coord, accum=None):
        x = self._coord(coord)
        f = np.zeros_like(x) if accum is None else self._coord(accum, copy=True)
        self._check(self.lib.mdy_forces(self.h, _f64(x), _f64(f)))
        return f

    # -- CGMethod ------------------------------------------------------------------
    def minimize(self, params, coord=None, trace_cap=0):
        """params: ConjugateGradientParams or _abi.CGParams. Returns (coord,
        force, MinimizationResult-like dict). Raises MdyError on failures the
        reference turns into exceptions."""
        if isinstance(params, ConjugateGradientParams):
            params = params.to_c(self.net.n_beads)
        x = self._coord(coord, copy=True)
        f = np.zeros_like(x)
        self._check(self.lib.mdy_set_trace(self.h, int(trace_cap)))
        res = _abi.MinResult()
        rc = self.lib.mdy_minimize(self.h, C.byref(params), _f64(x), _f64(f), C.byref(res))
        if rc not in (0,):
            raise MdyError(rc, self.lib.mdy_last_error(self.h).decode())
        return x, f, res

    # -- device-pointer forms (coordinates already in device memory) ------------
    def build_neighbors_dev(self, d_coord: int):
        """Neighbor lists from device coordinates (a device address, e.g. a
        torch tensor's data_ptr())."""
        self._check(self.lib.mdy_build_neighbors_dev(self.h, C.c_void_p(d_coord)))

    def minimize_dev(self, params, d_coord: int, d_force_out: int = 0, trace_cap=0):
        """Minimizes the device coordinates in place; returns the MinResult."""
        if isinstance(params, ConjugateGradientParams):
            params = params.to_c(self.net.n_beads)
        self._check(self.lib.mdy_set_trace(self.h, int(trace_cap)))
        res = _abi.MinResult()
        rc = self.lib.mdy_minimize_dev(self.h, C.byref(params), C.c_void_p(d_coord),
                                       C.c_void_p(d_force_out) if d_force_out else None, C.byref(res))
        if rc != 0:
            raise MdyError(rc, self.lib.mdy_last_error(self.h).decode())
        return res

    def energy_dev(self, d_coord: int, d_out: int, stream: int = 0):
        """d_out: MDY_NFF per-FF energies then the computeEnergy total."""
        self._check(self.lib.mdy_energy_dev(self.h, C.c_void_p(d_coord), C.c_void_p(d_out),
                                            C.c_void_p(stream) if stream else None))

    def energy_dev_noalloc(self, d_coord: int, d_out: int, stream: int = 0):
        """energy_dev without allocation (scratch in a static device array
        shared by every call on the GPU; calls that overlap on one GPU race)."""
        self._check(self.lib.mdy_energy_dev_noalloc(self.h, C.c_void_p(d_coord), C.c_void_p(d_out),
                                                    C.c_void_p(stream) if stream else None))

    def forces_dev(self, d_coord: int, d_force: int, accumulate=True, stream: int = 0):
        self._check(self.lib.mdy_forces_dev(self.h, C.c_void_p(d_coord), C.c_void_p(d_force), 1 if accumulate else 0,
                                            C.c_void_p(stream) if stream else None))

    def stretch_forces(self, coord):
        """assignallforcemags outputs at coord: (linker, motor, branch [n x 3])."""
        x = self._coord(coord)
        n = self.net
        nbr = max(len(n.bs_k), len(n.bdh_k), len(n.bp_k))
        lk, mt, br = np.zeros(max(len(n.ls_k), 1)), np.zeros(max(len(n.ms_k), 1)), np.zeros(max(3 * nbr, 1))
        self._check(self.lib.mdy_stretch_forces(self.h, _f64(x), _f64(lk), _f64(mt), _f64(br)))
        return lk[: len(n.ls_k)], mt[: len(n.ms_k)], br[: 3 * nbr].reshape(-1, 3)

    def load_forces(self, coord, mon_size=(2.7,), n_mon=(40,), max_mon=None):
        """computeLoadForces at coord (boundary neighbor list of the last
        build): (plus, minus), each n_beads x max_mon."""
        x = self._coord(coord)
        ms = np.ascontiguousarray(mon_size, dtype=np.float64)
        nm = np.ascontiguousarray(n_mon, dtype=np.int32)
        mm = int(max_mon if max_mon is not None else nm.max())
        nb = self.net.n_beads
        lp, lm = np.zeros(nb * mm), np.zeros(nb * mm)
        self._check(self.lib.mdy_load_forces(self.h, _f64(x), len(ms), _f64(ms),
                                             nm.ctypes.data_as(C.POINTER(C.c_int32)), mm, _f64(lp), _f64(lm)))
        return lp.reshape(nb, mm), lm.reshape(nb, mm)

    def hvp_dev(self, d_coord: int, d_v: int, d_hv: int, stream: int = 0):
        """H v of the exact-gradient force field at the device coordinates."""
        self._check(self.lib.mdy_hvp_dev(self.h, C.c_void_p(d_coord), C.c_void_p(d_v), C.c_void_p(d_hv),
                                         C.c_void_p(stream) if stream else None))

    def hvp_dev_async(self, d_coord: int, d_v: int, d_hv: int, stream: int = 0):
        """H v enqueued without waiting; a device failure is raised by
        sync_errors (or the next call that waits)."""
        self._check(self.lib.mdy_hvp_dev_async(self.h, C.c_void_p(d_coord), C.c_void_p(d_v), C.c_void_p(d_hv),
                                               C.c_void_p(stream) if stream else None))

    def sync_errors(self, stream: int = 0):
        """Waits for the stream and raises the failure an asynchronous call left."""
        self._check(self.lib.mdy_sync_errors(self.h, C.c_void_p(stream) if stream else None))

    def energy_jvp_dev(self, d_coord: int, d_w: int, d_out: int, stream: int = 0):
        """d_out[k] = grad E_k . w per force field, then the sum (MDY_NFF + 1 doubles)."""
        self._check(self.lib.mdy_energy_jvp_dev(self.h, C.c_void_p(d_coord), C.c_void_p(d_w), C.c_void_p(d_out),
                                                C.c_void_p(stream) if stream else None))

    def quadrature_pair_count(self, coord=None) -> int:
        """Excluded-volume pairs that MDY_COPLANAR_QUADRATURE evaluates by
        quadrature at coord (default: the network's coordinates)."""
        x = self._coord(coord)
        n = C.c_int32(0)
        self._check(self.lib.mdy_quadrature_pair_count(self.h, _f64(x), C.byref(n)))
        return int(n.value)

    def search_dir(self):
        """CGMethod::searchDir after the last minimization (see mdy_get_search_dir)."""
        d = np.zeros(3 * self.net.n_beads)
        self._check(self.lib.mdy_get_search_dir(self.h, _f64(d)))
        return d

    def trace(self):
        n = C.c_int64()
        self._check(self.lib.mdy_get_trace(self.h, None, None, None, None, None, 0, C.byref(n)))
        m = int(n.value)
        arrs = [np.zeros(max(m, 1)) for _ in range(4)] + [np.zeros(max(m, 1), dtype=np.int32)]
        self._check(self.lib.mdy_get_trace(self.h, *[_f64(a) for a in arrs[:4]],
                                           arrs[4].ctypes.data_as(C.POINTER(C.c_int32)), m, C.byref(n)))
        m = min(m, int(n.value))
        return dict(lambda_=arrs[0][:m], energy=arrs[1][:m], maxf=arrs[2][:m], beta=arrs[3][:m], safe=arrs[4][:m])

    def set_profile(self, cap):
        self._check(self.lib.mdy_set_profile(self.h, int(cap)))

    def profile(self):
        """(kind, ticks) of the phases of the last minimization; ticks are
        100 MHz device wall-clock stamps taken at each phase's barrier exit."""
        n = C.c_int64()
        self._check(self.lib.mdy_get_profile(self.h, None, None, 0, C.byref(n)))
        m = int(n.value)
        kind = np.zeros(max(m, 1), dtype=np.int32)
        ticks = np.zeros(max(m, 1), dtype=np.int64)
        self._check(self.lib.mdy_get_profile(self.h, kind.ctypes.data_as(C.POINTER(C.c_int32)),
                                             ticks.ctypes.data_as(C.POINTER(C.c_int64)), m, C.byref(n)))
        m = min(m, int(n.value))
        return kind[:m], ticks[:m]

    def profile_detail(self):
        """[n x 5] stamps per phase: barrier exit, latest arrival, block 0's
        work start, end of its main loop, its arrival (100 MHz ticks)."""
        n = C.c_int64()
        self._check(self.lib.mdy_get_profile_detail(self.h, None, 0, C.byref(n)))
        m = int(n.value)
        rec = np.zeros(5 * max(m, 1), dtype=np.int64)
        self._check(self.lib.mdy_get_profile_detail(self.h, rec.ctypes.data_as(C.POINTER(C.c_int64)), m,
                                                    C.byref(n)))
        return rec[: 5 * m].reshape(-1, 5)

    def last_ms(self):
        t = C.c_double()
        self._check(self.lib.mdy_last_timing(self.h, C.byref(t)))
        return t.value


def minimize_batch(ctxs, params, d_coords, d_forces=None, threads_per_block=64, weights=None, first_pass_calls=0,
                   stream=0):
    """Batched minimization (mdy_minimize_batch): ctxs[i] holds network i with
    its neighbor lists; d_coords[i] its device coordinates (updated in place).
    params: one CGParams / ConjugateGradientParams for all, or one per
    network. weights (optional): block shares. first_pass_calls > 0: passes
    with force-call caps growing 4x; a capped network is suspended and resumed
    on more blocks. first_pass_calls = -k: passes end by completions (once
    max(1, m / k) of a pass's m networks have finished, the others are
    suspended and resumed on all the blocks). Returns (results, per-network
    completion ms, total ms)."""
    n = len(ctxs)
    lib = load_library()
    if not isinstance(params, (list, tuple)):
        params = [params] * n
    ps = (_abi.CGParams * n)()
    for i, p in enumerate(params):
        ps[i] = p.to_c(ctxs[i].net.n_beads) if isinstance(p, ConjugateGradientParams) else p
    hs = (C.c_void_p * n)(*[c.h.value if isinstance(c.h, C.c_void_p) else c.h for c in ctxs])
    xs = (C.c_void_p * n)(*[int(d) for d in d_coords])
    fs = (C.c_void_p * n)(*[int(d) if d else None for d in d_forces]) if d_forces is not None else None
    res = (_abi.MinResult * n)()
    ms = np.zeros(n)
    tot = C.c_double()
    w = np.ascontiguousarray(weights, dtype=np.float64) if weights is not None else None
    if w is not None and w.shape != (n,):
        raise ValueError("one weight per network")
    rc = lib.mdy_minimize_batch(hs, n, ps, xs, fs, res, _f64(ms), int(threads_per_block),
                                _f64(w) if w is not None else None, int(first_pass_calls),
                                C.c_void_p(stream) if stream else None, C.byref(tot))
    if rc != 0:
        raise MdyError(rc, lib.mdy_last_error(ctxs[0].h).decode())
    return list(res), ms, tot.value


class ForceFieldManager:
    """ForceFieldManager over the device: vectorizeAllForceFields uploads the
    tables and rebuilds the neighbor lists from the given coordinates."""

    def __init__(self, ctx: Context):
        self.ctx = ctx

    def vectorizeAllForceFields(self, net: Network, coord=None):
        self.ctx.set_network(net)
        self.ctx.build_neighbors(coord)

    def computeEnergy(self, coord):
        return self.ctx.energy(coord)[0]

    def computeIndividualEnergies(self, coord):
        tot, per = self.ctx.energy(coord)
        return dict(zip(_abi.FF_NAMES, per.tolist())), float(np.sum(per))

    def computeForces(self, coord, force, numVar=None):
        force[:] = 0.0
        force[:] = self.ctx.forces(coord, force)
        return force


class CGMethod:
    def __init__(self, ctx: Context):
        self.ctx = ctx

    def minimize(self, params: ConjugateGradientParams, coord, trace_cap=0):
        return self.ctx.minimize(params, coord, trace_cap=trace_cap)
