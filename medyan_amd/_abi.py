"""ctypes mirror of include/mdy.h (the C ABI of libmdy_hip.so).

Only plain data layouts live here; loading the library and the host-side
mirror of MEDYAN's ForceFieldManager / CGMethod interface are in
medyan_amd.mechanics.
"""
import ctypes as C

MDY_ABI_VERSION = 5

MDY_OK = 0
ERRORS = {
    -1: "MDY_ERR_ARG", -2: "MDY_ERR_HIP", -3: "MDY_ERR_STATE", -4: "MDY_ERR_NONFINITE",
    -5: "MDY_ERR_CAPACITY", -6: "MDY_ERR_NOGPU", -7: "MDY_ERR_LAUNCH", -8: "MDY_ERR_FORCE",
    -9: "MDY_ERR_RECOVERY", -10: "MDY_ERR_TIMEOUT", -11: "MDY_ERR_QUADRATIC",
}
MDY_ERR_ARG = -1
MDY_ERR_HIP = -2
MDY_ERR_STATE = -3
MDY_ERR_NONFINITE = -4
MDY_ERR_CAPACITY = -5
MDY_ERR_NOGPU = -6
MDY_ERR_LAUNCH = -7
MDY_ERR_FORCE = -8
MDY_ERR_RECOVERY = -9
MDY_ERR_TIMEOUT = -10
MDY_ERR_QUADRATIC = -11
MDY_LS_BACKTRACKING, MDY_LS_QUADRATIC = 0, 1
MDY_EXVOL_INTEGRAL, MDY_EXVOL_MONOMER = 0, 1
MDY_BRDIH_COSINE, MDY_BRDIH_QUADRATIC = 0, 1

FF_NAMES = ["FilamentStretching", "FilamentBending", "LinkerStretching", "MotorGhostStretching",
            "BranchingStretching", "BranchingBending", "BranchingDihedral", "BranchingPosition",
            "CylinderExclVolume", "BoundaryCylinderRepulsion"]
MDY_NFF = 10
(MDY_FF_FS, MDY_FF_FB, MDY_FF_LS, MDY_FF_MS, MDY_FF_BS, MDY_FF_BB, MDY_FF_BDH, MDY_FF_BP, MDY_FF_EV,
 MDY_FF_BD) = range(MDY_NFF)

MDY_CG_STEEPEST, MDY_CG_FLETCHER_RIEVES, MDY_CG_POLAK_RIBIERE = 0, 1, 2
MDY_PREC_FP64, MDY_PREC_MIXED = 0, 1
MDY_COPLANAR_REFERENCE, MDY_COPLANAR_EXACT, MDY_COPLANAR_QUADRATURE = 0, 1, 2
MDY_ERROR_LINESEARCH = 1

_i32p = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_i16p = C.POINTER(C.c_int16)
_u8p = C.POINTER(C.c_uint8)
_f64p = C.POINTER(C.c_double)


class Cylinders(C.Structure):
    _fields_ = [("n", C.c_int32), ("bead1", _i32p), ("bead2", _i32p), ("id", _i64p),
                ("filament", _i32p), ("position", _i32p), ("type", _i16p),
                ("eq_length", _f64p), ("exvol_k", _f64p), ("branch_child", _i32p),
                ("minus_end", _u8p), ("mon_minus", _i32p), ("mon_plus", _i32p), ("mon_size", _f64p),
                ("mon_interval", _i32p)]


class FFTables(C.Structure):
    _fields_ = [
        ("fs_n", C.c_int32), ("fs_beadset", _i32p), ("fs_k", _f64p), ("fs_l0", _f64p),
        ("fb_n", C.c_int32), ("fb_beadset", _i32p), ("fb_k", _f64p), ("fb_theta0", _f64p),
        ("ls_n", C.c_int32), ("ls_beadset", _i32p), ("ls_k", _f64p), ("ls_l0", _f64p),
        ("ls_pos1", _f64p), ("ls_pos2", _f64p),
        ("ms_n", C.c_int32), ("ms_beadset", _i32p), ("ms_k", _f64p), ("ms_l0", _f64p),
        ("ms_pos1", _f64p), ("ms_pos2", _f64p),
        ("bs_n", C.c_int32), ("bs_beadset", _i32p), ("bs_k", _f64p), ("bs_l0", _f64p),
        ("bs_pos", _f64p),
        ("bb_n", C.c_int32), ("bb_beadset", _i32p), ("bb_k", _f64p), ("bb_theta0", _f64p),
        ("bdh_n", C.c_int32), ("bdh_beadset", _i32p), ("bdh_k", _f64p), ("bdh_pos", _f64p),
        ("bp_n", C.c_int32), ("bp_beadset", _i32p), ("bp_k", _f64p), ("bp_pos", _f64p),
    ]


class Boundary(C.Structure):
    _fields_ = [("n_planes", C.c_int32), ("plane", _f64p), ("k", _f64p), ("screen", _f64p),
                ("cutoff", C.c_float)]


class System(C.Structure):
    _fields_ = [("n_beads", C.c_int32), ("cyl", Cylinders), ("ff", FFTables), ("bnd", Boundary),
                ("exvol_cutoff", C.c_float), ("exvol_enabled", C.c_int32),
                ("boundary_enabled", C.c_int32), ("exvol_type", C.c_int32),
                ("br_dihedral_type", C.c_int32)]


class CGParams(C.Structure):
    _fields_ = [("method", C.c_int32), ("max_dist", C.c_double), ("lambda_max", C.c_double),
                ("grad_tol", C.c_double), ("energy_rel_tol", C.c_double),
                ("energy_call_limit", C.c_int32), ("force_call_limit", C.c_int32),
                ("n_beads_limit", C.c_int32), ("try_to_recover", C.c_int32), ("line_search", C.c_int32),
                ("report_on_call_limit", C.c_int32)]


class MinResult(C.Structure):
    _fields_ = [("energies_before", C.c_double * (MDY_NFF + 1)),
                ("energies_after", C.c_double * (MDY_NFF + 1)),
                ("error_flags", C.c_uint32), ("num_energy_call", C.c_int32),
                ("energy_call_limit", C.c_int32), ("num_force_call", C.c_int32),
                ("force_call_limit", C.c_int32), ("num_iter", C.c_int32),
                ("grad_inf_norm", C.c_double), ("grad_tol", C.c_double),
                ("energy_rel_change", C.c_double), ("energy_rel_tol", C.c_double),
                ("num_energy_eval", C.c_int32), ("num_safe_mode", C.c_int32),
                ("num_recovery", C.c_int32), ("status", C.c_int32)]

    def as_dict(self):
        d = {f: getattr(self, f) for f, _ in self._fields_}
        d["energies_before"] = list(self.energies_before)
        d["energies_after"] = list(self.energies_after)
        return d


class Opts(C.Structure):
    _fields_ = [("precision", C.c_int32), ("coplanar_mode", C.c_int32),
                ("block_size", C.c_int32), ("blocks_per_cu", C.c_int32)]


class GenParams(C.Structure):
    _fields_ = [("seed", C.c_uint64), ("n_filaments", C.c_int32), ("cyl_per_filament", C.c_int32),
                ("box", C.c_double), ("n_linkers", C.c_int32), ("n_motors", C.c_int32),
                ("n_branches", C.c_int32), ("cyl_size", C.c_double), ("jitter", C.c_double),
                ("min_sep", C.c_double), ("pair_cut", C.c_double), ("rest_sigma", C.c_double),
                ("k_stretch", C.c_double), ("k_bend", C.c_double), ("theta_bend", C.c_double),
                ("k_exvol", C.c_double), ("exvol_cutoff", C.c_float),
                ("k_bd", C.c_double), ("screen_bd", C.c_double), ("bd_cutoff", C.c_float),
                ("k_linker", C.c_double), ("k_motor", C.c_double),
                ("k_br_stretch", C.c_double), ("l_br_stretch", C.c_double),
                ("k_br_bend", C.c_double), ("theta_br_bend", C.c_double),
                ("k_br_dihedral", C.c_double), ("k_br_position", C.c_double)]


class Network(C.Structure):
    _fields_ = [("sys", System), ("coord", _f64p), ("box", C.c_double), ("n_filaments", C.c_int32),
                ("fil_first_bead", _i32p), ("fil_ncyl", _i32p), ("storage", C.c_void_p)]


def default_cg_params(n_beads, grad_tol=1.0, max_dist=1.0, lambda_max=0.01,
                      force_call_limit=100000, method=MDY_CG_POLAK_RIBIERE,
                      energy_rel_tol=1e-9, try_to_recover=1, energy_call_limit=0,
                      line_search=MDY_LS_BACKTRACKING, report_on_call_limit=1):
    """CG parameters of the benchmark workloads: GRADIENTTOLERANCE 1, MAXDISTANCE 1,
    LAMBDAMAX 0.01 (the shipped example values), energy-change tolerance 1e-9."""
    p = CGParams()
    p.method = method
    p.max_dist = max_dist
    p.lambda_max = lambda_max
    p.grad_tol = grad_tol
    p.energy_rel_tol = energy_rel_tol
    p.energy_call_limit = energy_call_limit
    p.force_call_limit = force_call_limit
    p.n_beads_limit = n_beads
    p.try_to_recover = try_to_recover
    p.line_search = line_search
    p.report_on_call_limit = report_on_call_limit
    return p
