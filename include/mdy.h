/*
 * mdy.h -- C ABI of libmdy_hip.so, the MI355X (gfx950) implementation of
 * MEDYAN's mechanical-equilibration path: force-field energies/forces,
 * cylinder neighbor lists and the Polak-Ribiere conjugate-gradient minimizer.
 *
 * Conventions
 *   - Every entry point returns MDY_OK (0) or a negative MDY_ERR_* code;
 *     mdy_last_error() gives a message. No C++ exception crosses the ABI.
 *   - Host pointers are owned by the caller. A context owns its device memory.
 *   - Coordinates and forces are AoS doubles, bead b at [3b, 3b+1, 3b+2],
 *     exactly the layout of MEDYAN's serializeDof
 *     (src/Structure/DofSerializer.hpp:64-187).
 *   - Interaction tables use MEDYAN's vectorized layout: a beadSet entry is a
 *     coordinate OFFSET (3 * bead index), as written by each force field's
 *     vectorize() (e.g. src/Mechanics/ForceField/Filament/FilamentStretching.cpp:35-36).
 *   - One context = one HIP device + one stream. Not reentrant: calls on one
 *     context must come from one thread at a time (MEDYAN's caller is a single
 *     thread, src/Controller/MController.h:109-211).
 */
#ifndef MDY_H
#define MDY_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MDY_ABI_VERSION 5

enum {
    MDY_OK = 0,
    MDY_ERR_ARG = -1,        /* invalid argument or size */
    MDY_ERR_HIP = -2,        /* HIP runtime error */
    MDY_ERR_STATE = -3,      /* call out of order (e.g. no network set) */
    MDY_ERR_NONFINITE = -4,  /* max force is not finite (CGMethod::maxF throws) */
    MDY_ERR_CAPACITY = -5,   /* caller buffer too small */
    MDY_ERR_NOGPU = -6,      /* no HIP device */
    MDY_ERR_LAUNCH = -7,     /* cooperative launch not possible */
    MDY_ERR_FORCE = -8,      /* excluded-volume force stayed non-finite after the
                                quadrature fallback (reference exits) */
    MDY_ERR_RECOVERY = -9,   /* recovery move had infinite step (reference throws) */
    MDY_ERR_TIMEOUT = -10,   /* device-side grid barrier watchdog fired */
    MDY_ERR_QUADRATIC = -11  /* quadratic line search: positive ideal energy change (the reference exits) */
};

/* Force-field slots, in the order MController::initializeFF registers them
 * (src/Controller/MController.cpp:65-188). Energies are summed in this order. */
enum {
    MDY_FF_FS = 0, /* FilamentStretching<Harmonic>           */
    MDY_FF_FB,     /* FilamentBending<Cosine>                */
    MDY_FF_LS,     /* LinkerStretching<Harmonic>             */
    MDY_FF_MS,     /* MotorGhostStretching<Harmonic>         */
    MDY_FF_BS,     /* BranchingStretching<Harmonic>          */
    MDY_FF_BB,     /* BranchingBending<Cosine>               */
    MDY_FF_BDH,    /* BranchingDihedral<Cosine>              */
    MDY_FF_BP,     /* BranchingPosition<Cosine>              */
    MDY_FF_EV,     /* CylinderExclVolume<CylinderExclVolRepulsion> */
    MDY_FF_BD,     /* BoundaryCylinderRepulsion<Exp>         */
    MDY_NFF
};

/* Cylinder table: what HybridCylinderCylinderNL / BoundaryCylinderNL read from
 * the cylinder database (src/Structure/Cylinder.cpp:30-49). */
typedef struct mdy_cylinders {
    int32_t n;
    const int32_t *bead1;        /* coord offset of getFirstBead()                  */
    const int32_t *bead2;        /* coord offset of getSecondBead()                 */
    const int64_t *id;           /* getId(): half list keeps pairs with id(c) > id(n) */
    const int32_t *filament;     /* parent filament id                              */
    const int32_t *position;     /* position on filament                            */
    const int16_t *type;         /* filament type                                   */
    const double  *eq_length;    /* MCylinder::getEqLength                          */
    const double  *exvol_k;      /* MCylinder::getExVolConst                        */
    const int32_t *branch_child; /* index of getBranchingCylinder() or -1           */
    const uint8_t *minus_end;    /* isMinusEnd()                                    */
    /* monomer-sum excluded volume only (CylinderVolumeMon::vectorize,
     * Volume/CylinderVolumeMon.hpp:149-163); may be NULL otherwise */
    const int32_t *mon_minus;    /* getFirstBead()->monomerSerial                   */
    const int32_t *mon_plus;     /* getSecondBead()->monomerSerial                  */
    const double  *mon_size;     /* GeoParams::monomerSize[type]                    */
    const int32_t *mon_interval; /* volumeExclusionMonomerInterval[type], >= 1      */
} mdy_cylinders;

/* Vectorized bonded/cross-link tables (each FF's vectorize()). */
typedef struct mdy_ff_tables {
    /* FilamentStretching::vectorize  (FilamentStretching.cpp:26-46) */
    int32_t fs_n; const int32_t *fs_beadset; const double *fs_k; const double *fs_l0;
    /* FilamentBending::vectorize     (FilamentBending.cpp:30-60) */
    int32_t fb_n; const int32_t *fb_beadset; const double *fb_k; const double *fb_theta0;
    /* LinkerStretching::vectorize    (LinkerStretching.cpp:30-55) */
    int32_t ls_n; const int32_t *ls_beadset; const double *ls_k; const double *ls_l0;
    const double *ls_pos1; const double *ls_pos2;
    /* MotorGhostStretching::vectorize (MotorGhostStretching.cpp:30-60) */
    int32_t ms_n; const int32_t *ms_beadset; const double *ms_k; const double *ms_l0;
    const double *ms_pos1; const double *ms_pos2;
    /* BranchingStretching::vectorize (BranchingStretching.cpp:29-50) */
    int32_t bs_n; const int32_t *bs_beadset; const double *bs_k; const double *bs_l0;
    const double *bs_pos;
    /* BranchingBending::vectorize    (BranchingBending.cpp:29-48) */
    int32_t bb_n; const int32_t *bb_beadset; const double *bb_k; const double *bb_theta0;
    /* BranchingDihedral::vectorize   (BranchingDihedral.cpp:32-60): mother first and
     * second bead, daughter first and second bead; kdih; adjusted position */
    int32_t bdh_n; const int32_t *bdh_beadset; const double *bdh_k; const double *bdh_pos;
    /* BranchingPosition::vectorize   (BranchingPosition.cpp:27-50): mother first and
     * second bead, daughter first bead; kpos; adjusted position */
    int32_t bp_n; const int32_t *bp_beadset; const double *bp_k; const double *bp_pos;
} mdy_ff_tables;

/* Plane boundary elements (PlaneBoundaryElement, src/Structure/BoundaryElementImpl.h:30-119):
 * plane[4i..4i+3] = (a, b, c, d) of a*x + b*y + c*z + d = 0. */
typedef struct mdy_boundary {
    int32_t n_planes;
    const double *plane;
    const double *k;       /* getRepulsionConst   */
    const double *screen;  /* getScreeningLength  */
    float cutoff;          /* BoundaryCylinderNL rMax (stored as float in the reference) */
} mdy_boundary;

/* Everything a minimization needs besides coordinates. */
typedef struct mdy_system {
    int32_t n_beads;           /* independent DOF = 3 * n_beads */
    mdy_cylinders cyl;
    mdy_ff_tables ff;
    mdy_boundary bnd;
    float exvol_cutoff;        /* VolumeCutoff, rMax of the hybrid NL (float in the reference) */
    int32_t exvol_enabled;
    int32_t boundary_enabled;
    int32_t exvol_type;        /* MDY_EXVOL_*: cylinderVolumeExclusionFFType */
    int32_t br_dihedral_type;  /* MDY_BRDIH_*: BRDIHEDRALFFTYPE */
} mdy_system;
/* Branching dihedral form (BranchingFF.h:64-81): COSINE,
 * kdih (1 - cos theta) (BranchingDihedralCosine.cpp:211-661), or QUADRATIC,
 * kdih theta^2 (BranchingDihedralQuadratic.cpp:26-398). Both occupy the
 * MDY_FF_BDH slot and the same table (BranchingDihedral::vectorize). */
enum { MDY_BRDIH_COSINE = 0, MDY_BRDIH_QUADRATIC = 1 };
/* Cylinder excluded volume form (MController::initializeFF,
 * Controller/MController.cpp:128-153): the analytic integral
 * (CylinderExclVolume<CylinderExclVolRepulsion>) or the monomer-pair sum
 * (CylinderVolumeMon, Volume/CylinderVolumeMon.hpp:31-100). Both occupy the
 * MDY_FF_EV slot. */
enum { MDY_EXVOL_INTEGRAL = 0, MDY_EXVOL_MONOMER = 1 };

/* ConjugateGradientParams (src/Mechanics/Minimizer/CGMethod.hpp:44-61). */
enum { MDY_CG_STEEPEST = 0, MDY_CG_FLETCHER_RIEVES = 1, MDY_CG_POLAK_RIBIERE = 2 };
typedef struct mdy_cg_params {
    int32_t method;               /* MDY_CG_*                                          */
    double  max_dist;             /* MAXDISTANCE                                       */
    double  lambda_max;           /* LAMBDAMAX                                         */
    double  grad_tol;             /* GRADIENTTOLERANCE (component-wise infinity norm)  */
    double  energy_rel_tol;       /* energy-change-relative-tolerance (default 1e-9)   */
    int32_t energy_call_limit;    /* 0 = no limit                                      */
    int32_t force_call_limit;     /* effective limit is max(this, 3 * n_beads_limit)   */
    int32_t n_beads_limit;        /* Bead::numBeads() used by the call limit           */
    int32_t try_to_recover;       /* tryToRecoverInLineSearchError                     */
    int32_t line_search;          /* MDY_LS_*: lineSearchAlgorithm (CGMethod.hpp:51)   */
    int32_t report_on_call_limit; /* reportOnCallLimit (CGMethod.hpp:59; the reference's
                                     default is true, set 1): when the force-call limit
                                     ends the minimization, restore the lowest-energy
                                     coordinates backed up near convergence
                                     (CGMethod.hpp:1589-1608). MController sets it
                                     false for interrupted segments (MController.h:171). */
} mdy_cg_params;
/* BACKTRACKING (backtrackingLineSearch, the default) or QUADRATIC
 * (quadraticLineSearchV2 + quadraticoptimization outside safe mode,
 * CGMethod.hpp:522-765, 1319-1331). */
enum { MDY_LS_BACKTRACKING = 0, MDY_LS_QUADRATIC = 1 };

/* MinimizationResult (src/Mechanics/Minimizer/MinimizationTypes.hpp:10-64). */
#define MDY_ERROR_LINESEARCH 1u
typedef struct mdy_min_result {
    double   energies_before[MDY_NFF + 1]; /* per FF, then total */
    double   energies_after[MDY_NFF + 1];
    uint32_t error_flags;
    int32_t  num_energy_call;
    int32_t  energy_call_limit;
    int32_t  num_force_call;
    int32_t  force_call_limit;
    int32_t  num_iter;
    double   grad_inf_norm;
    double   grad_tol;
    double   energy_rel_change;
    double   energy_rel_tol;
    /* additions of this implementation */
    int32_t  num_energy_eval;       /* energy evaluations actually performed (E(x) is cached) */
    int32_t  num_safe_mode;         /* times safe mode was entered                         */
    int32_t  num_recovery;          /* recovery moves                                      */
    int32_t  status;                /* MDY_OK or the error that stopped the minimization   */
} mdy_min_result;

/* ------------------------------------------------------------------------- */
/* Context                                                                    */
/* ------------------------------------------------------------------------- */
enum { MDY_PREC_FP64 = 0, MDY_PREC_MIXED = 1 };
/* MDY_COPLANAR_REFERENCE: the reference's force path (coplanar pairs perturbed);
 * MDY_COPLANAR_EXACT: the exact gradient of the closed-form energy everywhere;
 * MDY_COPLANAR_QUADRATURE: as EXACT, but excluded-volume pairs whose closed
 *   form is ill-conditioned (squared normalized triple product below 1e-3:
 *   nearly coplanar or nearly parallel cylinders) are evaluated by composite
 *   Gauss-Legendre quadrature of the defining integral, so that their
 *   derivatives are not rounding noise. No counterpart in the reference; for
 *   the differentiable form. One-shot evaluations (energy, forces, H v,
 *   energy directional derivatives) only: FP64 contexts, no minimization. */
enum { MDY_COPLANAR_REFERENCE = 0, MDY_COPLANAR_EXACT = 1, MDY_COPLANAR_QUADRATURE = 2 };

typedef struct mdy_opts {
    int32_t precision;       /* MDY_PREC_*                                                 */
    int32_t coplanar_mode;   /* MDY_COPLANAR_REFERENCE: perturb c2 like the reference force */
    int32_t block_size;      /* threads per block of the persistent CG kernel (0 = auto)   */
    int32_t blocks_per_cu;   /* resident blocks per CU for the persistent kernel (0 = max) */
} mdy_opts;

typedef struct mdy_ctx mdy_ctx;

int         mdy_abi_version(void);
int         mdy_device_count(void);
void        mdy_default_opts(mdy_opts *opts);
int         mdy_create(mdy_ctx **out, int device, const mdy_opts *opts);
void        mdy_destroy(mdy_ctx *ctx);
const char *mdy_last_error(const mdy_ctx *ctx);
int         mdy_set_opts(mdy_ctx *ctx, const mdy_opts *opts);

/* Upload a system (tables, cylinders, boundary). Replaces ForceFieldManager::
 * vectorizeAllForceFields for the bonded / cross-link force fields
 * (src/Mechanics/ForceField/ForceFieldManager.cpp:35-93). */
int mdy_set_network(mdy_ctx *ctx, const mdy_system *sys);

/* Neighbor-list rebuild + excluded-volume / boundary vectorization from the
 * given coordinates: HybridCylinderCylinderNL::reset (HybridNeighborListImpl.cpp:527-562),
 * BoundaryCylinderNL::reset (NeighborListImpl.cpp:801-809), then
 * CylinderExclVolume::vectorize (CylinderExclVolume.cpp:34-179) and
 * BoundaryCylinderRepulsion::vectorize (BoundaryCylinderRepulsion.cpp:33-88). */
int mdy_build_neighbors(mdy_ctx *ctx, const double *coord);
/* Same from device-resident coordinates. */
int mdy_build_neighbors_dev(mdy_ctx *ctx, const double *d_coord);
/* CGMethod::searchDir after the context's last minimization (3 * n_beads
 * doubles to host memory): the direction after the last update, shifted by
 * beta after the final accepted step and reset to the force when conjugacy is
 * lost (CGMethod.hpp:1495-1507, CGMethod.cpp:356-360). MController::run passes
 * it with CGMethod::LAMBDATOL (MDY_LAMBDATOL) to its line-search diagnosis
 * after a line-search error (Controller/MController.h:142-152). Where a reset
 * made it the force itself, components that are exactly zero may differ in
 * sign from the reference's F + 0 * d. MDY_ERR_STATE before a minimization. */
int mdy_get_search_dir(mdy_ctx *ctx, double *dir_out);
#define MDY_LAMBDATOL 1e-8   /* CGMethod::LAMBDATOL (CGMethod.hpp:99) */

/* Cylinder-cylinder neighbor pairs (index of c, index of neighbor n,
 * id(c) > id(n)), sorted by (c, n): the list HybridCylinderCylinderNL holds,
 * mother-daughter pairs included. */
int mdy_get_pairs(mdy_ctx *ctx, int32_t *pairs, int64_t cap, int64_t *n);
/* Excluded-volume interaction table built from those pairs
 * (CylinderExclVolume::vectorize, CylinderExclVolume.cpp:34-179: pairs whose
 * neighbor is the cylinder's mother dropped): 4 bead offsets per pair and
 * kR = k * l1 * l2. In the monomer-sum form (CylinderVolumeMon::vectorize,
 * CylinderVolumeMon.hpp:126-185: mother-daughter pairs dropped in either
 * order) kR = max(k1, k2) * seg1 * seg2, seg = interval * monomer size.
 * beadset / krep may be NULL to query the count. */
int mdy_get_exvol_table(mdy_ctx *ctx, int32_t *beadset, double *krep, int64_t cap, int64_t *n);
/* Boundary interaction table: per-plane counts and bead offsets (reference order). */
int mdy_get_boundary_table(mdy_ctx *ctx, int32_t *nneighbors, int32_t *beadset, int64_t cap, int64_t *n);

/* ForceFieldManager::computeEnergy (ForceFieldManager.cpp:96-207): *e_total is
 * +INFINITY if a force field returned the -1 sentinel. e_per_ff (nullable)
 * receives MDY_NFF individual energies (computeIndividualEnergies). */
int mdy_energy(mdy_ctx *ctx, const double *coord, double *e_total, double *e_per_ff);
/* ForceFieldManager::computeForces (ForceFieldManager.cpp:242-290) but with
 * ForceField::computeForces semantics: forces are ADDED to force_accum. */
int mdy_forces(mdy_ctx *ctx, const double *coord, double *force_accum);

/* Device-pointer forms on a caller stream (hipStream_t passed as void*; NULL =
 * the context stream). d_out receives MDY_NFF energies then the total. */
int mdy_energy_dev(mdy_ctx *ctx, const double *d_coord, double *d_out, void *stream);
int mdy_forces_dev(mdy_ctx *ctx, const double *d_coord, double *d_force, int accumulate, void *stream);
/* mdy_energy_dev without any allocation: no hipMalloc, hipFree or allocating
 * library call on the call path. No reference counterpart; for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops). The
 * block partials, the quadrature-fallback queue and its counters, the
 * per-pair quadrature energies and the error word live in one static device
 * array of about 580 KB per GPU (2048 block records of 256 B, 4096 queue
 * entries), held until the process ends and shared by every call on that
 * GPU: calls that overlap on one GPU race. The counters are cleared by a
 * kernel at the start of each call. The context's network and neighbor list
 * must already be built (those calls allocate). Returns MDY_ERR_CAPACITY if
 * the launch would need more than 2048 block records. Results equal
 * mdy_energy_dev's. */
int mdy_energy_dev_noalloc(mdy_ctx *ctx, const double *d_coord, double *d_out, void *stream);
/* Hessian-vector product H v = d/dt grad E(x + t v) at t = 0 of the
 * exact-gradient force field (no coplanar perturbation); writes H v. */
int mdy_hvp_dev(mdy_ctx *ctx, const double *d_coord, const double *d_v, double *d_hv, void *stream);
/* Same product, enqueued on the stream without waiting for it (no host
 * synchronization per call, so successive products run back to back). A
 * device-side failure (a degenerate excluded-volume pair: MDY_ERR_FORCE)
 * stays in the context until mdy_sync_errors, or the next call that waits,
 * reports and clears it. No reference counterpart (the reference has no
 * H v). */
int mdy_hvp_dev_async(mdy_ctx *ctx, const double *d_coord, const double *d_v, double *d_hv, void *stream);
/* Waits for the stream (NULL = the context's) and returns the failure an
 * asynchronous call left, clearing it; MDY_OK if none. */
int mdy_sync_errors(mdy_ctx *ctx, void *stream);
/* Directional derivatives of the individual energies of the exact-gradient
 * force field: d_out[k] = grad E_k(x) . w for the MDY_NFF force fields, then
 * their sum. No reference counterpart (derivative of
 * ForceFieldManager::computeIndividualEnergies, ForceFieldManager.cpp:209-227);
 * used for parameter sensitivities of minimized structures. */
int mdy_energy_jvp_dev(mdy_ctx *ctx, const double *d_coord, const double *d_w, double *d_out, void *stream);
/* Number of excluded-volume pairs that the MDY_COPLANAR_QUADRATURE mode
 * evaluates by quadrature at host coordinates coord (0 in the other modes).
 * No reference counterpart; reports the mode's extent. */
int mdy_quadrature_pair_count(mdy_ctx *ctx, const double *coord, int32_t *count);

/* Post-minimization auxiliary outputs (MController::afterMinimization_,
 * Controller/MController.h:240-268), at coord:
 * ForceFieldManager::assignallforcemags -- linker[n_linkers] and
 * motor[n_motors] stretch forces f0 * d = k (d - l0)
 * (LinkerStretchingHarmonic.cpp:293, MotorGhostStretchingHarmonic.cpp:339),
 * branch3[3 * n_branches] branch-point forces: the branching stretching,
 * dihedral and position contributions added in force-field order
 * (BranchingStretchingHarmonic.cpp:273-275, BranchingDihedralCosine.cpp:503-505,
 * BranchingPositionCosine.cpp:445-447); entry i of the three branching tables
 * is the same branch point, n_branches the largest table. Each output nullable. */
int mdy_stretch_forces(mdy_ctx *ctx, const double *coord, double *linker, double *motor, double *branch3);
/* ForceFieldManager::computeLoadForces (ForceFieldManager.cpp:292-309) with
 * BoundaryCylinderRepulsion::computeLoadForces (BoundaryCylinderRepulsion.cpp:124-189):
 * for every cylinder in a boundary plane's neighbor list (as of the last
 * neighbor-list build), the plus end loads its second bead and the minus end
 * its first, at n_mon[type] monomer positions spaced mon_size[type] ahead
 * along the cylinder: load += max(-n.dir, 0) * K exp(-r/lambda) / lambda,
 * planes in order. load_plus / load_minus: n_beads x max_mon doubles,
 * entry [bead * max_mon + i], zero where no load applies. */
int mdy_load_forces(mdy_ctx *ctx, const double *coord, int n_types, const double *mon_size, const int32_t *n_mon,
                    int max_mon, double *load_plus, double *load_minus);

/* CGMethod::minimize<method> with backtracking line search
 * (src/Mechanics/Minimizer/CGMethod.hpp:826-1743), run on the device without
 * returning to the host. coord_inout: 3*n_beads doubles (host). force_out
 * (nullable) receives the final forces (ptrOutForce). */
int mdy_minimize(mdy_ctx *ctx, const mdy_cg_params *params, double *coord_inout,
                 double *force_out, mdy_min_result *result);
/* Same, coordinates already on the device (d_coord updated in place). */
int mdy_minimize_dev(mdy_ctx *ctx, const mdy_cg_params *params, double *d_coord,
                     double *d_force_out, mdy_min_result *result);
/* Batched form: n independent minimizations (CGMethod::minimize per network,
 * as above) in one cooperative launch. ctxs[i] holds network i with its
 * neighbor lists; all contexts share one device and one precision and appear
 * once. The resident blocks (threads_per_block = 64, 128 or 256; 0 = 64) are
 * split into one cluster per network in proportion to weights[i] (nullable:
 * default the per-evaluation work, interactions + 2 x beads), at least one
 * block each; each cluster runs its own minimization with its own barrier.
 * Networks beyond the resident block count run in further launches.
 * first_pass_calls > 0 runs the batch in passes: pass k suspends each
 * minimization at the top of an iteration after first_pass_calls * 4^(k-1)
 * force calls, and the next pass resumes the suspended ones where they
 * stopped, on the blocks the finished networks released (uncapped once each
 * gets 64 blocks or more). first_pass_calls = -k (k = 1..1024) ends the
 * passes by completions instead: once max(1, m / k) of a launch's m networks
 * have finished, the others suspend at the top of their next iteration and
 * the next pass re-splits the blocks among them, down to one network on all
 * blocks. Results equal those of one uncapped pass (a
 * trajectory does not depend on its cluster size); 0 = one pass.
 * params, d_coord (device, updated in place), results: n entries each;
 * d_force_out (nullable array; entries nullable) receives final forces;
 * net_ms (nullable, n entries): each network's completion time, from the
 * start of the first launch; total_ms (nullable): the duration of the launches
 * (HIP events on the stream). stream NULL = the stream of
 * ctxs[0]. No reference counterpart: MController::run minimizes one network
 * per call (MController.h:109-211). */
int mdy_minimize_batch(mdy_ctx *const *ctxs, int n, const mdy_cg_params *params, double *const *d_coord,
                       double *const *d_force_out, mdy_min_result *results, double *net_ms,
                       int threads_per_block, const double *weights, int first_pass_calls, void *stream,
                       double *total_ms);

/* Per-iteration trace of the last minimization (lambda, energy after the line
 * search, max force, beta, safe-mode flag), for parity checks. */
int mdy_get_trace(mdy_ctx *ctx, double *lambda, double *energy, double *maxf,
                  double *beta, int32_t *safe, int64_t cap, int64_t *n);
int mdy_set_trace(mdy_ctx *ctx, int64_t cap);

/* Timing of the last minimization (device events), milliseconds. */
int mdy_last_timing(mdy_ctx *ctx, double *ms_total);
/* Phase profile of the last minimization: the device timestamp (100 MHz
 * wall clock) at the end of each grid-wide phase and the phase kind
 * (1 energy, 2 energy fallback, 3 energy and forces, 4 force fallback,
 * 5 gather, 6 reduction, 7 restore). Recording is off unless a capacity is set. */
int mdy_set_profile(mdy_ctx *ctx, int64_t cap);
int mdy_get_profile(mdy_ctx *ctx, int32_t *kind, int64_t *ticks, int64_t cap, int64_t *n);
/* Per phase 5 stamps: barrier exit, latest block arrival at the barrier,
 * block 0's work start, end of its main loop, and its arrival. */
int mdy_get_profile_detail(mdy_ctx *ctx, int64_t *rec, int64_t cap, int64_t *n);

/* ------------------------------------------------------------------------- */
/* Seeded synthetic networks (the generator behind the benchmark workloads)   */
/* ------------------------------------------------------------------------- */
typedef struct mdy_gen_params {
    uint64_t seed;
    int32_t  n_filaments;       /* total filaments, daughters included          */
    int32_t  cyl_per_filament;
    double   box;               /* cubic box edge (nm); planes inset 25 nm      */
    int32_t  n_linkers;
    int32_t  n_motors;
    int32_t  n_branches;        /* daughters are the last n_branches filaments  */
    double   cyl_size;          /* 108 nm                                       */
    double   jitter;            /* Gaussian bead jitter sigma (nm)              */
    double   min_sep;           /* rejection: min segment distance at placement */
    double   pair_cut;          /* linker/motor binding-point distance cap      */
    double   rest_sigma;        /* rest length = d0 + N(0, rest_sigma)          */
    double   k_stretch, k_bend, theta_bend;
    double   k_exvol;  float exvol_cutoff;
    double   k_bd, screen_bd; float bd_cutoff;
    double   k_linker, k_motor;
    double   k_br_stretch, l_br_stretch, k_br_bend, theta_br_bend;
    double   k_br_dihedral, k_br_position;   /* 0 = no dihedral / position tables */
} mdy_gen_params;

typedef struct mdy_network {
    mdy_system sys;
    double    *coord;        /* 3*n_beads */
    double     box;
    int32_t    n_filaments;
    int32_t   *fil_first_bead;
    int32_t   *fil_ncyl;
    void      *storage;      /* owned */
} mdy_network;

void mdy_gen_defaults(mdy_gen_params *p);
int  mdy_generate(const mdy_gen_params *p, mdy_network **out);
void mdy_network_free(mdy_network *net);

#ifdef __cplusplus
}
#endif
#endif /* MDY_H */
