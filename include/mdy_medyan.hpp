// mdy_medyan.hpp -- C++ host mirror of MEDYAN's mechanics interface over the
// C ABI in mdy.h (header-only, C++17).
//
// Names, argument meaning and error behaviour follow the reference:
//   ForceFieldManager::vectorizeAllForceFields / computeEnergy /
//     computeIndividualEnergies / computeForces / assignallforcemags /
//     computeLoadForces
//     (src/Mechanics/ForceField/ForceFieldManager.h:48-216, .cpp:35-309)
//   CGMethod::minimize (src/Mechanics/Minimizer/CGMethod.hpp:826-836)
//   ConjugateGradientParams (CGMethod.hpp:44-61)
//   MinimizationResult, EnergyReport (MinimizationTypes.hpp:10-64)
// computeEnergy returns +inf when a force field reports an invalid energy;
// a non-finite max force throws std::runtime_error("maxF error") like
// CGMethod::maxF (CGMethod.hpp:201-211); a non-finite excluded-volume force
// after its quadrature fallback throws (the reference exits); a line-search
// failure without recovery sets MinimizationResult::errorLineSearch.
// Everything runs on the HIP device of the context; there is no CPU path.
#pragma once

#include <cmath>
#include <cstdint>
#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "mdy.h"

namespace mdy_medyan {

enum class ConjugateGradientDescentSearch { steepest = MDY_CG_STEEPEST, fletcherRieves = MDY_CG_FLETCHER_RIEVES,
                                            polakRibiere = MDY_CG_POLAK_RIBIERE };

struct ConjugateGradientParams {
    ConjugateGradientDescentSearch descentSearchMethod = ConjugateGradientDescentSearch::polakRibiere;
    double maxDist = 1.0;
    double lambdaMax = 0.01;
    double gradTol = 1.0;
    double energyRelTol = 1e-9;
    int energyCallLimit = 0;
    int forceCallLimit = 10000;
    bool tryToRecoverInLineSearchError = true;
    // "BACKTRACKING" or "QUADRATIC"; CGMethod::minimize (CGMethod.hpp:1319-1331)
    // takes any other name as QUADRATIC
    std::string lineSearchAlgorithm = "BACKTRACKING";
    // restore the lowest-energy backup when the call limit ends the run
    // (CGMethod.hpp:59, 1589-1608)
    bool reportOnCallLimit = true;
};

struct EnergyReport {
    struct Each { const char *name; double energy; };
    double total = 0;
    std::vector<Each> individual;
};

struct MinimizationResult {
    static constexpr std::uint_fast32_t errorLineSearch = 1u << 0;
    EnergyReport energiesBefore, energiesAfter;
    std::uint_fast32_t errorFlags = 0;
    int numEnergyCall = 0, energyCallLimit = 0, numForceCall = 0, forceCallLimit = 0;
    int numCGMethodCall = 0;
    double gradInfNorm = INFINITY, gradTol = 0, energyRelChange = INFINITY, energyRelTol = 0;
    int numIter = 0;   // CG iterations (not part of the reference struct)

    bool success() const { return errorFlags == 0; }
    bool callLimitExceeded() const {
        return (energyCallLimit != 0 && numEnergyCall > energyCallLimit) ||
               (forceCallLimit != 0 && numForceCall > forceCallLimit);
    }
    bool gradInfNormConverged() const { return gradInfNorm < gradTol; }
    bool energyRelChangeConverged() const { return energyRelChange < energyRelTol; }
    bool converged() const { return gradInfNormConverged() || energyRelChangeConverged(); }
};

inline const char *const kForceFieldNames[MDY_NFF] = {
    "FilamentStretching", "FilamentBending", "LinkerStretching", "MotorGhostStretching",
    "BranchingStretching", "BranchingBending", "BranchingDihedral", "BranchingPosition",
    "CylinderExclVolume", "BoundaryCylinderRepulsion"};

class Error : public std::runtime_error {
public:
    Error(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }
private:
    int code_;
};

// One device context (mdy_ctx): device-resident tables, neighbor lists and
// minimizer state. Not copyable; one caller thread at a time.
class ForceFieldManager {
public:
    explicit ForceFieldManager(int device = 0, const mdy_opts *opts = nullptr) {
        int rc = mdy_create(&ctx_, device, opts);
        if (rc != MDY_OK) throw Error(rc, "mdy_create failed (no HIP device?)");
    }
    ~ForceFieldManager() { mdy_destroy(ctx_); }
    ForceFieldManager(const ForceFieldManager &) = delete;
    ForceFieldManager &operator=(const ForceFieldManager &) = delete;

    // Upload the tables and rebuild the neighbor lists from `coord`
    // (vectorizeAllForceFields + the neighbor-list reset of MController::run).
    void vectorizeAllForceFields(const mdy_system &sys, const double *coord) {
        check(mdy_set_network(ctx_, &sys));
        check(mdy_build_neighbors(ctx_, coord));
        numDof_ = 3 * sys.n_beads;
    }
    // Neighbor lists only (the tables are unchanged).
    void updateNeighborLists(const double *coord) { check(mdy_build_neighbors(ctx_, coord)); }

    // verbose prints each force field's energy; an invalid energy (-1 from a
    // force field) makes the total +inf and names the culprit, as
    // ForceFieldManager::computeEnergy does (ForceFieldManager.cpp:137-162)
    double computeEnergy(const double *coord, bool verbose = false) const {
        double e = 0, per[MDY_NFF];
        check(mdy_energy(ctx_, coord, &e, per));
        for (int k = 0; k < MDY_NFF; k++) {
            if (verbose) std::cout << kForceFieldNames[k] << " energy = " << per[k] << std::endl;
            if (per[k] <= -1) {
                std::cout << "Energy = " << per[k] << std::endl;
                std::cout << "Energy of system became infinite. Try adjusting minimization parameters." << std::endl;
                std::cout << "The culprit was ... " << kForceFieldNames[k] << std::endl;
                break;
            }
        }
        return e;
    }
    EnergyReport computeIndividualEnergies(const double *coord) const {
        double tot = 0, per[MDY_NFF];
        check(mdy_energy(ctx_, coord, &tot, per));
        EnergyReport r;
        for (int k = 0; k < MDY_NFF; k++) {
            r.individual.push_back({kForceFieldNames[k], per[k]});
            r.total += per[k];
        }
        return r;
    }
    // Zero-fills force[0, numVar), then every force field adds its forces.
    void computeForces(const double *coord, double *force, int numVar) const {
        for (int i = 0; i < numVar; i++) force[i] = 0.0;
        check(mdy_forces(ctx_, coord, force));
    }

    // assignallforcemags (MController::afterMinimization_): stretch forces of
    // linkers and motors, force vectors of branch points, at coord.
    struct ForceMags {
        std::vector<double> linker, motor, branch;   // branch: 3 per branch point
    };
    ForceMags assignallforcemags(const double *coord, int nLinkers, int nMotors, int nBranches) const {
        ForceMags m;
        m.linker.assign(nLinkers, 0.0);
        m.motor.assign(nMotors, 0.0);
        m.branch.assign(3 * (size_t)nBranches, 0.0);
        check(mdy_stretch_forces(ctx_, coord, m.linker.data(), m.motor.data(), m.branch.data()));
        return m;
    }
    // computeLoadForces: per bead, the boundary load on monomer positions
    // ahead of a plus end (plus) or minus end (minus), [bead * maxMon + i].
    void computeLoadForces(const double *coord, const std::vector<double> &monSize, const std::vector<int32_t> &numMon,
                           std::vector<double> &plus, std::vector<double> &minus) const {
        int maxMon = 0;
        for (int32_t v : numMon) maxMon = v > maxMon ? v : maxMon;
        plus.assign((size_t)(numDof_ / 3) * maxMon, 0.0);
        minus.assign(plus.size(), 0.0);
        if (maxMon == 0) return;
        check(mdy_load_forces(ctx_, coord, (int)monSize.size(), monSize.data(), numMon.data(), maxMon, plus.data(),
                              minus.data()));
    }

    mdy_ctx *ctx() const { return ctx_; }
    int numDof() const { return numDof_; }

    void check(int rc) const {
        if (rc == MDY_OK) return;
        if (rc == MDY_ERR_NONFINITE) throw Error(rc, "maxF error");
        throw Error(rc, std::string("mdy: ") + mdy_last_error(ctx_));
    }

private:
    mdy_ctx *ctx_ = nullptr;
    int numDof_ = 0;
};

class CGMethod {
public:
    // Minimizes coord[0, numDof) in place with the manager's force fields;
    // outForce (nullable) receives the final forces.
    // CGMethod::searchDir and LAMBDATOL after minimize, which MController::run
    // hands to its line-search diagnosis (MController.h:142-152)
    std::vector<double> searchDir;
    double LAMBDATOL = MDY_LAMBDATOL;

    MinimizationResult minimize(const ConjugateGradientParams &p, std::vector<double> &coord, int numDof,
                                ForceFieldManager &ffm, std::vector<double> *outForce = nullptr) {
        if ((int)coord.size() < numDof || numDof != ffm.numDof())
            throw Error(MDY_ERR_ARG, "coordinate vector does not match the network");
        mdy_cg_params cp{};
        cp.method = (int32_t)p.descentSearchMethod;
        cp.max_dist = p.maxDist;
        cp.lambda_max = p.lambdaMax;
        cp.grad_tol = p.gradTol;
        cp.energy_rel_tol = p.energyRelTol;
        cp.energy_call_limit = p.energyCallLimit;
        cp.force_call_limit = p.forceCallLimit;
        cp.n_beads_limit = numDof / 3;
        cp.try_to_recover = p.tryToRecoverInLineSearchError ? 1 : 0;
        cp.line_search = p.lineSearchAlgorithm == "BACKTRACKING" ? MDY_LS_BACKTRACKING : MDY_LS_QUADRATIC;
        cp.report_on_call_limit = p.reportOnCallLimit ? 1 : 0;
        if (outForce) outForce->assign(numDof, 0.0);
        mdy_min_result r{};
        ffm.check(mdy_minimize(ffm.ctx(), &cp, coord.data(), outForce ? outForce->data() : nullptr, &r));
        MinimizationResult m;
        for (int k = 0; k < MDY_NFF; k++) {
            m.energiesBefore.individual.push_back({kForceFieldNames[k], r.energies_before[k]});
            m.energiesAfter.individual.push_back({kForceFieldNames[k], r.energies_after[k]});
        }
        m.energiesBefore.total = r.energies_before[MDY_NFF];
        m.energiesAfter.total = r.energies_after[MDY_NFF];
        m.errorFlags = r.error_flags;
        m.numEnergyCall = r.num_energy_call;
        m.energyCallLimit = r.energy_call_limit;
        m.numForceCall = r.num_force_call;
        m.forceCallLimit = r.force_call_limit;
        m.numCGMethodCall = 1;
        m.gradInfNorm = r.grad_inf_norm;
        m.gradTol = r.grad_tol;
        m.energyRelChange = r.energy_rel_change;
        m.energyRelTol = r.energy_rel_tol;
        m.numIter = r.num_iter;
        searchDir.assign(numDof, 0.0);
        ffm.check(mdy_get_search_dir(ffm.ctx(), searchDir.data()));
        return m;
    }
};

}  // namespace mdy_medyan
