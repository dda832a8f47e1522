// Drives the C++ mirror of MEDYAN's mechanics interface (include/mdy_medyan.hpp)
// the way MController::run drives the reference (Controller/MController.h:109-211):
// vectorize, energies, forces, minimize. Needs a HIP device.
// Built as a shared object whose mdy_shim_run() returns 0 when every check
// holds (a failed check returns its line number; an escaped exception -1),
// so it runs inside the test process: a second process on the same GPU
// could not make the cooperative launches the minimizer uses while the test
// process holds the device's cooperative-launch resource.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "mdy_medyan.hpp"

#define REQUIRE(c)                                                              \
    do {                                                                        \
        if (!(c)) {                                                             \
            std::fprintf(stderr, "FAILED %s:%d: %s\n", __FILE__, __LINE__, #c); \
            return __LINE__;                                                    \
        }                                                                       \
    } while (0)

using namespace mdy_medyan;

static int run() {
    mdy_gen_params gp;
    mdy_gen_defaults(&gp);
    gp.n_filaments = 20;
    gp.n_linkers = 40;
    gp.n_motors = 5;
    gp.seed = 7;
    mdy_network *net = nullptr;
    REQUIRE(mdy_generate(&gp, &net) == MDY_OK);
    const int ndof = 3 * net->sys.n_beads;
    std::vector<double> coord(net->coord, net->coord + ndof);

    ForceFieldManager ffm(0);
    ffm.vectorizeAllForceFields(net->sys, coord.data());

    // computeEnergy == sum of computeIndividualEnergies for a valid configuration
    const double e = ffm.computeEnergy(coord.data());
    EnergyReport rep = ffm.computeIndividualEnergies(coord.data());
    REQUIRE(rep.individual.size() == MDY_NFF);
    REQUIRE(std::fabs(e - rep.total) <= 1e-12 * std::fabs(e));

    // computeForces zero-fills, so a dirty buffer gives the same forces
    std::vector<double> f1(ndof, 0.0), f2(ndof, 123.0);
    ffm.computeForces(coord.data(), f1.data(), ndof);
    ffm.computeForces(coord.data(), f2.data(), ndof);
    REQUIRE(f1 == f2);

    // minimize to the benchmark tolerance
    const std::vector<double> start = coord;
    ConjugateGradientParams p;
    p.gradTol = 1.0;
    p.maxDist = 1.0;
    p.lambdaMax = 0.01;
    p.forceCallLimit = 100000;
    std::vector<double> out;
    MinimizationResult r = CGMethod{}.minimize(p, coord, ndof, ffm, &out);
    REQUIRE(r.success());
    REQUIRE(r.converged());
    REQUIRE(r.energiesAfter.total < r.energiesBefore.total);
    double mf = 0;
    for (double v : out) mf = std::max(mf, std::fabs(v));
    REQUIRE(mf == r.gradInfNorm);
    REQUIRE(std::fabs(ffm.computeEnergy(coord.data()) - r.energiesAfter.total) <= 1e-12 * std::fabs(r.energiesAfter.total));

    // post-minimization outputs
    auto mags = ffm.assignallforcemags(coord.data(), net->sys.ff.ls_n, net->sys.ff.ms_n, net->sys.ff.bs_n);
    REQUIRE((int)mags.linker.size() == net->sys.ff.ls_n);
    std::vector<double> lp, lm;
    ffm.computeLoadForces(coord.data(), {2.7}, {40}, lp, lm);
    REQUIRE(lp.size() == (size_t)(ndof / 3) * 40);

    // misuse is reported like the reference reports it: by exception
    bool threw = false;
    try {
        std::vector<double> wrong(ndof - 3);
        CGMethod{}.minimize(p, wrong, ndof - 3, ffm);
    } catch (const Error &err) {
        threw = err.code() == MDY_ERR_ARG;
    }
    REQUIRE(threw);

    // the quadratic line search from the same start
    std::vector<double> cq = start;
    p.lineSearchAlgorithm = "QUADRATIC";
    CGMethod cg;
    MinimizationResult rq = cg.minimize(p, cq, ndof, ffm);
    REQUIRE(rq.success());
    REQUIRE((int)cg.searchDir.size() == ndof && cg.LAMBDATOL == 1e-8);
    double sdotf = 0;
    std::vector<double> fq(ndof);
    ffm.computeForces(cq.data(), fq.data(), ndof);
    for (int i = 0; i < ndof; i++) sdotf += cg.searchDir[i] * fq[i];
    REQUIRE(sdotf > 0);   // a descent direction at the returned coordinates
    REQUIRE(rq.converged());
    REQUIRE(rq.energiesAfter.total < rq.energiesBefore.total);

    mdy_network_free(net);
    std::printf("ALL OK: %d CG iterations, max|F| %.4f pN\n", r.numIter, r.gradInfNorm);
    return 0;
}

extern "C" int mdy_shim_run() {
    try {
        return run();
    } catch (const std::exception &e) {
        std::fprintf(stderr, "FAILED: exception %s\n", e.what());
        return -1;
    }
}
