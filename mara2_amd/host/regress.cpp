// regress.cpp -- the reference's regression programs (src/Problems.cpp:35-671)
// written against the host layer, running every step on the GPU.
//
//   mara_regress <program> <problem> <scheme> <out.bin> [max_steps
//                [[viscous=]code.hsaco kernel [parameters...]]]
//   program: regress-1d | regress-rel | regress-2d | regress-mhd | sod
//   The optional code object (include/mara_source.h) is a source-term
//   function set with GpuMethodOfLinesTVD::setSourceTermsCodeObject, or with
//   the viscous= prefix a viscous-flux function set with
//   GpuMethodOfLinesTVD::setViscousFluxCodeObject.
//
// Writes the initial primitives (after the ghost fill), the final ones and
// the loop's clock to out.bin, so a test can replay the same initial data on
// the oracle:  int32 header[8] = {N1, N2, N3, nq, steps, 0, 0, 0},
// float64 t_final_reached, float64 P0[...], float64 P[...].
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "mara2.hpp"

using namespace mara2;

namespace {

struct Problem
{
    double finalTime;
    std::shared_ptr<BoundaryCondition> bc;
    InitialDataFunction idf;
    InitialDataFunction ivp = nullptr;  // vector potential, when B comes from one
};

std::vector<double> pick(bool left, std::vector<double> a, std::vector<double> b) { return left ? a : b; }

// src/Problems.cpp:35-89
std::map<std::string, Problem> problems1d()
{
    auto periodic = std::make_shared<PeriodicBoundaryCondition>();
    auto outflow = std::make_shared<OutflowBoundaryCondition>();
    std::map<std::string, Problem> p;
    p["Shocktube1"] = {0.100, outflow, [](double x, double, double) { return pick(x < 0.5, {1.000, 0.000, 0.0, 0.0, 1.000}, {0.125, 0.000, 0.0, 0.0, 0.100}); }};
    p["Shocktube2"] = {0.100, outflow, [](double x, double, double) { return pick(x < 0.5, {1.000, -2.000, 0.0, 0.0, 0.400}, {1.000, 2.000, 0.0, 0.0, 0.400}); }};
    p["Shocktube3"] = {0.005, outflow, [](double x, double, double) { return pick(x < 0.5, {1.0, 0.0, 0.0, 0.0, 1e+3}, {1.0, 0.0, 0.0, 0.0, 1e-2}); }};
    p["Shocktube4"] = {0.01, outflow, [](double x, double, double) { return pick(x < 0.5, {1.0, 0.0, 0.0, 0.0, 1e-2}, {1.0, 0.0, 0.0, 0.0, 1e+2}); }};
    p["Shocktube5"] = {0.01, outflow, [](double x, double, double) { return pick(x < 0.5, {5.99924, 19.59750, 0.0, 0.0, 460.894}, {5.99924, -6.19633, 0.0, 0.0, 46.095}); }};
    p["ContactWave"] = {0.1, outflow, [](double x, double, double) { return pick(x < 0.5, {1.0, 0.0, 0.7, 0.2, 1.0}, {0.1, 0.0, 0.7, 0.2, 1.0}); }};
    p["DensityWave"] = {1.0, periodic, [](double x, double, double) { return std::vector<double>{1.0 + 0.1 * std::sin(4 * M_PI * x), 1.0, 0.0, 0.0, 1.0}; }};
    return p;
}

// src/Problems.cpp:224-237
std::map<std::string, Problem> problemsRel()
{
    auto outflow = std::make_shared<OutflowBoundaryCondition>();
    std::map<std::string, Problem> p;
    p["RelativisticShocktube1"] = {0.15, outflow, [](double x, double, double) { return pick(x < 0.5, {1.000, 0.0, 0.0, 0.0, 1.0, 0, 0, 0}, {0.125, 0.0, 0.0, 0.0, 0.1, 0, 0, 0}); }};
    return p;
}

// src/Problems.cpp:350-376
std::map<std::string, Problem> problems2d()
{
    auto periodic = std::make_shared<PeriodicBoundaryCondition>();
    auto outflow = std::make_shared<OutflowBoundaryCondition>();
    std::map<std::string, Problem> p;
    p["Shocktube2D"] = {0.100, outflow, [](double x, double y, double) { return pick(x + y < 1., {1.000, 0.000, 0.0, 0.0, 1.000}, {0.125, 0.000, 0.0, 0.0, 0.100}); }};
    p["ContactWave2D"] = {0.1, periodic, [](double x, double y, double) { return pick(x + y < 1., {1.0, 0.0, 0.0, 0.0, 1.0}, {0.1, 0.0, 0.0, 0.0, 1.0}); }};
    p["DensityWave2D"] = {1.0, periodic, [](double x, double y, double) { return std::vector<double>{1.0 + 0.1 * std::sin(4 * M_PI * (x + y)), 1.0, 1.0, 0.0, 1.0}; }};
    return p;
}

// src/Problems.cpp:489-552; FieldLoop and AbcField take B from a vector
// potential on faces (runProblem, src/Problems.cpp:660-667)
std::map<std::string, Problem> problemsMhd()
{
    auto periodic = std::make_shared<PeriodicBoundaryCondition>();
    std::map<std::string, Problem> p;
    p["CylindricalBlast"] = {0.1, periodic, [](double x, double y, double) {
        x -= 0.5;
        y -= 0.5;
        const double R = std::sqrt(x * x + y * y);
        return pick(R < 0.125, {1.000, 0., 0., 0., 1.000, 1.33, 1.33, 0.}, {0.125, 0., 0., 0., 0.100, 1.33, 1.33, 0.});
    }};
    // a weak loop of B_phi = 3 B0 k (kR)^2 exp(-(kR)^3) advected diagonally;
    // A_z = B0 exp(-(kR)^3)
    const double B0 = 1e-5, kl = 10.0;
    p["FieldLoop"] = {1.0, periodic,
        [=](double x, double y, double) {
            x -= 0.5;
            y -= 0.5;
            const double R = std::sqrt(x * x + y * y);
            const double bphi = B0 * 3 * kl * std::pow(kl * R, 2) * std::exp(-std::pow(kl * R, 3));
            return std::vector<double>{1.0, 1.0, 1.0, 0.0, 1.0, bphi * (-y / R), bphi * (x / R), 0.0};
        },
        [=](double x, double y, double) {
            x -= 0.5;
            y -= 0.5;
            const double R = std::sqrt(x * x + y * y);
            return std::vector<double>{0.0, 0.0, B0 * std::exp(-std::pow(kl * R, 3))};
        }};
    // the Arnold-Beltrami-Childress field with (A, B, C) = (1, 1, 0), k = 2 pi,
    // and its vector potential B / k
    auto abc = [](double x, double y, double z) {
        const double A = 1.0, B = 1.0, C = 0.0, k = 2 * M_PI;
        return std::array<double, 3>{C * std::cos(k * z) - B * std::sin(k * y), A * std::cos(k * x) - C * std::sin(k * z),
                                     B * std::cos(k * y) - A * std::sin(k * x)};
    };
    p["AbcField"] = {0.1, periodic,
        [=](double x, double y, double z) {
            const auto b = abc(x, y, z);
            return std::vector<double>{1., 0., 0., 0., 1., b[0], b[1], b[2]};
        },
        [=](double x, double y, double z) {
            const auto b = abc(x, y, z);
            const double k = 2 * M_PI;
            return std::vector<double>{b[0] / k, b[1] / k, b[2] / k};
        }};
    return p;
}

// The ThermalConvection program (src/Problems/ThermalConvection.cpp:176-300)
// with its default user values: a polytrope rho = r^-1.5, p = K rho^gamma
// (K = r0 g0 (gamma - 1) / gamma) at rest on a spherical wedge, r in
// [1, 100] with logarithmic edges, theta within pi/12 of the equator, phi in
// [0, 0.1].  Its gravity, heating and geometric source terms come from the
// code object given on the command line (mara2_amd/sources/
// thermal_convection.hip).  The program's acoustic-driver boundary is not
// part of this path: outflow stands in for it.
std::map<std::string, Problem> problemsSpherical()
{
    std::map<std::string, Problem> p;
    const double g0 = 0.2, d0 = 1.0, gamma = 5. / 3, gm = 5. / 3, r0 = 1.0;
    const double K = r0 * g0 * (gm - 1) / gm / std::pow(d0, gm - 1);
    p["ThermalConvection"] = {0.5, std::make_shared<OutflowBoundaryCondition>(), [=](double r, double, double) {
        const double d = d0 * std::pow(r, -1.5);
        return std::vector<double>{d, 0, 0, 0, K * std::pow(d, gamma)};
    }};
    return p;
}

int usage()
{
    std::cerr << "usage: mara_regress <regress-1d|regress-rel|regress-2d|regress-mhd|sod|thermal-convection> <problem> <scheme> <out.bin>"
                 " [max_steps [[viscous=]code.hsaco kernel [parameters...]]]\n";
    return 2;
}

} // namespace

int main(int argc, const char* argv[])
{
    if (argc < 5) return usage();
    const std::string program = argv[1], pname = argv[2], sname = argv[3], out = argv[4];
    const int maxSteps = argc > 5 ? std::atoi(argv[5]) : 1 << 30;
    // optional source terms: <code object> <kernel> [parameters...], or a
    // viscous-flux function: viscous=<code object> <kernel> [parameters...]
    std::string srcPath, srcKernel;
    std::vector<double> srcParams;
    bool viscous = false;
    if (argc > 7)
    {
        srcPath = argv[6];
        if (srcPath.rfind("viscous=", 0) == 0)
        {
            viscous = true;
            srcPath = srcPath.substr(8);
        }
        srcKernel = argv[7];
        for (int a = 8; a < argc; ++a) srcParams.push_back(std::atof(argv[a]));
    }

    // scheme: pcmN / plmN with N the Runge-Kutta order (src/Problems.cpp:102-135)
    if (sname.size() != 4) return usage();
    const bool plm = sname.compare(0, 3, "plm") == 0;
    const int order = sname[3] - '0';

    Shape cs{{128, 1, 1, 1, 1}}, bs{{2, 0, 0, 0, 0}};
    std::shared_ptr<ConservationLaw> cl;
    std::map<std::string, Problem> table;
    double cfl = 0.5;
    auto ss = std::make_shared<GpuMethodOfLinesTVD>();
    std::shared_ptr<IntercellFluxScheme> fs = plm ? std::shared_ptr<IntercellFluxScheme>(std::make_shared<MethodOfLinesPlm>())
                                                  : std::make_shared<MethodOfLines>();
    if (program == "regress-1d")
    {
        cl = std::make_shared<NewtonianHydro>();
        table = problems1d();
        // runProblem overrides the scheme's flux scheme with PLM + HLLC
        // (src/Problems.cpp:162-174)
        fs = std::make_shared<MethodOfLinesPlm>();
        fs->setRiemannSolver(std::make_shared<HllcNewtonianHydroRiemannSolver>());
    }
    else if (program == "regress-rel")
    {
        cl = std::make_shared<RelativisticMHD>();
        table = problemsRel();
        cfl = 0.25;
    }
    else if (program == "regress-2d")
    {
        cs = {{128, 128, 1, 1, 1}};
        bs = {{2, 2, 0, 0, 0}};
        cl = std::make_shared<NewtonianHydro>();
        table = problems2d();
    }
    else if (program == "regress-mhd")
    {
        cs = {{64, 64, 1, 1, 1}};
        bs = {{2, 2, 0, 0, 0}};
        cl = std::make_shared<NewtonianMHD>();
        table = problemsMhd();
        ss->setDisableFieldCT(false);
    }
    else if (program == "sod")
    {
        // 1024 zones, gamma 1.4, PLM + HLLE, CFL 0.4, t = 0.2
        cs = {{1024, 1, 1, 1, 1}};
        cl = std::make_shared<NewtonianHydro>();
        cl->setGammaLawIndex(1.4);
        auto outflow = std::make_shared<OutflowBoundaryCondition>();
        table["Sod"] = {0.2, outflow, [](double x, double, double) { return pick(x < 0.5, {1.0, 0.0, 0.0, 0.0, 1.0}, {0.125, 0.0, 0.0, 0.0, 0.1}); }};
        cfl = 0.4;
    }
    else if (program == "thermal-convection")
    {
        cs = {{64, 64, 1, 1, 1}};
        bs = {{2, 2, 0, 0, 0}};
        cl = std::make_shared<NewtonianHydro>();
        table = problemsSpherical();
        fs = std::make_shared<MethodOfLinesPlm>();
        fs->setPlmTheta(2.0);
        fs->setRiemannSolver(std::make_shared<HllcNewtonianHydroRiemannSolver>());
    }
    else return usage();
    if (!table.count(pname)) return usage();
    const Problem& problem = table[pname];

    std::shared_ptr<MeshGeometry> mg;
    if (program == "thermal-convection")
    {
        auto sg = std::make_shared<SphericalMeshGeometry>();
        sg->setLowerUpper({{1.0, M_PI * 0.5 - M_PI / 12.0, 0}}, {{100.0, M_PI * 0.5 + M_PI / 12.0, 0.1}});
        sg->setUseLogarithmicRadialBinning(true);
        mg = sg;
    }
    else mg = std::make_shared<CartesianMeshGeometry>();
    auto mo = std::make_shared<MeshOperator>();
    auto fo = std::make_shared<FieldOperator>();
    auto md = std::make_shared<MeshData>(cs, bs, cl->getNumConserved());
    mg->setCellsShape(cs);
    mo->setMeshGeometry(mg);
    fo->setConservationLaw(cl);
    fo->setMeshGeometry(mg);
    ss->setRungeKuttaOrder(order);
    ss->setFieldOperator(fo);
    ss->setMeshOperator(mo);
    ss->setBoundaryCondition(problem.bc);
    ss->setIntercellFluxScheme(fs);
    if (!srcPath.empty() && viscous) ss->setViscousFluxCodeObject(srcPath, srcKernel, srcParams);
    else if (!srcPath.empty()) ss->setSourceTermsCodeObject(srcPath, srcKernel, srcParams);
    problem.bc->setConservationLaw(cl);

    SimulationStatus status;
    status.totalCellsInMesh = (long)mg->totalCellsInMesh();
    auto L = mo->linearCellDimension();
    md->assignPrimitive(mo->generate(problem.idf, MeshLocation::cell));
    if (problem.ivp)
    {
        // B = div(CT(F(A))) over the whole array, guards included
        CellCenteredFieldCT ct;
        md->setMagneticIndex(cl->getIndexFor(ConservationLaw::VariableType::magnetic));
        const Array A = mo->generate(problem.ivp, MeshLocation::face, MeshOperator::VectorMode::scalars, bs);
        const Array G = ct.generateGodunovFluxes(ct.vectorPotentialToFluxes(A), 0);
        md->assignMagneticField(mo->divergence(G, 1.0), MeshLocation::cell, MeshData::includeGuard);
    }
    md->applyBoundaryCondition(*problem.bc);
    const Array P0 = md->P;

    auto timestep = [&]() {
        const double dt1 = cfl * fo->getCourantTimestep(*md, L);
        const double dt2 = problem.finalTime - status.simulationTime;
        return dt1 < dt2 ? dt1 : dt2;
    };
    auto condition = [&]() { return status.simulationTime < problem.finalTime && status.simulationIter < maxSteps; };
    auto advance = [&](double dt) { ss->advance(*md, 0.0, dt); };
    double kzps = 0.0;
    try
    {
        maraMainLoop(status, timestep, condition, advance, &kzps);
    }
    catch (const std::exception& e)
    {
        std::cerr << "mara_regress: " << e.what() << "\n";
        return 1;
    }

    FILE* f = std::fopen(out.c_str(), "wb");
    if (!f) return 1;
    const int header[8] = {md->P.size(0), md->P.size(1), md->P.size(2), md->P.size(3), status.simulationIter, 0, 0, 0};
    std::fwrite(header, sizeof(int), 8, f);
    std::fwrite(&status.simulationTime, sizeof(double), 1, f);
    std::fwrite(P0.data(), sizeof(double), P0.size(), f);
    std::fwrite(md->P.data(), sizeof(double), md->P.size(), f);
    std::fclose(f);
    std::printf("%s %s %s: %d steps to t=%.6f, mean kzps %.1f\n", program.c_str(), pname.c_str(), sname.c_str(),
                status.simulationIter, status.simulationTime, kzps);
    return 0;
}
