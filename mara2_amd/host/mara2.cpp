// mara2.cpp -- the host layer of mara2.hpp.  Only descriptor mapping, host
// arrays and the ghost fill of initial data live here; every step runs in
// the HIP library through the C ABI.
#include "mara2.hpp"

#include <chrono>
#include <cmath>
#include <cstring>

namespace mara2 {

// ---------------------------------------------------------------------------
Array::Array(Shape s) : shape_(s)
{
    std::size_t n = 1;
    for (int a = 0; a < 5; ++a) n *= (std::size_t)(shape_[a] < 1 ? 1 : shape_[a]);
    for (int a = 0; a < 5; ++a) if (shape_[a] < 1) shape_[a] = 1;
    data_.assign(n, 0.0);
}

Array::Array(int n1, int n2, int n3, int n4, int n5) : Array(Shape{{n1, n2, n3, n4, n5}}) {}

double& Array::operator()(int i, int j, int k, int q, int r)
{
    return data_[((((std::size_t)i * shape_[1] + j) * shape_[2] + k) * shape_[3] + q) * shape_[4] + r];
}

double Array::operator()(int i, int j, int k, int q, int r) const
{
    return data_[((((std::size_t)i * shape_[1] + j) * shape_[2] + k) * shape_[3] + q) * shape_[4] + r];
}

// ---------------------------------------------------------------------------
std::vector<std::string> ConservationLaw::getPrimitiveNames() const
{
    std::vector<std::string> names;
    for (int q = 0; q < getNumConserved(); ++q) names.push_back(getPrimitiveName(q));
    return names;
}

int NewtonianHydro::getIndexFor(VariableType type) const
{
    switch (type)
    {
        case VariableType::density: return 0;
        case VariableType::velocity: return 1;
        case VariableType::pressure: return 4;
        default: return -1;
    }
}

std::string NewtonianHydro::getPrimitiveName(int q) const
{
    static const char* names[] = {"density", "velocity1", "velocity2", "velocity3", "pressure"};
    return q >= 0 && q < 5 ? names[q] : "";
}

int NewtonianMHD::getIndexFor(VariableType type) const
{
    switch (type)
    {
        case VariableType::density: return 0;
        case VariableType::velocity: return 1;
        case VariableType::pressure: return 4;
        case VariableType::magnetic: return 5;
        default: return -1;
    }
}

std::string NewtonianMHD::getPrimitiveName(int q) const
{
    static const char* names[] = {"density", "velocity1", "velocity2", "velocity3", "pressure",
                                  "magnetic1", "magnetic2", "magnetic3"};
    if (q >= 0 && q < 8) return names[q];
    return "auxiliary" + std::to_string(q - 7);  // NewtonianMHD::getPrimitiveName
}

// ---------------------------------------------------------------------------
CartesianMeshGeometry::CartesianMeshGeometry() : shape{{128, 1, 1, 1, 1}}, lower{{0, 0, 0}}, upper{{1, 1, 1}}
{
    cacheSpacing();
}

CartesianMeshGeometry::CartesianMeshGeometry(Shape S) : shape{{S[0], S[1], S[2], 1, 1}}, lower{{0, 0, 0}}, upper{{1, 1, 1}}
{
    cacheSpacing();
}

void CartesianMeshGeometry::setCellsShape(Shape S)
{
    shape[0] = S[0];
    shape[1] = S[1];
    shape[2] = S[2];
    cacheSpacing();
}

void CartesianMeshGeometry::setLowerUpper(Coordinate L, Coordinate U)
{
    lower = L;
    upper = U;
    cacheSpacing();
}

Coordinate CartesianMeshGeometry::coordinateAtIndex(double i, double j, double k) const
{
    return Coordinate{{lower[0] + dx[0] * (i + 0.5), lower[1] + dx[1] * (j + 0.5), lower[2] + dx[2] * (k + 0.5)}};
}

void MeshGeometry::describe(mara_desc& d) const
{
    const Shape cs = cellsShape();
    const Coordinate lo = lowerCorner(), hi = upperCorner();
    for (int a = 0; a < 3; ++a)
    {
        d.cells[a] = cs[a];
        d.lower[a] = lo[a];
        d.upper[a] = hi[a];
    }
    d.geometry = geometryCode();
    d.log_radial = logRadial();
}

// ---------------------------------------------------------------------------
SphericalMeshGeometry::SphericalMeshGeometry() : shape{{128, 1, 1, 1, 1}}, lower{{1.0, 0.0, 0.0}}, upper{{10.0, M_PI, 2 * M_PI}} {}

SphericalMeshGeometry::SphericalMeshGeometry(Shape S)
    : shape{{S[0], S[1], S[2], 1, 1}}, lower{{1.0, 0.0, 0.0}}, upper{{10.0, M_PI, 2 * M_PI}}
{
}

SphericalMeshGeometry::SphericalMeshGeometry(int ni, int nj, int nk)
    : shape{{ni, nj, nk, 1, 1}}, lower{{1.0, 0.0, 0.0}}, upper{{10.0, M_PI, 2 * M_PI}}
{
}

void SphericalMeshGeometry::setCellsShape(Shape S)
{
    shape[0] = S[0];
    shape[1] = S[1];
    shape[2] = S[2];
}

// edge n of an axis, fractional n for centres (:190-205)
double SphericalMeshGeometry::getEdge(double n, int axis) const
{
    if (axis == 0 && logarithmic) return lower[0] * std::pow(upper[0] / lower[0], n / shape[0]);
    return lower[axis] + (upper[axis] - lower[axis]) * n / shape[axis];
}

Coordinate SphericalMeshGeometry::coordinateAtIndex(double i, double j, double k) const
{
    return Coordinate{{getEdge(i + 0.5, 0), getEdge(j + 0.5, 1), getEdge(k + 0.5, 2)}};
}

unsigned long SphericalMeshGeometry::totalCellsInMesh() const
{
    return (unsigned long)shape[0] * shape[1] * shape[2];
}

// (:77-91); with linear edges the reference averages radial edge i with
// theta edge i + 1, which the GPU step reproduces
double SphericalMeshGeometry::cellLength(int i, int j, int k, int axis) const
{
    const double ri = logarithmic ? std::sqrt(getEdge(i, 0) * getEdge(i + 1, 0)) : 0.5 * (getEdge(i, 0) + getEdge(i + 1, 1));
    const double qi = 0.5 * (getEdge(j, 1) + getEdge(j + 1, 1));
    if (axis == 0) return 1. * (getEdge(i + 1, 0) - getEdge(i, 0));
    if (axis == 1) return ri * (getEdge(j + 1, 1) - getEdge(j, 1));
    if (axis == 2) return ri * (getEdge(k + 1, 2) - getEdge(k, 2)) * std::sin(qi);
    throw std::logic_error("Axis argument not 0, 1, or 2");
}

// (:106-123): the face on the low side of cell (i, j, k)
double SphericalMeshGeometry::faceArea(int i, int j, int k, int axis) const
{
    const double r0 = getEdge(i, 0), r1 = getEdge(i + 1, 0);
    const double q0 = getEdge(j, 1), q1 = getEdge(j + 1, 1);
    const double p0 = getEdge(k, 2), p1 = getEdge(k + 1, 2);
    if (axis == 0) return -r0 * r0 * (p1 - p0) * (std::cos(q1) - std::cos(q0));
    if (axis == 1) return +1. / 3 * (r1 * r1 * r1 - r0 * r0 * r0) * std::sin(q0) * (p1 - p0);
    if (axis == 2) return -1. / 3 * (r1 * r1 * r1 - r0 * r0 * r0) * (std::cos(q1) - std::cos(q0));
    throw std::logic_error("Invalid axis");
}

double SphericalMeshGeometry::cellVolume(int i, int j, int k) const
{
    const double r0 = getEdge(i, 0), r1 = getEdge(i + 1, 0);
    const double q0 = getEdge(j, 1), q1 = getEdge(j + 1, 1);
    const double p0 = getEdge(k, 2), p1 = getEdge(k + 1, 2);
    return -1. / 3 * (r1 * r1 * r1 - r0 * r0 * r0) * (std::cos(q1) - std::cos(q0)) * (p1 - p0);
}

double SphericalMeshGeometry::meshVolume() const
{
    return -1. / 3 * (upper[0] * upper[0] * upper[0] - lower[0] * lower[0] * lower[0])
           * (std::cos(upper[1]) - std::cos(lower[1])) * (upper[2] - lower[2]);
}

// ---------------------------------------------------------------------------
unsigned long CartesianMeshGeometry::totalCellsInMesh() const
{
    return (unsigned long)shape[0] * shape[1] * shape[2];
}

void CartesianMeshGeometry::cacheSpacing()
{
    for (int n = 0; n < 3; ++n) dx[n] = (upper[n] - lower[n]) / shape[n];
}

// ---------------------------------------------------------------------------
// Host ghost fill for initial data (src/BoundaryConditions.cpp:8-176); the
// step's own ghost fills run on the device.
void BoundaryCondition::applySimple(Array& A, Shape boundaryShape) const
{
    for (int axis = 0; axis < 3; ++axis)
    {
        if (boundaryShape[axis] > 0)
        {
            apply(A, MeshLocation::cell, MeshBoundary::left, axis, boundaryShape[axis]);
            apply(A, MeshLocation::cell, MeshBoundary::right, axis, boundaryShape[axis]);
        }
    }
}

namespace {
void copy_layer(Array& A, int axis, int dst, int src, int negate)
{
    int n[3] = {A.size(0), A.size(1), A.size(2)};
    n[axis] = 1;
    for (int i = 0; i < n[0]; ++i)
    for (int j = 0; j < n[1]; ++j)
    for (int k = 0; k < n[2]; ++k)
    {
        int d[3] = {i, j, k}, s[3] = {i, j, k};
        d[axis] = dst;
        s[axis] = src;
        for (int q = 0; q < A.size(3); ++q)
        {
            double v = A(s[0], s[1], s[2], q);
            if (q == negate) v *= -1;
            A(d[0], d[1], d[2], q) = v;
        }
    }
}
}

void PeriodicBoundaryCondition::apply(Array& A, MeshLocation, MeshBoundary b, int axis, int g) const
{
    if (A.size(axis) == 1) throw std::runtime_error("Attempt to apply boundary condition on flattened axis " + std::to_string(axis));
    const int N = A.size(axis);
    for (int n = 0; n < g; ++n)
        if (b == MeshBoundary::left) copy_layer(A, axis, n, N - 2 * g + n, -1);
        else copy_layer(A, axis, N - g + n, g + n, -1);
}

void OutflowBoundaryCondition::apply(Array& A, MeshLocation, MeshBoundary b, int axis, int g) const
{
    if (A.size(axis) == 1) throw std::runtime_error("Attempt to apply boundary condition on flattened axis " + std::to_string(axis));
    const int N = A.size(axis);
    for (int n = 0; n < g; ++n)
        if (b == MeshBoundary::left) copy_layer(A, axis, n, g, -1);
        else copy_layer(A, axis, N - 1 - n, N - g - 1, -1);
}

void ReflectingBoundaryCondition::apply(Array& A, MeshLocation, MeshBoundary b, int axis, int g) const
{
    if (A.size(axis) == 1) throw std::logic_error("Attempt to apply boundary condition on flattened axis " + std::to_string(axis));
    if (!conservationLaw) throw std::logic_error("ReflectingBoundaryCondition needs an instance of ConservationLaw");
    const int ivel = conservationLaw->getIndexFor(ConservationLaw::VariableType::velocity);
    const int N = A.size(axis);
    for (int n = 0; n < g; ++n)
        if (b == MeshBoundary::left) copy_layer(A, axis, n, 2 * g - n - 1, ivel + axis);
        else copy_layer(A, axis, N - 1 - n, N - 2 * g + n, ivel + axis);
}

// ---------------------------------------------------------------------------
Array MeshOperator::linearCellDimension() const
{
    if (!geometry) throw std::logic_error("MeshOperator has no MeshGeometry instance");
    const Shape s = geometry->cellsShape();
    Array L(s[0], s[1], s[2]);
    for (int i = 0; i < s[0]; ++i)
    for (int j = 0; j < s[1]; ++j)
    for (int k = 0; k < s[2]; ++k)
    {
        double m = geometry->cellLength(i, j, k, 0);
        if (geometry->cellLength(i, j, k, 1) < m) m = geometry->cellLength(i, j, k, 1);
        if (geometry->cellLength(i, j, k, 2) < m) m = geometry->cellLength(i, j, k, 2);
        L(i, j, k) = m;
    }
    return L;
}

Array MeshOperator::generate(InitialDataFunction F, MeshLocation location, Shape bs) const
{
    if (!geometry) throw std::logic_error("MeshOperator has no MeshGeometry instance");
    if (location != MeshLocation::cell) throw std::logic_error("Unsupported mesh location for MeshOperator::generate");
    const int nq = (int)F(0.0, 0.0, 0.0).size();
    const Shape s = geometry->cellsShape();
    Array P(s[0] + 2 * bs[0], s[1] + 2 * bs[1], s[2] + 2 * bs[2], nq);
    for (int i = 0; i < P.size(0); ++i)
    for (int j = 0; j < P.size(1); ++j)
    for (int k = 0; k < P.size(2); ++k)
    {
        const Coordinate x = geometry->coordinateAtIndex(i - bs[0], j - bs[1], k - bs[2]);
        const std::vector<double> v = F(x[0], x[1], x[2]);
        for (int q = 0; q < nq; ++q) P(i, j, k, q) = v[q];
    }
    return P;
}

Array MeshOperator::generate(InitialDataFunction F, MeshLocation location, VectorMode, Shape bs) const
{
    if (!geometry) throw std::logic_error("MeshOperator has no MeshGeometry instance");
    if (location != MeshLocation::face) return generate(F, location, bs);
    const int nq = (int)F(0.0, 0.0, 0.0).size();
    const Shape s = geometry->cellsShape();
    Array A(s[0] + 2 * bs[0] + 1, s[1] + 2 * bs[1] + 1, s[2] + 2 * bs[2] + 1, nq, 3);
    for (int i = 0; i < A.size(0); ++i)
    for (int j = 0; j < A.size(1); ++j)
    for (int k = 0; k < A.size(2); ++k)
    {
        const double c[3] = {double(i - bs[0]), double(j - bs[1]), double(k - bs[2])};
        for (int a = 0; a < 3; ++a)
        {
            // the face of axis a on the left of cell c: half a cell back along a
            const Coordinate x = geometry->coordinateAtIndex(c[0] - (a == 0 ? 0.5 : 0.0), c[1] - (a == 1 ? 0.5 : 0.0),
                                                             c[2] - (a == 2 ? 0.5 : 0.0));
            const std::vector<double> v = F(x[0], x[1], x[2]);
            for (int q = 0; q < nq; ++q) A(i, j, k, q, a) = v[q];
        }
    }
    return A;
}

Array MeshOperator::divergence(const Array& flux, double factor) const
{
    if (!geometry) throw std::logic_error("MeshOperator has no MeshGeometry instance");
    if (flux.size(4) != 3) throw std::logic_error("Attempt to compute divergence from data whose size(4) != 3");
    const int nq = flux.size(3);
    Array D(flux.size(0) - 1, flux.size(1) - 1, flux.size(2) - 1, nq);
    for (int i = 0; i < D.size(0); ++i)
    for (int j = 0; j < D.size(1); ++j)
    for (int k = 0; k < D.size(2); ++k)
    {
        const double a0 = geometry->faceArea(i, j, k, 0), a1 = geometry->faceArea(i, j, k, 1);
        const double a2 = geometry->faceArea(i, j, k, 2);
        const double f = factor / geometry->cellVolume(i, j, k);
        for (int q = 0; q < nq; ++q)
        {
            double s = +flux(i + 1, j, k, q, 0) * a0;
            s -= flux(i, j, k, q, 0) * a0;
            s += flux(i, j + 1, k, q, 1) * a1;
            s -= flux(i, j, k, q, 1) * a1;
            s += flux(i, j, k + 1, q, 2) * a2;
            s -= flux(i, j, k, q, 2) * a2;
            D(i, j, k, q) = f * s;
        }
    }
    return D;
}

// ---------------------------------------------------------------------------
Array CellCenteredFieldCT::vectorPotentialToFluxes(const Array& A) const
{
    Array F(A.size(0), A.size(1), A.size(2), 3, 3);
    // flux of B_q through faces of axis a: (e_a x A)_q with a sign, i.e.
    // F(q, a) = -eps(a, q, r) A_r for the cyclic (a, q, r)
    for (int i = 0; i < A.size(0); ++i)
    for (int j = 0; j < A.size(1); ++j)
    for (int k = 0; k < A.size(2); ++k)
        for (int a = 0; a < 3; ++a)
        {
            const int q = (a + 1) % 3, r = (a + 2) % 3;
            F(i, j, k, q, a) = -A(i, j, k, r, a);
            F(i, j, k, r, a) = +A(i, j, k, q, a);
        }
    return F;
}

Array CellCenteredFieldCT::generateGodunovFluxes(const Array& F, int magneticIndex) const
{
    Array G = F;
    correctGodunovFluxes(G, magneticIndex);
    return G;
}

// Toth (2000) eq. 25: the flux of B_t through faces of axis a becomes the
// average of the four edge values around the face, each edge value the mean
// of the two adjacent flux estimates of the same electric field (+F of B_t
// on axis a, -F of B_a on axis t).  The normal component's own flux is zero.
void CellCenteredFieldCT::correctGodunovFluxes(Array& F, int mi) const
{
    const Array G = F;  // read the uncorrected fluxes throughout
    const int n[3] = {F.size(0), F.size(1), F.size(2)};
    for (int i = 0; i < n[0]; ++i)
    for (int j = 0; j < n[1]; ++j)
    for (int k = 0; k < n[2]; ++k)
        for (int a = 0; a < 3; ++a) F(i, j, k, mi + a, a) = 0.0;
    for (int a = 0; a < 3; ++a)
    {
        if (n[a] < 3) continue;
        for (int t = 0; t < 3; ++t)
        {
            if (t == a) continue;
            int e_a[3] = {0, 0, 0}, e_t[3] = {0, 0, 0};
            e_a[a] = 1;
            e_t[t] = 1;
            int lo[3] = {0, 0, 0}, hi[3] = {n[0], n[1], n[2]};
            lo[a] = 1;
            lo[t] = 1;
            hi[t] = n[t] - 1;
            for (int i = lo[0]; i < hi[0]; ++i)
            for (int j = lo[1]; j < hi[1]; ++j)
            for (int k = lo[2]; k < hi[2]; ++k)
            {
                auto g = [&](int q, int ax, int di, int dj, int dk) { return G(i + di, j + dj, k + dk, q, ax); };
                const int bt = mi + t, ba = mi + a;
                double s = 2 * g(bt, a, 0, 0, 0);
                s += g(bt, a, e_t[0], e_t[1], e_t[2]);
                s += g(bt, a, -e_t[0], -e_t[1], -e_t[2]);
                s -= g(ba, t, e_t[0], e_t[1], e_t[2]);
                s -= g(ba, t, e_t[0] - e_a[0], e_t[1] - e_a[1], e_t[2] - e_a[2]);
                s -= g(ba, t, 0, 0, 0);
                s -= g(ba, t, -e_a[0], -e_a[1], -e_a[2]);
                F(i, j, k, bt, a) = 0.125 * s;
            }
        }
    }
}

// ---------------------------------------------------------------------------
MeshData::MeshData(Shape cs, Shape bs, int nq) : boundaryShape(bs)
{
    P = Array(cs[0] + 2 * bs[0], cs[1] + 2 * bs[1], cs[2] + 2 * bs[2], nq);
}

void MeshData::assignPrimitive(const Array& Pi)
{
    const Shape& b = boundaryShape;
    for (int i = 0; i < Pi.size(0); ++i)
    for (int j = 0; j < Pi.size(1); ++j)
    for (int k = 0; k < Pi.size(2); ++k)
    for (int q = 0; q < Pi.size(3); ++q)
        P(i + b[0], j + b[1], k + b[2], q) = Pi(i, j, k, q);
    deviceVersion = 0;  // host copy replaced
}

void MeshData::assignMagneticField(const Array& B, MeshLocation location, int flags)
{
    if (location != MeshLocation::cell) throw std::logic_error("MeshData::assignMagneticField: cell location only");
    if (magneticIndex < 0) throw std::logic_error("MeshData::assignMagneticField: no magnetic index set");
    const Shape& b = boundaryShape;
    const int o[3] = {flags & includeGuard ? 0 : b[0], flags & includeGuard ? 0 : b[1], flags & includeGuard ? 0 : b[2]};
    for (int i = 0; i < B.size(0); ++i)
    for (int j = 0; j < B.size(1); ++j)
    for (int k = 0; k < B.size(2); ++k)
        for (int d = 0; d < 3; ++d) P(i + o[0], j + o[1], k + o[2], magneticIndex + d) = B(i, j, k, d);
    deviceVersion = 0;
}

// ---------------------------------------------------------------------------
namespace {
std::string status_text(int s) { return mara_error_string(s); }

void throw_for(int status, const char* where, const mara_report* r)
{
    if (status == MARA_OK) return;
    std::string msg = std::string(where) + ": " + status_text(status);
    if (r && r->c2p_failures)
        msg += " (" + std::to_string(r->c2p_failures) + " zones, first at zone index [" +
               std::to_string(r->first_failure[0]) + " " + std::to_string(r->first_failure[1]) + " " +
               std::to_string(r->first_failure[2]) + "])";
    switch (status)
    {
        case MARA_ERR_GUARD_TOO_SMALL:
        case MARA_ERR_UNSUPPORTED:
        case MARA_ERR_INVALID_ARGUMENT: throw std::logic_error(msg);
        case MARA_ERR_INVERSION_FAILURE: throw ConservationLaw::StateFailure(msg);
        default: throw std::runtime_error(msg);
    }
}
}

GpuMethodOfLinesTVD::GpuMethodOfLinesTVD()
{
    fluxScheme = std::make_shared<MethodOfLines>();
}

GpuMethodOfLinesTVD::~GpuMethodOfLinesTVD()
{
    if (block) mara_block_destroy(block);
}

void GpuMethodOfLinesTVD::setRungeKuttaOrder(int order)
{
    if (0 >= order || order > 3) throw std::logic_error("Runge-Kutta order not 1, 2, or 3");
    rungeKuttaOrder = order;
}

int GpuMethodOfLinesTVD::getStencilSize() const
{
    if (!fluxScheme) throw std::logic_error("No IntercellFluxScheme instance");
    return fluxScheme->getStencilSize();
}

mara_desc GpuMethodOfLinesTVD::describe(const MeshData& md) const
{
    if (!fieldOperator) throw std::logic_error("No FieldOperator instance");
    if (!meshOperator) throw std::logic_error("No MeshOperator instance");
    if (!boundaryCondition) throw std::logic_error("No BoundaryCondition instance");
    if (!fluxScheme) throw std::logic_error("No IntercellFluxScheme instance");
    auto law = fieldOperator->getConservationLaw();
    if (!law) throw std::logic_error("FieldOperator has no ConservationLaw");
    auto geom = meshOperator->getMeshGeometry();
    if (!geom) throw std::logic_error("MeshOperator has no MeshGeometry instance");
    mara_desc d;
    std::memset(&d, 0, sizeof d);
    d.law = law->lawCode();
    d.gamma = law->getGammaLawIndex();
    d.pressure_floor = law->getPressureFloor();
    d.cooling_rate = law->getCoolingRate();
    d.recon = fluxScheme->reconCode();
    d.plm_theta = fluxScheme->getPlmTheta();
    d.riemann = fluxScheme->getRiemannSolver() ? fluxScheme->getRiemannSolver()->solverCode() : MARA_RIEMANN_HLLE;
    d.bc = boundaryCondition->bcCode();
    const Shape cs = geom->cellsShape();
    geom->describe(d);
    for (int a = 0; a < 3; ++a)
    {
        d.guards[a] = md.boundaryShape[a];
        if (md.P.size(a) != cs[a] + 2 * md.boundaryShape[a])
            throw std::logic_error("MeshData shape does not match the mesh geometry");
    }
    if (md.P.size(3) != law->getNumConserved())
        throw std::logic_error("MeshData field count does not match the conservation law");
    d.rk_order = rungeKuttaOrder;
    d.field_ct = disableFieldCT ? 0 : 1;
    d.precision = precision;
    return d;
}

void GpuMethodOfLinesTVD::check(int status, const char* where) const
{
    throw_for(status, where, &report);
}

mara_block* GpuMethodOfLinesTVD::blockFor(const MeshData& md) const
{
    const mara_desc d = describe(md);
    if (block && (blockOwner != &md || std::memcmp(&d, &blockDesc, sizeof d) != 0))
    {
        mara_block_destroy(block);
        block = nullptr;
    }
    if (!block)
    {
        // the block is kept only once its hooks are on it: a failed hook
        // leaves no cached block that would step without them
        mara_block* fresh = nullptr;
        check(mara_block_create(&d, &fresh), "mara_block_create");
        try
        {
            applySource(fresh);
        }
        catch (...)
        {
            mara_block_destroy(fresh);
            throw;
        }
        block = fresh;
        blockDesc = d;
        blockOwner = &md;
    }
    else if (!hooksApplied)
        applySource(block);
    return block;
}

void GpuMethodOfLinesTVD::applySource(mara_block* b) const
{
    hooksApplied = false;
    if (!sourcePath.empty() && fieldOperator && fieldOperator->getConservationLaw() &&
        fieldOperator->getConservationLaw()->getCoolingRate() > 0.0)
        throw std::logic_error("a source-terms function and NewtonianMHD cooling on one scheme are not supported");
    if (sourcePath.empty())
        check(mara_block_set_source(b, nullptr, nullptr, nullptr, 0), "mara_block_set_source");
    else
        check(mara_block_set_source(b, sourcePath.c_str(), sourceKernel.c_str(), sourceParams.data(),
                                    (int)sourceParams.size()),
              "mara_block_set_source");
    check(mara_block_set_source_aux(b, sourceAux.data()), "mara_block_set_source_aux");
    if (viscousPath.empty())
        check(mara_block_set_viscous_flux(b, nullptr, nullptr, nullptr, 0), "mara_block_set_viscous_flux");
    else
        check(mara_block_set_viscous_flux(b, viscousPath.c_str(), viscousKernel.c_str(), viscousParams.data(),
                                          (int)viscousParams.size()),
              "mara_block_set_viscous_flux");
    hooksApplied = true;
}

void GpuMethodOfLinesTVD::setSourceTermsCodeObject(const std::string& path, const std::string& kernel,
                                                   std::vector<double> params)
{
    if (params.size() > MARA_SOURCE_MAX_PARAMS) throw std::logic_error("at most 16 source-term parameters");
    auto law = fieldOperator ? fieldOperator->getConservationLaw() : nullptr;
    if (!path.empty() && law && law->getCoolingRate() > 0.0)
        throw std::logic_error("a source-terms function and NewtonianMHD cooling on one scheme are not supported");
    const std::string oldPath = sourcePath, oldKernel = sourceKernel;
    const std::vector<double> oldParams = sourceParams;
    sourcePath = path;
    sourceKernel = kernel;
    sourceParams = std::move(params);
    if (!block) return;
    try
    {
        applySource(block);
    }
    catch (...)
    {
        // keep the previous hooks; the next use of the block puts them back
        sourcePath = oldPath;
        sourceKernel = oldKernel;
        sourceParams = oldParams;
        throw;
    }
}

void GpuMethodOfLinesTVD::setViscousFluxCodeObject(const std::string& path, const std::string& kernel,
                                                   std::vector<double> params)
{
    if (params.size() > MARA_SOURCE_MAX_PARAMS) throw std::logic_error("at most 16 viscous-flux parameters");
    const std::string oldPath = viscousPath, oldKernel = viscousKernel;
    const std::vector<double> oldParams = viscousParams;
    viscousPath = path;
    viscousKernel = kernel;
    viscousParams = std::move(params);
    if (!block) return;
    try
    {
        applySource(block);
    }
    catch (...)
    {
        viscousPath = oldPath;
        viscousKernel = oldKernel;
        viscousParams = oldParams;
        throw;
    }
}

void GpuMethodOfLinesTVD::setStarParticleData(const std::array<double, 8>& t)
{
    sourceAux = t;
    if (block) check(mara_block_set_source_aux(block, sourceAux.data()), "mara_block_set_source_aux");
}

mara_shape GpuMethodOfLinesTVD::autotune(const MeshData& md, int repeats) const
{
    mara_block* b = blockFor(md);
    if (!resident || md.deviceVersion == 0) check(mara_block_upload_prim(b, md.P.data()), "mara_block_upload_prim");
    mara_shape s{};
    check(mara_block_autotune(b, repeats, &s, nullptr), "mara_block_autotune");
    return s;
}

void GpuMethodOfLinesTVD::syncToDevice(MeshData& md) const
{
    mara_block* b = blockFor(md);
    check(mara_block_upload_prim(b, md.P.data()), "mara_block_upload_prim");
    md.deviceVersion = 1;
}

void GpuMethodOfLinesTVD::syncToHost(MeshData& md) const
{
    mara_block* b = blockFor(md);
    check(mara_block_download_prim(b, md.P.data()), "mara_block_download_prim");
}

void GpuMethodOfLinesTVD::advance(MeshData& md, double t0, double dt) const
{
    mara_block* b = blockFor(md);
    if (!resident || md.deviceVersion == 0)
        syncToDevice(md);
    std::memset(&report, 0, sizeof report);
    const int st = mara_block_advance(b, t0, dt, &report);
    md.deviceVersion += 1;
    if (!resident) check(mara_block_download_prim(b, md.P.data()), "mara_block_download_prim");
    check(st, "MethodOfLinesTVD::advance");
}

int GpuMethodOfLinesTVD::run(MeshData& md, double cfl, double tFinal, int maxSteps, double& t) const
{
    mara_block* b = blockFor(md);
    if (!resident || md.deviceVersion == 0) syncToDevice(md);
    int steps = 0;
    std::memset(&report, 0, sizeof report);
    const int st = mara_block_run(b, cfl, tFinal, maxSteps, &t, &steps, &report);
    md.deviceVersion += 1;
    if (!resident) check(mara_block_download_prim(b, md.P.data()), "mara_block_download_prim");
    check(st, "MethodOfLinesTVD::run");
    return steps;
}

// ---------------------------------------------------------------------------
FieldOperator::~FieldOperator()
{
    if (block) mara_block_destroy(block);
}

double FieldOperator::getCourantTimestep(const MeshData& md, const Array& L) const
{
    if (!law) throw std::logic_error("FieldOperator has no ConservationLaw");
    if (!geometry) throw std::logic_error("FieldOperator needs the MeshGeometry (setMeshGeometry)");
    const Shape cs = geometry->cellsShape();
    if (L.size(0) != cs[0] || L.size(1) != cs[1] || L.size(2) != cs[2])
        throw std::logic_error("Primitive and linear cell dimension arrays have different sizes");
    // the device evaluates linearCellDimension of the mesh itself: one length
    // on a Cartesian mesh, the geometry's per-cell lengths on a spherical one
    const bool cartesian = geometry->getType() == "cartesian";
    for (std::size_t n = 1; cartesian && n < L.size(); ++n)
        if (L.data()[n] != L.data()[0]) throw std::logic_error("non-uniform linear cell dimension is not supported");
    mara_desc d;
    std::memset(&d, 0, sizeof d);
    d.law = law->lawCode();
    d.gamma = law->getGammaLawIndex();
    d.pressure_floor = law->getPressureFloor();
    d.cooling_rate = law->getCoolingRate();
    d.recon = MARA_RECON_PCM;
    d.plm_theta = 1.5;
    d.bc = MARA_BC_OUTFLOW;
    d.rk_order = 1;
    geometry->describe(d);
    for (int a = 0; a < 3; ++a) d.guards[a] = md.boundaryShape[a];
    if (block && std::memcmp(&d, &blockDesc, sizeof d) != 0)
    {
        mara_block_destroy(block);
        block = nullptr;
    }
    if (!block)
    {
        throw_for(mara_block_create(&d, &block), "FieldOperator::getCourantTimestep", nullptr);
        blockDesc = d;
    }
    double dt = 0.0;
    int st = mara_block_upload_prim(block, md.P.data());
    if (st == MARA_OK) st = mara_block_courant_dt(block, &dt);
    throw_for(st, "FieldOperator::getCourantTimestep", nullptr);
    return dt;
}

// ---------------------------------------------------------------------------
int maraMainLoop(SimulationStatus& status, std::function<double()> timestep, std::function<bool()> condition,
                 std::function<void(double)> advance, double* meanKzps)
{
    double sessionKzps = 0.0;
    int sessionIter = 0;
    while (condition())
    {
        const auto t0 = std::chrono::steady_clock::now();
        const double dt = timestep();
        advance(dt);
        status.simulationTime += dt;
        status.simulationIter += 1;
        const double age = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        sessionKzps += 1e-3 * status.totalCellsInMesh / age;
        sessionIter += 1;
    }
    if (meanKzps) *meanKzps = sessionIter ? sessionKzps / sessionIter : 0.0;
    return sessionIter;
}

} // namespace mara2
