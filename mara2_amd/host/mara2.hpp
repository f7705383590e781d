// mara2.hpp -- C++ host layer over the C ABI (include/mara_gpu.h) that keeps
// Mara2's plugin interfaces for the time step: SolutionScheme,
// IntercellFluxScheme, RiemannSolver, ConservationLaw, BoundaryCondition,
// MeshGeometry, MeshOperator, FieldOperator and MeshData (src/Mara.hpp:44-532).
//
// The classes keep the reference's names, setter names, argument meaning and
// exception types, so a driver written against them reads like the
// reference's src/Problems.cpp.  Collaborators are configuration objects:
// GpuMethodOfLinesTVD maps the injected ones to a mara_desc (by dynamic_cast,
// as the reference's own code dispatches on them) and runs the whole step on
// the GPU.  Configurations the GPU step does not implement (source terms,
// viscous fluxes, non-Cartesian meshes, DrivenMHDBoundary, MPI decomposition)
// are not constructible here, and an unsupported combination raises
// std::logic_error -- there is no CPU fallback.
#pragma once

#include <array>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/mara_gpu.h"

namespace mara2 {

using Shape = std::array<int, 5>;
using Coordinate = std::array<double, 3>;
using AreaElement = std::array<double, 3>;
using InitialDataFunction = std::function<std::vector<double>(double x, double y, double z)>;
enum class MeshLocation { vert, edge, face, cell };
enum class MeshBoundary { left, right };

// ---------------------------------------------------------------------------
// A row-major 5-D host array with the reference's index order (i, j, k, q, r)
// and zero initialisation (the layout of Cow::Array used by MeshData).
// ---------------------------------------------------------------------------
class Array
{
public:
    Array() : shape_{{0, 0, 0, 0, 0}} {}
    explicit Array(Shape s);
    Array(int n1, int n2 = 1, int n3 = 1, int n4 = 1, int n5 = 1);
    double& operator()(int i, int j = 0, int k = 0, int q = 0, int r = 0);
    double operator()(int i, int j = 0, int k = 0, int q = 0, int r = 0) const;
    double* data() { return data_.data(); }
    const double* data() const { return data_.data(); }
    int size(int axis) const { return shape_[axis]; }
    std::size_t size() const { return data_.size(); }
    const Shape& shape() const { return shape_; }
private:
    Shape shape_;
    std::vector<double> data_;
};

// ---------------------------------------------------------------------------
// Physics plug-ins (configuration objects)
// ---------------------------------------------------------------------------
class ConservationLaw
{
public:
    enum class VariableType { density, velocity, pressure, magnetic };
    class StateFailure : public std::runtime_error
    {
    public:
        explicit StateFailure(const std::string& what) : std::runtime_error(what) {}
    };
    virtual ~ConservationLaw() {}
    virtual void setPressureFloor(double pf) { pressureFloor = pf; }
    virtual void setGammaLawIndex(double gm) { gammaLawIndex = gm; }
    virtual int getNumConserved() const = 0;
    virtual int getIndexFor(VariableType type) const = 0;
    virtual std::string getPrimitiveName(int fieldIndex) const = 0;
    std::vector<std::string> getPrimitiveNames() const;
    double getGammaLawIndex() const { return gammaLawIndex; }
    double getPressureFloor() const { return pressureFloor; }
    virtual double getCoolingRate() const { return -1.0; }  // NewtonianMHD only
    virtual int lawCode() const = 0;
protected:
    double gammaLawIndex = 5. / 3;  // reference default (src/ConservationLaws.cpp:233,351,611)
    double pressureFloor = -1.0;
};

class NewtonianHydro : public ConservationLaw
{
public:
    int getNumConserved() const override { return 5; }
    int getIndexFor(VariableType type) const override;
    std::string getPrimitiveName(int fieldIndex) const override;
    int lawCode() const override { return MARA_LAW_HYDRO; }
};

class NewtonianMHD : public ConservationLaw
{
public:
    // NewtonianMHD::setCoolingRate (src/ConservationLaws.cpp:354-358): a
    // positive rate adds the auxiliary ninth field ("auxiliary1") and the
    // cooling source (mara_desc.cooling_rate)
    void setCoolingRate(double cr) { coolingRate = cr; }
    double getCoolingRate() const override { return coolingRate; }
    int getNumConserved() const override { return coolingRate > 0.0 ? 9 : 8; }
    int getIndexFor(VariableType type) const override;
    std::string getPrimitiveName(int fieldIndex) const override;
    int lawCode() const override { return MARA_LAW_MHD; }
private:
    double coolingRate = -1.0;  // reference default: off
};

class RelativisticMHD : public NewtonianMHD
{
public:
    double getCoolingRate() const override { return -1.0; }
    int getNumConserved() const override { return 8; }
    int lawCode() const override { return MARA_LAW_RMHD; }
};

class RiemannSolver
{
public:
    virtual ~RiemannSolver() {}
    virtual int solverCode() const = 0;
};
class HlleRiemannSolver : public RiemannSolver { public: int solverCode() const override { return MARA_RIEMANN_HLLE; } };
class HllcNewtonianHydroRiemannSolver : public RiemannSolver { public: int solverCode() const override { return MARA_RIEMANN_HLLC; } };

class IntercellFluxScheme
{
public:
    virtual ~IntercellFluxScheme() {}
    virtual void setPlmTheta(double) {}
    virtual void setRiemannSolver(std::shared_ptr<RiemannSolver> rs) { riemannSolver = rs; }
    virtual int getStencilSize() const = 0;
    virtual int reconCode() const = 0;
    double getPlmTheta() const { return plmTheta; }
    std::shared_ptr<RiemannSolver> getRiemannSolver() const { return riemannSolver; }
protected:
    double plmTheta = 1.5;
    std::shared_ptr<RiemannSolver> riemannSolver = std::make_shared<HlleRiemannSolver>();
};

// src/IntercellFluxSchemes.cpp:35-63 (piecewise constant), default HLLE
class MethodOfLines : public IntercellFluxScheme
{
public:
    int getStencilSize() const override { return 1; }
    int reconCode() const override { return MARA_RECON_PCM; }
};

// src/IntercellFluxSchemes.cpp:69-117, theta 1.5, default HLLE
class MethodOfLinesPlm : public IntercellFluxScheme
{
public:
    void setPlmTheta(double t) override { plmTheta = t; }
    int getStencilSize() const override { return 2; }
    int reconCode() const override { return MARA_RECON_PLM; }
};

// The MethodOfLinesPlm structure with Reconstruction::weno5 (WENO5_FD_C2R /
// C2L, ImprovedShenZha10, A = 50; src/Reconstruction.cpp:68-169).
class MethodOfLinesWeno5 : public IntercellFluxScheme
{
public:
    int getStencilSize() const override { return 3; }
    int reconCode() const override { return MARA_RECON_WENO5; }
};

// ---------------------------------------------------------------------------
// Mesh
// ---------------------------------------------------------------------------
class MeshGeometry
{
public:
    virtual ~MeshGeometry() {}
    virtual void setCellsShape(Shape) = 0;
    virtual void setLowerUpper(Coordinate, Coordinate) {}
    virtual Shape cellsShape() const = 0;
    virtual Coordinate coordinateAtIndex(double i, double j, double k) const = 0;
    virtual unsigned long totalCellsInMesh() const = 0;
    virtual double cellLength(int i, int j, int k, int axis) const = 0;
    virtual double faceArea(int i, int j, int k, int axis) const = 0;
    virtual double cellVolume(int i, int j, int k) const = 0;
    virtual Coordinate lowerCorner() const = 0;
    virtual Coordinate upperCorner() const = 0;
    virtual std::string getType() const = 0;
    // the mesh's part of a block descriptor: cells, lower, upper, geometry
    void describe(mara_desc& d) const;
protected:
    virtual int geometryCode() const = 0;
    virtual int logRadial() const { return 0; }
};

// src/CartesianMeshGeometry.cpp:61-157
class CartesianMeshGeometry : public MeshGeometry
{
public:
    CartesianMeshGeometry();
    explicit CartesianMeshGeometry(Shape S);
    void setCellsShape(Shape S) override;
    void setLowerUpper(Coordinate L, Coordinate U) override;
    Shape cellsShape() const override { return shape; }
    Coordinate coordinateAtIndex(double i, double j, double k) const override;
    unsigned long totalCellsInMesh() const override;
    double cellLength(int, int, int, int axis) const override { return dx[axis]; }
    // src/CartesianMeshGeometry.cpp:79-92: uniform areas and volume
    double faceArea(int, int, int, int axis) const override { return axis == 0 ? dx[1] * dx[2] : axis == 1 ? dx[2] * dx[0] : dx[0] * dx[1]; }
    double cellVolume(int, int, int) const override { return dx[0] * dx[1] * dx[2]; }
    Coordinate lowerCorner() const override { return lower; }
    Coordinate upperCorner() const override { return upper; }
    std::string getType() const override { return "cartesian"; }
protected:
    int geometryCode() const override { return MARA_GEOM_CARTESIAN; }
private:
    void cacheSpacing();
    Shape shape;
    Coordinate lower, upper;
    double dx[3];
};

// src/SphericalMeshGeometry.cpp: axes (r, theta, phi); linear edges, or
// logarithmic radial ones.  The GPU step evaluates the same areas, volumes
// and lengths from the descriptor (MARA_GEOM_SPHERICAL).
class SphericalMeshGeometry : public MeshGeometry
{
public:
    SphericalMeshGeometry();
    explicit SphericalMeshGeometry(Shape S);
    SphericalMeshGeometry(int ni, int nj, int nk);
    void setUseLogarithmicRadialBinning(bool shouldUse) { logarithmic = shouldUse; }
    void setCellsShape(Shape S) override;
    void setLowerUpper(Coordinate L, Coordinate U) override { lower = L; upper = U; }
    Shape cellsShape() const override { return shape; }
    Coordinate coordinateAtIndex(double i, double j, double k) const override;
    unsigned long totalCellsInMesh() const override;
    double cellLength(int i, int j, int k, int axis) const override;
    double faceArea(int i, int j, int k, int axis) const override;
    double cellVolume(int i, int j, int k) const override;
    double meshVolume() const;
    Coordinate lowerCorner() const override { return lower; }
    Coordinate upperCorner() const override { return upper; }
    std::string getType() const override { return "spherical"; }
    double getEdge(double n, int axis) const;
protected:
    int geometryCode() const override { return MARA_GEOM_SPHERICAL; }
    int logRadial() const override { return logarithmic ? 1 : 0; }
private:
    Shape shape;
    Coordinate lower, upper;
    bool logarithmic = false;
};

class BoundaryCondition
{
public:
    virtual ~BoundaryCondition() {}
    virtual void apply(Array& A, MeshLocation location, MeshBoundary boundary, int axis, int numGuard) const = 0;
    void applySimple(Array& A, Shape boundaryShape) const;  // src/BoundaryConditions.cpp:8-18
    virtual int bcCode() const = 0;
    virtual void setConservationLaw(std::shared_ptr<ConservationLaw> law) { conservationLaw = law; }
protected:
    std::shared_ptr<ConservationLaw> conservationLaw;
};

class PeriodicBoundaryCondition : public BoundaryCondition
{
public:
    void apply(Array& A, MeshLocation, MeshBoundary boundary, int axis, int numGuard) const override;
    int bcCode() const override { return MARA_BC_PERIODIC; }
};

class OutflowBoundaryCondition : public BoundaryCondition
{
public:
    void apply(Array& A, MeshLocation, MeshBoundary boundary, int axis, int numGuard) const override;
    int bcCode() const override { return MARA_BC_OUTFLOW; }
};

class ReflectingBoundaryCondition : public BoundaryCondition
{
public:
    void apply(Array& A, MeshLocation, MeshBoundary boundary, int axis, int numGuard) const override;
    int bcCode() const override { return MARA_BC_REFLECTING; }
};

class MeshOperator
{
public:
    explicit MeshOperator(std::shared_ptr<MeshGeometry> geometry = nullptr) : geometry(geometry) {}
    void setMeshGeometry(std::shared_ptr<MeshGeometry> g) { geometry = g; }
    std::shared_ptr<MeshGeometry> getMeshGeometry() const { return geometry; }
    // src/MeshOperator.cpp:292-307
    Array linearCellDimension() const;
    // src/MeshOperator.cpp:47-195, cell location
    Array generate(InitialDataFunction F, MeshLocation location, Shape boundaryShape = {{0, 0, 0, 0, 0}}) const;
    // face location with every component on the faces of every axis
    // (VectorMode::scalars): an array (N1+1, N2+1, N3+1, nq, 3), guards
    // included, component q on the faces of axis a at the face centres
    enum class VectorMode { scalars };
    Array generate(InitialDataFunction F, MeshLocation location, VectorMode mode, Shape boundaryShape) const;
    // src/MeshOperator.cpp:323-381: factor / V * (sum over axes of F A on the
    // right face minus F A on the left), flux (M1+1, M2+1, M3+1, nq, 3) ->
    // (M1, M2, M3, nq)
    Array divergence(const Array& flux, double factor) const;
private:
    std::shared_ptr<MeshGeometry> geometry;
};

// The host half of CellCenteredFieldCT (src/CellCenteredFieldCT.cpp:14-157)
// that initial data uses: fluxes of B from a vector potential on faces, and
// Toth's flux-CT correction of them; the step's own flux-CT runs on the device.
class CellCenteredFieldCT
{
public:
    // A (…, 3, 3) [component, face axis] -> the induction fluxes F = -n x A
    Array vectorPotentialToFluxes(const Array& A) const;
    // a corrected copy of F (correctGodunovFluxes with B at magneticIndex)
    Array generateGodunovFluxes(const Array& F, int magneticIndex) const;
    void correctGodunovFluxes(Array& F, int magneticIndex) const;
};

// src/MeshData.cpp:8-28: primitives with guard zones, reference layout.
class MeshData
{
public:
    enum { includeGuard = 1 };
    MeshData(Shape cellsShape, Shape boundaryShape, int numComponents);
    void assignPrimitive(const Array& interiorP);
    void setMagneticIndex(int index) { magneticIndex = index; }
    // cell-centred B (M1, M2, M3, 3) into the magnetic fields of P, over the
    // whole array with includeGuard, else over the interior
    void assignMagneticField(const Array& B, MeshLocation location, int flags);
    int magneticIndex = -1;
    void applyBoundaryCondition(BoundaryCondition& bc) { bc.applySimple(P, boundaryShape); }
    Shape getBoundaryShape() const { return boundaryShape; }
    Array P;
    Shape boundaryShape;
    // incremented by every device step; lets the scheme keep state resident
    unsigned long long deviceVersion = 0;
};

// ---------------------------------------------------------------------------
// Operators and the solution scheme, executed on the GPU
// ---------------------------------------------------------------------------
class FieldOperator
{
public:
    explicit FieldOperator(std::shared_ptr<ConservationLaw> law = nullptr) : law(law) {}
    virtual ~FieldOperator();
    FieldOperator(const FieldOperator&) = delete;
    FieldOperator& operator=(const FieldOperator&) = delete;
    void setConservationLaw(std::shared_ptr<ConservationLaw> cl) { law = cl; }
    std::shared_ptr<ConservationLaw> getConservationLaw() const { return law; }
    // min over interior zones of L / max lambda (src/FieldOperator.cpp:125-169);
    // P is the MeshData primitive array with guards, L from
    // MeshOperator::linearCellDimension.  Evaluated on the GPU.
    virtual double getCourantTimestep(const MeshData& md, const Array& L) const;
    void setMeshGeometry(std::shared_ptr<MeshGeometry> g) { geometry = g; }
private:
    std::shared_ptr<ConservationLaw> law;
    std::shared_ptr<MeshGeometry> geometry;
    mutable mara_block* block = nullptr;  // reused while the descriptor is unchanged
    mutable mara_desc blockDesc{};
};

class SolutionScheme
{
public:
    virtual ~SolutionScheme() {}
    virtual int getStencilSize() const = 0;
    virtual void advance(MeshData& solution, double t0, double dt) const = 0;
    virtual void setBoundaryCondition(std::shared_ptr<BoundaryCondition> bc) { boundaryCondition = bc; }
    virtual void setFieldOperator(std::shared_ptr<FieldOperator> fo) { fieldOperator = fo; }
    virtual void setMeshOperator(std::shared_ptr<MeshOperator> mo) { meshOperator = mo; }
    virtual void setIntercellFluxScheme(std::shared_ptr<IntercellFluxScheme> fs) { fluxScheme = fs; }
protected:
    std::shared_ptr<BoundaryCondition> boundaryCondition;
    std::shared_ptr<FieldOperator> fieldOperator;
    std::shared_ptr<MeshOperator> meshOperator;
    std::shared_ptr<IntercellFluxScheme> fluxScheme;
};

// MethodOfLinesTVD (src/SolutionSchemes.cpp:70-275) on the GPU.
class GpuMethodOfLinesTVD : public SolutionScheme
{
public:
    GpuMethodOfLinesTVD();
    ~GpuMethodOfLinesTVD() override;
    GpuMethodOfLinesTVD(const GpuMethodOfLinesTVD&) = delete;
    GpuMethodOfLinesTVD& operator=(const GpuMethodOfLinesTVD&) = delete;
    void setRungeKuttaOrder(int order);          // logic_error outside 1..3
    void setDisableFieldCT(bool disable) { disableFieldCT = disable; }
    void setPrecision(int precision) { this->precision = precision; }  // MARA_PREC_*
    // Resident mode keeps the state on the device between steps: advance()
    // uploads only when md.P was replaced on the host (syncToDevice) and
    // downloads only on syncToHost().  Off by default (strict drop-in: every
    // advance reads and writes md.P).
    void setResident(bool r) { resident = r; }
    void syncToHost(MeshData& md) const;
    void syncToDevice(MeshData& md) const;
    int getStencilSize() const override;
    void advance(MeshData& solution, double t0, double dt) const override;
    // the device-driven loop for resident runs: steps of dt = cfl * courant
    // until t_final or max_steps; returns steps taken, advances t
    int run(MeshData& solution, double cfl, double tFinal, int maxSteps, double& t) const;
    mara_desc describe(const MeshData& solution) const;
    const mara_report& lastReport() const { return report; }
    // setSourceTermsFunction (src/SolutionSchemes.hpp:36) on the GPU: the
    // per-zone function compiled into a gfx950 code object written with
    // include/mara_source.h, its kernel name and up to 16 parameters; an
    // empty path removes it.  setStarParticleData stands for the 8 auxiliary
    // values the reference passes to the function (its star-particle data).
    void setSourceTermsCodeObject(const std::string& path, const std::string& kernel,
                                  std::vector<double> params = {});
    void setStarParticleData(const std::array<double, 8>& t);
    // setViscousFluxFunction (src/SolutionSchemes.cpp:91-94) on the GPU: the
    // per-face function in a code object (MARA_VISCOUS_FLUX_KERNEL of
    // include/mara_source.h), its kernel name and up to 16 parameters; an
    // empty path removes it.  The C ABI refuses it in the mixed form and for
    // a magnetic law with flux-CT (mara_block_set_viscous_flux).
    void setViscousFluxCodeObject(const std::string& path, const std::string& kernel,
                                  std::vector<double> params = {});
    // launch-shape autotuner on the block of md (mara_block_autotune); the
    // state is left as it is
    mara_shape autotune(const MeshData& md, int repeats = 5) const;
private:
    mara_block* blockFor(const MeshData& md) const;
    void applySource(mara_block* b) const;
    void check(int status, const char* where) const;
    std::string sourcePath, sourceKernel;
    std::vector<double> sourceParams;
    std::array<double, 8> sourceAux{};
    std::string viscousPath, viscousKernel;
    std::vector<double> viscousParams;
    int rungeKuttaOrder = 1;
    bool disableFieldCT = false;
    int precision = MARA_PREC_FP64;
    bool resident = false;
    mutable mara_block* block = nullptr;
    mutable mara_desc blockDesc{};
    mutable const MeshData* blockOwner = nullptr;
    mutable bool hooksApplied = false;  // the block carries the current source and viscous hooks
    mutable mara_report report{};
};

// maraMainLoop without scheduler and logger (src/Mara.cpp:42-88): while
// condition(): dt = timestep(); advance(dt); t += dt.  Returns the number
// of iterations; *kzps receives the mean kilo-zone-updates per second.
struct SimulationStatus
{
    double simulationTime = 0.0;
    int simulationIter = 0;
    long totalCellsInMesh = 0;
};

int maraMainLoop(SimulationStatus& status, std::function<double()> timestep, std::function<bool()> condition,
                 std::function<void(double)> advance, double* meanKzps = nullptr);

} // namespace mara2
