// C++ host layer over libsphgpu with the reference's SPH package interface.
//
// A Spheral user drives the SPH hydro through the Physics-package entry points
// of SPHHydroBase<Dimension> (src/SPH/SPHHydroBase.hh:52-160,
// src/Physics/Physics.hh:51-159) and through SynchronousRK2
// (src/Integrator/SynchronousRK2.hh:32-35, Integrator.hh:74-221).  This header
// gives those classes with the same names, constructor arguments and method
// signatures; every computation runs on the GPU behind the C-ABI of sphgpu.h,
// and the header itself holds no physics.
//
// The containers are minimal stand-ins with the reference's names: a NodeList
// (FluidNodeList, src/NodeList/FluidNodeLists.py:12-35) owns its fields, State
// and StateDerivatives hold named fields (the reference's FieldLists)
// under the reference's names (src/Hydro/HydroFieldNames.cc:10-69)
// with the "delta " / "new " prefixes of IncrementFieldList / ReplaceFieldList
// (src/DataBase/IncrementFieldList.hh:48, ReplaceFieldList.hh:52).  A field is
// a flat std::vector<double> in the reference's element layout: Vector = nDim,
// SymTensor = nDim(nDim+1)/2 (xx, xy, .., yy, ..), Tensor = nDim^2 (row-major)
// doubles per node; internal nodes first, then the NodeList's own ghost nodes.
// With several NodeLists a field holds each NodeList's (internal) rows in turn,
// in the DataBase's order.
// Failed checks and device errors throw VERIFYError, as the reference's
// VERIFY2 does (src/Utilities/DBC.hh:160-169).
#pragma once

#include <algorithm>
#include <array>
#include <cstdint>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "sphgpu.h"

namespace Spheral {

struct VERIFYError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
inline void VERIFY2(bool ok, const std::string& msg) {
  if (!ok) throw VERIFYError(msg);
}

template <int ND>
struct Dim {
  static constexpr int nDim = ND;
  static constexpr int nSym = ND * (ND + 1) / 2;
  using Scalar = double;
  using Vector = std::array<double, ND>;
  using SymTensor = std::array<double, nSym>;
  using Tensor = std::array<double, ND * ND>;
};
using Dim1 = Dim<1>;
using Dim2 = Dim<2>;
using Dim3 = Dim<3>;

namespace HydroFieldNames {
inline const std::string mass = "mass";
inline const std::string position = "position";
inline const std::string velocity = "velocity";
inline const std::string H = "H";
inline const std::string velocityGradient = "velocity gradient";
inline const std::string internalVelocityGradient = "internal velocity gradient";
inline const std::string hydroAcceleration = "delta velocity hydro";
inline const std::string massDensity = "mass density";
inline const std::string normalization = "normalization";
inline const std::string specificThermalEnergy = "specific thermal energy";
inline const std::string maxViscousPressure = "max viscous pressure";
inline const std::string effectiveViscousPressure = "effective viscous pressure";
inline const std::string viscousWork = "viscous work rate";
inline const std::string XSPHDeltaV = "XSPH delta vi";
inline const std::string XSPHWeightSum = "XSPH weight sum";
inline const std::string massSecondMoment = "mass second moment";
inline const std::string weightedNeighborSum = "weighted neighbor sum";
inline const std::string pressure = "pressure";
inline const std::string soundSpeed = "sound speed";
inline const std::string omegaGradh = "grad h corrections";
inline const std::string M_SPHCorrection = "M SPH gradient correction";
}  // namespace HydroFieldNames

inline const std::string IncrementPrefix = "delta ";  // IncrementFieldList::prefix()
inline const std::string ReplacePrefix = "new ";      // ReplaceFieldList::prefix()

// Named flat fields: the FieldLists of the DataBase's NodeLists.
class FieldStore {
 public:
  std::vector<double>& field(const std::string& name) {
    auto it = mFields.find(name);
    VERIFY2(it != mFields.end(), "no field named \"" + name + "\"");
    return it->second;
  }
  const std::vector<double>& field(const std::string& name) const {
    auto it = mFields.find(name);
    VERIFY2(it != mFields.end(), "no field named \"" + name + "\"");
    return it->second;
  }
  // Make or resize a field of rows x comps doubles; new entries take `fill`.
  std::vector<double>& enroll(const std::string& name, size_t rows, int comps, double fill = 0.0) {
    auto& f = mFields[name];
    f.resize(rows * size_t(comps), fill);
    return f;
  }
  // Set a field to a copy of `values`.
  void enroll(const std::string& name, const std::vector<double>& values) { mFields[name] = values; }
  bool registered(const std::string& name) const { return mFields.count(name) != 0; }
  std::vector<std::string> keys() const {
    std::vector<std::string> k;
    for (const auto& kv : mFields) k.push_back(kv.first);
    return k;
  }

 private:
  std::map<std::string, std::vector<double>> mFields;
};

template <typename Dimension> class State : public FieldStore {};
template <typename Dimension> class StateDerivatives : public FieldStore {};

// src/Material/EquationOfState.hh:24-27
enum class MaterialPressureMinType { PressureFloor = 0, ZeroPressure = 1 };

// GammaLawGas(gamma, mu, constants, minimumPressure, maximumPressure,
// minPressureType) (src/Material/GammaLawGas.hh:26-31); the pressure and sound
// speed are set on the device (src/Material/GammaLawGas.cc:163-223).
template <typename Dimension>
class GammaLawGas {
 public:
  GammaLawGas(double gamma, double mu, double minimumPressure = -1.0e200, double maximumPressure = 1.0e200,
              MaterialPressureMinType minPressureType = MaterialPressureMinType::PressureFloor)
      : mGamma(gamma), mMu(mu), mPmin(minimumPressure), mPmax(maximumPressure), mMinType(minPressureType) {
    VERIFY2(gamma > 1.0, "GammaLawGas needs gamma > 1");
  }
  double gamma() const { return mGamma; }
  double molecularWeight() const { return mMu; }
  double minimumPressure() const { return mPmin; }
  double maximumPressure() const { return mPmax; }
  MaterialPressureMinType minimumPressureType() const { return mMinType; }

 private:
  double mGamma, mMu, mPmin, mPmax;
  MaterialPressureMinType mMinType;
};

// FluidNodeList(name, eos, numInternal, numGhost, hmin, hmax, hminratio, nPerh,
// maxNumNeighbors, rhoMin, rhoMax) (src/NodeList/FluidNodeLists.py:12-35).
template <typename Dimension>
class NodeList {
 public:
  static constexpr int nDim = Dimension::nDim;

  NodeList(std::string name, const GammaLawGas<Dimension>& eos, int64_t numInternal = 0, int64_t numGhost = 0,
           double hmin = 1.0e-20, double hmax = 1.0e20, double hminratio = 0.1, double nPerh = 2.01,
           int maxNumNeighbors = 500, double rhoMin = 1.0e-10, double rhoMax = 1.0e10)
      : mName(std::move(name)), mEOS(eos), mHmin(hmin), mHmax(hmax), mHminratio(hminratio), mNPerh(nPerh),
        mMaxNumNeighbors(maxNumNeighbors), mRhoMin(rhoMin), mRhoMax(rhoMax) {
    numInternalNodes(numInternal);
    numGhostNodes(numGhost);
  }
  const std::string& name() const { return mName; }
  const GammaLawGas<Dimension>& equationOfState() const { return mEOS; }
  int64_t numInternalNodes() const { return mNumInternal; }
  int64_t numGhostNodes() const { return mNumGhost; }
  int64_t numNodes() const { return mNumInternal + mNumGhost; }
  void numInternalNodes(int64_t n) {
    VERIFY2(n >= 0, "negative node count");
    mNumInternal = n;
    resize();
  }
  void numGhostNodes(int64_t n) {
    VERIFY2(n >= 0, "negative node count");
    mNumGhost = n;
    resize();
  }
  double hmin() const { return mHmin; }
  double hmax() const { return mHmax; }
  double hminratio() const { return mHminratio; }
  double nodesPerSmoothingScale() const { return mNPerh; }
  int maxNumNeighbors() const { return mMaxNumNeighbors; }
  double rhoMin() const { return mRhoMin; }
  double rhoMax() const { return mRhoMax; }

  std::vector<double>& mass() { return mMass; }
  std::vector<double>& positions() { return mPos; }
  std::vector<double>& velocity() { return mVel; }
  std::vector<double>& Hfield() { return mH; }
  std::vector<double>& massDensity() { return mRho; }
  std::vector<double>& specificThermalEnergy() { return mEps; }
  const std::vector<double>& mass() const { return mMass; }
  const std::vector<double>& positions() const { return mPos; }
  const std::vector<double>& velocity() const { return mVel; }
  const std::vector<double>& Hfield() const { return mH; }
  const std::vector<double>& massDensity() const { return mRho; }
  const std::vector<double>& specificThermalEnergy() const { return mEps; }

 private:
  void resize() {
    const size_t n = size_t(numNodes());
    mMass.resize(n); mRho.resize(n); mEps.resize(n);
    mPos.resize(n * nDim); mVel.resize(n * nDim);
    mH.resize(n * Dimension::nSym);
  }
  std::string mName;
  GammaLawGas<Dimension> mEOS;
  int64_t mNumInternal = 0, mNumGhost = 0;
  double mHmin, mHmax, mHminratio, mNPerh;
  int mMaxNumNeighbors;
  double mRhoMin, mRhoMax;
  std::vector<double> mMass, mPos, mVel, mH, mRho, mEps;
};

// DataBase of fluid NodeLists (src/DataBase/DataBase.hh:68-75, 161-173).  A
// State field of several NodeLists is the concatenation of each NodeList's
// rows in the order they were appended; with more than one NodeList the ghost
// rows are the device's (the hydro's appended boundaries), so each NodeList
// holds internal rows only.
template <typename Dimension>
class DataBase {
 public:
  void appendNodeList(NodeList<Dimension>& nl) {
    for (const NodeList<Dimension>* p : mNodeLists) VERIFY2(p != &nl, "NodeList already in the DataBase");
    mNodeLists.push_back(&nl);
  }
  int numNodeLists() const { return int(mNodeLists.size()); }
  int numFluidNodeLists() const { return numNodeLists(); }
  const std::vector<NodeList<Dimension>*>& nodeListPtrs() const { return mNodeLists; }
  const std::vector<NodeList<Dimension>*>& fluidNodeListPtrs() const { return mNodeLists; }
  // The first NodeList; with several, the one whose scalar settings (nPerh,
  // hmin, hmax, hminratio, rhoMin, rhoMax, pressure limits) all of them share.
  NodeList<Dimension>& nodeList() const {
    VERIFY2(!mNodeLists.empty(), "no NodeList in the DataBase");
    return *mNodeLists.front();
  }
  int64_t numInternalNodes() const {
    int64_t n = 0;
    for (const NodeList<Dimension>* p : mNodeLists) n += p->numInternalNodes();
    return n;
  }
  int64_t numGhostNodes() const {
    int64_t n = 0;
    for (const NodeList<Dimension>* p : mNodeLists) n += p->numGhostNodes();
    return n;
  }
  int64_t numNodes() const { return numInternalNodes() + numGhostNodes(); }

  // The rows of one NodeList field (picked by `get`, `comps` doubles a row),
  // concatenated over the NodeLists.
  template <typename Get>
  std::vector<double> gather(Get get, int comps) const {
    std::vector<double> v;
    v.reserve(size_t(numNodes()) * size_t(comps));
    for (NodeList<Dimension>* p : mNodeLists) {
      const std::vector<double>& f = get(*p);
      v.insert(v.end(), f.begin(), f.begin() + std::ptrdiff_t(size_t(p->numNodes()) * size_t(comps)));
    }
    return v;
  }
  // The inverse of gather over the internal rows: each NodeList's internal
  // rows from `v`, which holds the internal rows of all of them.
  template <typename Get>
  void scatter(Get get, int comps, const std::vector<double>& v) const {
    VERIFY2(numGhostNodes() == 0 || mNodeLists.size() == 1, "several NodeLists hold internal rows only");
    size_t at = 0;
    for (NodeList<Dimension>* p : mNodeLists) {
      std::vector<double>& f = get(*p);
      const size_t k = size_t(p->numInternalNodes()) * size_t(comps);
      VERIFY2(at + k <= v.size() && f.size() >= k, "field sizes do not match the NodeLists");
      std::copy(v.begin() + std::ptrdiff_t(at), v.begin() + std::ptrdiff_t(at + k), f.begin());
      at += k;
    }
  }

 private:
  std::vector<NodeList<Dimension>*> mNodeLists;
};

template <typename Dimension> struct BSplineKernel {};

// TableKernel(kernel, numPoints) (src/Kernel/TableKernel.hh:29-32): the table
// is built on the device from the B-spline.
template <typename Dimension>
class TableKernel {
 public:
  explicit TableKernel(const BSplineKernel<Dimension>&, int numPoints = 1000) : mNumPoints(numPoints) {
    VERIFY2(numPoints >= 4, "TableKernel needs at least 4 points");
  }
  int numPoints() const { return mNumPoints; }
  double kernelExtent() const { return 2.0; }

 private:
  int mNumPoints;
};

// MonaghanGingoldViscosity(Clinear, Cquadratic, linearInExpansion,
// quadraticInExpansion) (src/ArtificialViscosity/MonaghanGingoldViscosity.hh:28-31);
// epsilon2 defaults to 1e-2 (ArtificialViscosity.cc:64-82).
template <typename Dimension>
class MonaghanGingoldViscosity {
 public:
  MonaghanGingoldViscosity(double Clinear, double Cquadratic, bool linearInExpansion = false,
                           bool quadraticInExpansion = false)
      : mCl(Clinear), mCq(Cquadratic), mLin(linearInExpansion), mQuad(quadraticInExpansion) {}
  double Cl() const { return mCl; }
  double Cq() const { return mCq; }
  bool linearInExpansion() const { return mLin; }
  bool quadraticInExpansion() const { return mQuad; }
  double epsilon2() const { return mEps2; }
  void epsilon2(double v) { mEps2 = v; }

 private:
  double mCl, mCq;
  bool mLin, mQuad;
  double mEps2 = 1.0e-2;
};

// The ideal-H smoothing scale (src/NodeList/SPHSmoothingScale.cc:120-197),
// evaluated on the device inside the derivative pass.
template <typename Dimension> struct SPHSmoothingScale {};

template <typename Dimension>
struct GeomPlane {
  typename Dimension::Vector point, normal;
  GeomPlane(const typename Dimension::Vector& p, const typename Dimension::Vector& n) : point(p), normal(n) {}
};
// ReflectingBoundary(plane) (src/Boundary/ReflectingBoundary.cc:40-175).
template <typename Dimension>
struct ReflectingBoundary {
  GeomPlane<Dimension> plane;
  explicit ReflectingBoundary(const GeomPlane<Dimension>& p) : plane(p) {}
};
// PeriodicBoundary(plane1, plane2) (src/Boundary/PeriodicBoundary.cc:66-135).
template <typename Dimension>
struct PeriodicBoundary {
  GeomPlane<Dimension> plane1, plane2;
  PeriodicBoundary(const GeomPlane<Dimension>& a, const GeomPlane<Dimension>& b) : plane1(a), plane2(b) {}
};

// src/Physics/GenericHydro.hh:32-45
enum class MassDensityType {
  SumDensity = 0, RigorousSumDensity = 1, HybridSumDensity = 2, IntegrateDensity = 3,
  VoronoiCellDensity = 4, SumVoronoiCellDensity = 5, CorrectedSumDensity = 6,
};
enum class HEvolutionType { IdealH = 0, IntegrateH = 1 };

template <typename Dimension> class SynchronousRK2;

template <typename Dimension>
class SPHHydroBase {
 public:
  using Scalar = double;
  using Vector = typename Dimension::Vector;
  using TimeStepType = std::pair<double, std::string>;
  static constexpr int nDim = Dimension::nDim;

  // The reference constructor (src/SPH/SPHHydroBase.hh:53-71).  `device` picks
  // the GPU; `precision` the pair arithmetic (SPHGPU_PRECISION_FP64 keeps the
  // reference's FP64 operation order); `rowOrder` the order of each node's
  // neighbours (SPHGPU_ROW_ORDER_REFERENCE: the reference's, results equal to
  // its scatter loop's; SPHGPU_ROW_ORDER_SEARCH: the faster search order, 1e-12).
  SPHHydroBase(const SPHSmoothingScale<Dimension>& smoothingScaleMethod,
               const MonaghanGingoldViscosity<Dimension>& Q, const TableKernel<Dimension>& W,
               const TableKernel<Dimension>& WPi, double filter, double cfl, bool useVelocityMagnitudeForDt,
               bool compatibleEnergyEvolution, bool evolveTotalEnergy, bool gradhCorrection, bool XSPH,
               bool correctVelocityGradient, bool sumMassDensityOverAllNodeLists, MassDensityType densityUpdate,
               HEvolutionType HUpdate, double epsTensile, double nTensile, const Vector& xmin, const Vector& xmax,
               int device = 0, int precision = SPHGPU_PRECISION_FP64, int rowOrder = SPHGPU_ROW_ORDER_REFERENCE)
      : mQ(Q), mW(W), mCfl(cfl), mUseVelMag(useVelocityMagnitudeForDt), mCompatible(compatibleEnergyEvolution),
        mTotalEnergy(evolveTotalEnergy), mGradh(gradhCorrection), mXSPH(XSPH), mCorrectVG(correctVelocityGradient),
        mSumAll(sumMassDensityOverAllNodeLists), mDensityUpdate(densityUpdate), mEpsTensile(epsTensile),
        mPrecision(precision), mXmin(xmin), mXmax(xmax) {
    (void)smoothingScaleMethod;
    VERIFY2(WPi.numPoints() == W.numPoints(), "the device pass takes WPi = W");
    VERIFY2(filter == 0.0, "the device pass takes filter = 0");
    VERIFY2(nTensile == 4.0, "the device pass takes nTensile = 4");
    VERIFY2(HUpdate == HEvolutionType::IdealH, "the device pass evolves H with IdealH");
    VERIFY2(densityUpdate == MassDensityType::RigorousSumDensity || densityUpdate == MassDensityType::SumDensity,
            "the device pass takes SumDensity or RigorousSumDensity");
    VERIFY2(!(evolveTotalEnergy && compatibleEnergyEvolution), "evolveTotalEnergy excludes compatibleEnergyEvolution");
    check(sphgpu_create(nDim, device, &mSim));
    check(sphgpu_set_table_bspline(mSim, W.numPoints()));
    check(sphgpu_set_row_order(mSim, rowOrder));
    check(sphgpu_set_precision(mSim, precision));
  }
  SPHHydroBase(const SPHHydroBase&) = delete;
  SPHHydroBase& operator=(const SPHHydroBase&) = delete;
  virtual ~SPHHydroBase() {
    if (mSim) sphgpu_destroy(mSim);
  }

  virtual std::string label() const { return "SPHHydroBase"; }

  // Physics::appendBoundary (src/Physics/Physics.hh:80).  The device makes the
  // ghost nodes of the appended boundaries each time it needs them.
  void appendBoundary(const ReflectingBoundary<Dimension>& b) {
    sphgpu_boundary c{};
    c.kind = SPHGPU_BOUNDARY_REFLECTING;
    copyPlane(b.plane, c.plane1);
    mBoundaries.push_back(c);
    check(sphgpu_set_boundaries(mSim, int(mBoundaries.size()), mBoundaries.data()));
  }
  void appendBoundary(const PeriodicBoundary<Dimension>& b) {
    sphgpu_boundary c{};
    c.kind = SPHGPU_BOUNDARY_PERIODIC;
    copyPlane(b.plane1, c.plane1);
    copyPlane(b.plane2, c.plane2);
    mBoundaries.push_back(c);
    check(sphgpu_set_boundaries(mSim, int(mBoundaries.size()), mBoundaries.data()));
  }
  void clearBoundaries() {
    mBoundaries.clear();
    check(sphgpu_set_boundaries(mSim, 0, nullptr));
  }
  size_t numBoundaries() const { return mBoundaries.size(); }

  // SPHHydroBase::initializeProblemStartup (src/SPH/SPHHydroBase.cc:162-200):
  // the hydro's own pressure, sound speed (from each NodeList's equation of
  // state) and grad-h correction (1) fields.
  virtual void initializeProblemStartup(DataBase<Dimension>& db) {
    const NodeList<Dimension>& nl = settings(db);
    const int64_t n = db.numNodes();
    const auto& eos = nl.equationOfState();
    mPressure.assign(size_t(n), 0.0);
    mSoundSpeed.assign(size_t(n), 0.0);
    mOmegaGradh.assign(size_t(n), 1.0);
    // every row, ghosts included, goes through the equation of state
    check(sphgpu_set_nodes(mSim, n, n));
    const auto rho = db.gather([](NodeList<Dimension>& l) -> auto& { return l.massDensity(); }, 1);
    const auto eps = db.gather([](NodeList<Dimension>& l) -> auto& { return l.specificThermalEnergy(); }, 1);
    check(sphgpu_upload(mSim, SPHGPU_F_MASS_DENSITY, rho.data(), n));
    check(sphgpu_upload(mSim, SPHGPU_F_SPECIFIC_THERMAL_ENERGY, eps.data(), n));
    setLists(db);
    check(sphgpu_eos_gamma_law(mSim, eos.gamma(), eos.minimumPressure(), eos.maximumPressure(),
                               eos.minimumPressureType() == MaterialPressureMinType::PressureFloor ? 1 : 0));
    check(sphgpu_download(mSim, SPHGPU_F_PRESSURE, mPressure.data(), n));
    check(sphgpu_download(mSim, SPHGPU_F_SOUND_SPEED, mSoundSpeed.data(), n));
  }

  // SPHHydroBase::registerState (src/SPH/SPHHydroBase.cc:238-336): the
  // NodeLists' fields and the hydro's pressure, sound speed and grad-h
  // correction, under their reference names.
  virtual void registerState(DataBase<Dimension>& db, State<Dimension>& state) {
    using NL = NodeList<Dimension>;
    VERIFY2(mPressure.size() == size_t(db.numNodes()), "initializeProblemStartup first");
    state.enroll(HydroFieldNames::position, db.gather([](NL& l) -> auto& { return l.positions(); }, nDim));
    state.enroll(HydroFieldNames::velocity, db.gather([](NL& l) -> auto& { return l.velocity(); }, nDim));
    state.enroll(HydroFieldNames::mass, db.gather([](NL& l) -> auto& { return l.mass(); }, 1));
    state.enroll(HydroFieldNames::massDensity, db.gather([](NL& l) -> auto& { return l.massDensity(); }, 1));
    state.enroll(HydroFieldNames::specificThermalEnergy,
                 db.gather([](NL& l) -> auto& { return l.specificThermalEnergy(); }, 1));
    state.enroll(HydroFieldNames::H, db.gather([](NL& l) -> auto& { return l.Hfield(); }, Dimension::nSym));
    state.enroll(HydroFieldNames::pressure, mPressure);
    state.enroll(HydroFieldNames::soundSpeed, mSoundSpeed);
    state.enroll(HydroFieldNames::omegaGradh, mOmegaGradh);
  }

  // SPHHydroBase::registerDerivatives (src/SPH/SPHHydroBase.cc:367-425): the
  // fields evaluateDerivatives writes, one value per internal node.
  virtual void registerDerivatives(DataBase<Dimension>& db, StateDerivatives<Dimension>& derivs) {
    const size_t n = size_t(db.numInternalNodes());
    for (const auto& d : derivFields()) derivs.enroll(d.name, n, derivComps(d.id));
  }

  // SPHHydroBase::initialize (src/SPH/SPHHydroBase.cc:432-466): with
  // gradhCorrection, the grad-h corrections of the state's nodes.
  virtual void initialize(Scalar, Scalar, const DataBase<Dimension>& db, State<Dimension>& state,
                          StateDerivatives<Dimension>&) {
    if (!mGradh) return;
    pushState(db, state);
    check(sphgpu_omega_gradh(mSim));
    pullInternal(db, state, HydroFieldNames::omegaGradh, SPHGPU_F_OMEGA_GRADH, 1);
  }

  // SPHHydroBase::evaluateDerivatives (src/SPH/SPHHydroBase.cc:473-928): the
  // state goes to the device, the connectivity is built there (ghosts of the
  // appended boundaries, or the NodeList's own ghost rows), the pass runs and
  // every derivative field is written to `derivs`.  The pair accelerations of
  // compatibleEnergyEvolution stay on the device for SynchronousRK2.
  virtual void evaluateDerivatives(Scalar, Scalar, const DataBase<Dimension>& db, const State<Dimension>& state,
                                   StateDerivatives<Dimension>& derivs) const {
    pushState(db, state);
    sphgpu_hydro_params p = params(settings(db));
    check(sphgpu_evaluate_derivatives(mSim, &p));
    const int64_t n = db.numInternalNodes();
    for (const auto& d : derivFields()) {
      auto& f = derivs.enroll(d.name, size_t(n), derivComps(d.id));
      check(sphgpu_download_derivative(mSim, d.id, f.data(), n));
    }
  }

  // GenericHydro::dt (src/Physics/GenericHydro.cc:114-331): cfl times the
  // smallest of the sound-speed, viscous, divergence, pairwise-velocity and
  // acceleration limits over the internal nodes, from the state and the
  // derivatives passed in (max viscous pressure, velocity gradient, hydro
  // acceleration).
  virtual TimeStepType dt(const DataBase<Dimension>& db, const State<Dimension>& state,
                          const StateDerivatives<Dimension>& derivs, Scalar) const {
    pushState(db, state);
    const int64_t n = db.numInternalNodes();
    const auto& mvp = derivs.field(HydroFieldNames::maxViscousPressure);
    const auto& DvDx = derivs.field(HydroFieldNames::velocityGradient);
    const auto& DvDt = derivs.field(HydroFieldNames::hydroAcceleration);
    VERIFY2(mvp.size() >= size_t(n) && DvDx.size() >= size_t(n) * nDim * nDim && DvDt.size() >= size_t(n) * nDim,
            "derivative fields are too short");
    // per node: max viscous pressure, trace of the velocity gradient, acceleration
    std::vector<double> in(size_t(n) * (2 + nDim));
    for (int64_t i = 0; i < n; ++i) {
      double* r = &in[size_t(i) * (2 + nDim)];
      r[0] = mvp[size_t(i)];
      double tr = 0.0;
      for (int k = 0; k < nDim; ++k) tr += DvDx[size_t(i) * nDim * nDim + size_t(k) * (nDim + 1)];
      r[1] = tr;
      for (int k = 0; k < nDim; ++k) r[2 + k] = DvDt[size_t(i) * nDim + size_t(k)];
    }
    check(sphgpu_set_dt_inputs(mSim, 0, n, in.data()));
    double v = 0.0;
    check(sphgpu_hydro_dt(mSim, settings(db).nodesPerSmoothingScale(), mCfl, mUseVelMag ? 1 : 0, &v));
    return {v, "SPHHydroBase: minimum over the nodes of the sound-speed, viscous, divergence, pairwise-velocity "
               "and acceleration limits"};
  }

  // SPHHydroBase::finalize (src/SPH/SPHHydroBase.cc:964-1105): the summed mass
  // density of the state's connectivity replaces the state's mass density.
  virtual void finalize(Scalar, Scalar, DataBase<Dimension>& db, State<Dimension>& state,
                        StateDerivatives<Dimension>&) {
    pushState(db, state);
    check(sphgpu_sum_density(mSim));
    pullInternal(db, state, HydroFieldNames::massDensity, SPHGPU_F_MASS_DENSITY, 1);
  }

  // Unordered pairs of the last evaluateDerivatives.
  int64_t pairCount() const {
    int64_t n = 0;
    check(sphgpu_pair_count(mSim, &n));
    return n;
  }

  double cfl() const { return mCfl; }
  bool useVelocityMagnitudeForDt() const { return mUseVelMag; }
  bool compatibleEnergyEvolution() const { return mCompatible; }
  bool evolveTotalEnergy() const { return mTotalEnergy; }
  bool gradhCorrection() const { return mGradh; }
  bool XSPH() const { return mXSPH; }
  bool correctVelocityGradient() const { return mCorrectVG; }
  bool sumMassDensityOverAllNodeLists() const { return mSumAll; }
  MassDensityType densityUpdate() const { return mDensityUpdate; }
  HEvolutionType HEvolution() const { return HEvolutionType::IdealH; }
  double epsilonTensile() const { return mEpsTensile; }
  double nTensile() const { return 4.0; }
  const Vector& xmin() const { return mXmin; }
  const Vector& xmax() const { return mXmax; }
  const std::vector<double>& pressure() const { return mPressure; }
  const std::vector<double>& soundSpeed() const { return mSoundSpeed; }
  const std::vector<double>& omegaGradh() const { return mOmegaGradh; }
  const TableKernel<Dimension>& kernel() const { return mW; }
  sphgpu_sim* sim() const { return mSim; }

 private:
  friend class SynchronousRK2<Dimension>;
  struct DerivField { std::string name; int id; };

  static int derivComps(int id) {
    switch (id) {
      case SPHGPU_D_DXDT: case SPHGPU_D_DVDT: case SPHGPU_D_XSPH_DELTA_V: return nDim;
      case SPHGPU_D_DVDX: case SPHGPU_D_LOCAL_DVDX: case SPHGPU_D_M: case SPHGPU_D_LOCAL_M: return nDim * nDim;
      case SPHGPU_D_DHDT: case SPHGPU_D_HIDEAL: case SPHGPU_D_MASS_SECOND_MOMENT: return Dimension::nSym;
      default: return 1;
    }
  }
  // The derivative fields of SPHHydroBase::initializeProblemStartup /
  // registerDerivatives (src/SPH/SPHHydroBase.cc:165-196).
  static const std::vector<DerivField>& derivFields() {
    static const std::vector<DerivField> f = {
        {ReplacePrefix + HydroFieldNames::massDensity, SPHGPU_D_RHO_SUM},
        {HydroFieldNames::normalization, SPHGPU_D_NORMALIZATION},
        {IncrementPrefix + HydroFieldNames::position, SPHGPU_D_DXDT},
        {IncrementPrefix + HydroFieldNames::massDensity, SPHGPU_D_DRHODT},
        {HydroFieldNames::hydroAcceleration, SPHGPU_D_DVDT},
        {IncrementPrefix + HydroFieldNames::specificThermalEnergy, SPHGPU_D_DEPSDT},
        {HydroFieldNames::velocityGradient, SPHGPU_D_DVDX},
        {HydroFieldNames::internalVelocityGradient, SPHGPU_D_LOCAL_DVDX},
        {HydroFieldNames::M_SPHCorrection, SPHGPU_D_M},
        {"local " + HydroFieldNames::M_SPHCorrection, SPHGPU_D_LOCAL_M},
        {IncrementPrefix + HydroFieldNames::H, SPHGPU_D_DHDT},
        {ReplacePrefix + HydroFieldNames::H, SPHGPU_D_HIDEAL},
        {HydroFieldNames::maxViscousPressure, SPHGPU_D_MAX_VISCOUS_PRESSURE},
        {HydroFieldNames::effectiveViscousPressure, SPHGPU_D_EFF_VISCOUS_PRESSURE},
        {HydroFieldNames::viscousWork, SPHGPU_D_VISCOUS_WORK},
        {HydroFieldNames::XSPHWeightSum, SPHGPU_D_XSPH_WEIGHT_SUM},
        {HydroFieldNames::XSPHDeltaV, SPHGPU_D_XSPH_DELTA_V},
        {HydroFieldNames::weightedNeighborSum, SPHGPU_D_WEIGHTED_NEIGHBOR_SUM},
        {HydroFieldNames::massSecondMoment, SPHGPU_D_MASS_SECOND_MOMENT},
    };
    return f;
  }

  void check(int rc) const {
    if (rc != SPHGPU_OK) throw VERIFYError(std::string("sphgpu: ") + sphgpu_last_error(mSim));
  }

  sphgpu_hydro_params params(const NodeList<Dimension>& nl) const {
    sphgpu_hydro_params p{};
    p.Cl = mQ.Cl(); p.Cq = mQ.Cq(); p.eps2 = mQ.epsilon2(); p.epsTensile = mEpsTensile;
    p.nPerh = nl.nodesPerSmoothingScale(); p.hmin = nl.hmin(); p.hmax = nl.hmax(); p.hminratio = nl.hminratio();
    p.XSPH = mXSPH; p.correctVelocityGradient = mCorrectVG; p.evolveTotalEnergy = mTotalEnergy;
    p.linearInExpansion = mQ.linearInExpansion(); p.quadraticInExpansion = mQ.quadraticInExpansion();
    p.compatibleEnergyEvolution = mCompatible; p.precision = mPrecision;
    return p;
  }

  // The NodeList whose scalar settings the device takes: with several
  // NodeLists all of them agree on nPerh, hmin, hmax, hminratio, rhoMin, rhoMax
  // and the pressure limits (the device holds one of each), and hold internal
  // rows only (their ghosts are the device's, from the appended boundaries).
  static const NodeList<Dimension>& settings(const DataBase<Dimension>& db) {
    const NodeList<Dimension>& a = db.nodeList();
    const auto& ea = a.equationOfState();
    for (const NodeList<Dimension>* p : db.nodeListPtrs()) {
      if (p == &a) continue;
      const auto& e = p->equationOfState();
      VERIFY2(a.numGhostNodes() == 0 && p->numGhostNodes() == 0,
              "with several NodeLists the ghost nodes are the hydro boundaries' (none in the NodeLists)");
      VERIFY2(p->nodesPerSmoothingScale() == a.nodesPerSmoothingScale() && p->hmin() == a.hmin() &&
                  p->hmax() == a.hmax() && p->hminratio() == a.hminratio() && p->rhoMin() == a.rhoMin() &&
                  p->rhoMax() == a.rhoMax(),
              "the device takes one nPerh, hmin, hmax, hminratio, rhoMin and rhoMax for all NodeLists");
      VERIFY2(e.minimumPressure() == ea.minimumPressure() && e.maximumPressure() == ea.maximumPressure() &&
                  e.minimumPressureType() == ea.minimumPressureType(),
              "the device takes one pressure range and floor type for all NodeLists");
    }
    return a;
  }
  // Each NodeList's internal rows and gamma to the device (after
  // sphgpu_set_nodes, which returns it to one NodeList, and before the ghosts).
  void setLists(const DataBase<Dimension>& db) const {
    if (db.numNodeLists() == 1) return;
    std::vector<int64_t> counts;
    std::vector<double> gammas;
    for (const NodeList<Dimension>* p : db.nodeListPtrs()) {
      counts.push_back(p->numInternalNodes());
      gammas.push_back(p->equationOfState().gamma());
    }
    check(sphgpu_set_node_lists(mSim, int(counts.size()), counts.data(), gammas.data(), mSumAll ? 1 : 0));
  }

  // The state's fields to the device, then the connectivity.  With appended
  // boundaries the device makes the ghost rows (and culls the unused ones, as
  // the Integrator does by default); otherwise the NodeList's ghost rows are
  // taken as they are.
  void pushState(const DataBase<Dimension>& db, const State<Dimension>& state) const {
    settings(db);
    const int64_t nInt = db.numInternalNodes();
    const int64_t rows = mBoundaries.empty() ? db.numNodes() : nInt;
    check(sphgpu_set_nodes(mSim, nInt, rows));
    struct F { const std::string* name; int id; int comps; };
    const F fields[] = {
        {&HydroFieldNames::position, SPHGPU_F_POSITION, nDim},
        {&HydroFieldNames::velocity, SPHGPU_F_VELOCITY, nDim},
        {&HydroFieldNames::mass, SPHGPU_F_MASS, 1},
        {&HydroFieldNames::massDensity, SPHGPU_F_MASS_DENSITY, 1},
        {&HydroFieldNames::specificThermalEnergy, SPHGPU_F_SPECIFIC_THERMAL_ENERGY, 1},
        {&HydroFieldNames::H, SPHGPU_F_H, Dimension::nSym},
        {&HydroFieldNames::pressure, SPHGPU_F_PRESSURE, 1},
        {&HydroFieldNames::soundSpeed, SPHGPU_F_SOUND_SPEED, 1},
        {&HydroFieldNames::omegaGradh, SPHGPU_F_OMEGA_GRADH, 1}};
    for (const F& f : fields) {
      const auto& v = state.field(*f.name);
      VERIFY2(v.size() >= size_t(rows) * size_t(f.comps), "state field \"" + *f.name + "\" is too short");
      check(sphgpu_upload(mSim, f.id, v.data(), rows));
    }
    setLists(db);
    if (!mBoundaries.empty()) check(sphgpu_set_ghost_nodes(mSim, mW.kernelExtent(), 1));
    else check(sphgpu_build_neighbors(mSim, mW.kernelExtent()));
  }
  // The internal rows of one resident field into the state (its ghost rows stay).
  void pullInternal(const DataBase<Dimension>& db, State<Dimension>& state, const std::string& name, int id,
                    int comps) const {
    const int64_t nInt = db.numInternalNodes();
    auto& v = state.field(name);
    VERIFY2(v.size() >= size_t(nInt) * size_t(comps), "state field \"" + name + "\" is too short");
    check(sphgpu_download(mSim, id, v.data(), nInt));
  }
  static void copyPlane(const GeomPlane<Dimension>& p, sphgpu_plane& c) {
    for (int k = 0; k < 3; ++k) {
      c.point[k] = k < nDim ? p.point[k] : 0.0;
      c.normal[k] = k < nDim ? p.normal[k] : 0.0;
    }
  }

  sphgpu_sim* mSim = nullptr;
  MonaghanGingoldViscosity<Dimension> mQ;
  TableKernel<Dimension> mW;
  double mCfl;
  bool mUseVelMag, mCompatible, mTotalEnergy, mGradh, mXSPH, mCorrectVG, mSumAll;
  MassDensityType mDensityUpdate;
  double mEpsTensile;
  int mPrecision;
  Vector mXmin, mXmax;
  std::vector<sphgpu_boundary> mBoundaries;
  std::vector<double> mPressure, mSoundSpeed, mOmegaGradh;
};

// SynchronousRK2(dataBase, physicsPackages) (src/Integrator/SynchronousRK2.hh:34-35)
// with the Integrator controls (src/Integrator/Integrator.hh:158-221).  The
// step (src/Integrator/SynchronousRK2.cc:78-144: ghosts and connectivity, dt
// vote, two derivative passes, state updates with the SPH policies, summed
// density, boundary enforcement) runs on the device, which holds the state
// between steps: the NodeList and the hydro's pressure / sound speed fields are
// written back after every step(), and after the last step of advance().  After
// changing NodeList fields between steps, call reinitializeDeviceState().
template <typename Dimension>
class SynchronousRK2 {
 public:
  using Scalar = double;
  static constexpr int nDim = Dimension::nDim;

  SynchronousRK2(DataBase<Dimension>& dataBase, const std::vector<SPHHydroBase<Dimension>*>& physicsPackages)
      : mDB(dataBase) {
    VERIFY2(physicsPackages.size() == 1 && physicsPackages[0] != nullptr,
            "the device integrator advances one SPH hydro package");
    mHydro = physicsPackages[0];
  }
  virtual ~SynchronousRK2() = default;

  virtual std::string label() const { return "SynchronousRK2"; }

  // Integrator::step(maxTime): one RK2 step, no further than maxTime.
  virtual void step(Scalar maxTime) {
    stepDevice(maxTime);
    pullState();
  }
  // Integrator::advance(goalTime) (src/Integrator/Integrator.cc:765-770):
  // steps until the current time reaches goalTime.
  virtual void advance(Scalar goalTime) {
    while (currentTime() < goalTime) stepDevice(goalTime);
    pullState();
  }
  // The next step starts from the NodeList's fields (and the hydro's pressure
  // and sound speed) instead of the device-resident state.
  void reinitializeDeviceState() { mResident = false; }

  Scalar currentTime() const { return mTime; }
  void currentTime(Scalar t) { mTime = t; if (mResident) syncClock(); }
  int currentCycle() const { return int(mCycle); }
  void currentCycle(int c) { mCycle = c; if (mResident) syncClock(); }
  Scalar lastDt() const { return mLastDt; }
  void lastDt(Scalar dt) { mLastDt = dt; if (mResident) syncClock(); }
  Scalar dtMin() const { return mDtMin; }
  void dtMin(Scalar v) { mDtMin = v; }
  Scalar dtMax() const { return mDtMax; }
  void dtMax(Scalar v) { mDtMax = v; }
  Scalar dtGrowth() const { return mDtGrowth; }
  void dtGrowth(Scalar v) { mDtGrowth = v; }
  bool cullGhostNodes() const { return mCull; }
  void cullGhostNodes(bool v) { mCull = v; }
  DataBase<Dimension>& accessDataBase() { return mDB; }

 private:
  void check(int rc) const { mHydro->check(rc); }

  sphgpu_integrator_params integratorParams() const {
    const NodeList<Dimension>& nl = SPHHydroBase<Dimension>::settings(mDB);
    const auto& eos = nl.equationOfState();
    sphgpu_integrator_params ip{};
    ip.cfl = mHydro->mCfl;
    ip.dtMin = mDtMin; ip.dtMax = mDtMax; ip.dtGrowth = mDtGrowth;
    ip.gamma = eos.gamma(); ip.Pmin = eos.minimumPressure(); ip.Pmax = eos.maximumPressure();
    ip.pressureFloor = eos.minimumPressureType() == MaterialPressureMinType::PressureFloor ? 1 : 0;
    ip.cullGhostNodes = mCull ? 1 : 0;
    ip.rhoMin = nl.rhoMin(); ip.rhoMax = nl.rhoMax();
    ip.kernelExtent = mHydro->mW.kernelExtent();
    ip.useVelocityMagnitudeForDt = mHydro->mUseVelMag ? 1 : 0;
    ip.gradhCorrection = mHydro->mGradh ? 1 : 0;
    return ip;
  }

  // The NodeLists' fields, concatenated in the DataBase's order, with the
  // hydro's pressure, sound speed and grad-h correction.
  void pushState() {
    VERIFY2(mDB.numGhostNodes() == 0, "the device integrator makes the ghost nodes of the hydro's boundaries");
    VERIFY2(mHydro->mPressure.size() == size_t(mDB.numNodes()), "initializeProblemStartup first");
    SPHHydroBase<Dimension>::settings(mDB);
    sphgpu_sim* s = mHydro->mSim;
    const int64_t n = mDB.numInternalNodes();
    check(sphgpu_set_nodes(s, n, n));
    for (const NodeField& f : nodeFields()) check(sphgpu_upload(s, f.id, mDB.gather(f.get, f.comps).data(), n));
    const auto mass = mDB.gather([](NodeList<Dimension>& l) -> auto& { return l.mass(); }, 1);
    check(sphgpu_upload(s, SPHGPU_F_MASS, mass.data(), n));
    check(sphgpu_upload(s, SPHGPU_F_PRESSURE, mHydro->mPressure.data(), n));
    check(sphgpu_upload(s, SPHGPU_F_SOUND_SPEED, mHydro->mSoundSpeed.data(), n));
    check(sphgpu_upload(s, SPHGPU_F_OMEGA_GRADH, mHydro->mOmegaGradh.data(), n));
    mHydro->setLists(mDB);
    syncClock();
    mResident = true;
  }
  void syncClock() { check(sphgpu_set_time(mHydro->mSim, mTime, mLastDt, mCycle)); }

  void stepDevice(Scalar maxTime) {
    if (!mResident) pushState();
    const sphgpu_hydro_params hp = mHydro->params(SPHHydroBase<Dimension>::settings(mDB));
    const sphgpu_integrator_params ip = integratorParams();
    double dt = 0.0;
    check(sphgpu_rk2_step(mHydro->mSim, &hp, &ip, maxTime, &dt));
    check(sphgpu_get_time(mHydro->mSim, &mTime, &mLastDt, &mCycle));
  }

  // The evolved fields back into each NodeList (its internal rows), and the
  // hydro's pressure, sound speed and grad-h correction.
  void pullState() {
    if (!mResident) return;
    sphgpu_sim* s = mHydro->mSim;
    const int64_t n = mDB.numInternalNodes();
    std::vector<double> v;
    for (const NodeField& f : nodeFields()) {
      v.resize(size_t(n) * size_t(f.comps));
      check(sphgpu_download(s, f.id, v.data(), n));
      mDB.scatter(f.get, f.comps, v);
    }
    check(sphgpu_download(s, SPHGPU_F_PRESSURE, mHydro->mPressure.data(), n));
    check(sphgpu_download(s, SPHGPU_F_SOUND_SPEED, mHydro->mSoundSpeed.data(), n));
    check(sphgpu_download(s, SPHGPU_F_OMEGA_GRADH, mHydro->mOmegaGradh.data(), n));
  }

  // The NodeList fields the step evolves (the mass stays).
  struct NodeField {
    int id, comps;
    std::vector<double>& (*get)(NodeList<Dimension>&);
  };
  static const std::vector<NodeField>& nodeFields() {
    using NL = NodeList<Dimension>;
    static const std::vector<NodeField> f = {
        {SPHGPU_F_POSITION, nDim, [](NL& l) -> std::vector<double>& { return l.positions(); }},
        {SPHGPU_F_VELOCITY, nDim, [](NL& l) -> std::vector<double>& { return l.velocity(); }},
        {SPHGPU_F_MASS_DENSITY, 1, [](NL& l) -> std::vector<double>& { return l.massDensity(); }},
        {SPHGPU_F_SPECIFIC_THERMAL_ENERGY, 1, [](NL& l) -> std::vector<double>& { return l.specificThermalEnergy(); }},
        {SPHGPU_F_H, Dimension::nSym, [](NL& l) -> std::vector<double>& { return l.Hfield(); }}};
    return f;
  }

  DataBase<Dimension>& mDB;
  SPHHydroBase<Dimension>* mHydro = nullptr;
  bool mResident = false;
  double mTime = 0.0, mLastDt = 1.0e-5;
  int64_t mCycle = 0;
  double mDtMin = 0.0, mDtMax = 3.40282347e+38, mDtGrowth = 2.0;
  bool mCull = true;
};

}  // namespace Spheral
