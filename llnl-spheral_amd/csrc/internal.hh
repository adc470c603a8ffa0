// Internal declarations of libsphgpu (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include <vector>
#include <functional>
#include <stdexcept>
#include <tuple>
#include <utility>
#include "../../include/sphgpu.h"

namespace sphgpu {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// Ensemble members (indices in row order) that cannot take the step, with
// their SPHGPU_MEMBER_* reasons: thrown by the neighbour build before any
// state of the step is written, so the others can be stepped without them.
struct MemberFailure {
  std::vector<int> members, reasons;
  void add(int m, int reason) { members.push_back(m); reasons.push_back(reason); }
};

#define SG_CHECK_HIP(expr)                                                                     \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess)                                                                      \
      throw ::sphgpu::Error(SPHGPU_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

#define SG_REQUIRE(cond, code, msg)                 \
  do {                                              \
    if (!(cond)) throw ::sphgpu::Error(code, msg);  \
  } while (0)

// Device buffer with grow-only semantics.
struct DBuf {
  void* p = nullptr;
  size_t bytes = 0;
  DBuf() = default;
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  ~DBuf() { release(); }
  void ensure(size_t nb) {
    if (nb <= bytes) return;
    if (p) SG_CHECK_HIP(hipFree(p));
    p = nullptr; bytes = 0;
    if (nb == 0) return;
    SG_CHECK_HIP(hipMalloc(&p, nb));
    bytes = nb;
  }
  // Grow to at least nb bytes keeping the first `keep` bytes (stream-ordered copy).
  void grow(size_t nb, size_t keep, hipStream_t st) {
    if (nb <= bytes) return;
    void* q = nullptr;
    SG_CHECK_HIP(hipMalloc(&q, nb));
    if (p && keep) SG_CHECK_HIP(hipMemcpyAsync(q, p, keep, hipMemcpyDeviceToDevice, st));
    if (p) {
      SG_CHECK_HIP(hipStreamSynchronize(st));
      SG_CHECK_HIP(hipFree(p));
    }
    p = q;
    bytes = nb;
  }
  void swap(DBuf& o) { std::swap(p, o.p); std::swap(bytes, o.bytes); }
  void release() { if (p) (void)hipFree(p); p = nullptr; bytes = 0; }
  template <class T> T* as() const { return static_cast<T*>(p); }
};

// Tabulated kernel as held on the host: per table point i1 the parabola
// coefficients of W and dW/deta, plus the nperh <-> Wsum tables.
struct TableHost {
  int n = 0;
  double extent = 2.0, step = 0.0;
  std::vector<double> Ak, Bk, Ck, Ag, Bg, Cg, nperh, wsum;
  double kernelValue(double eta, double Hdet) const;
  double gradValue(double eta, double Hdet) const;
};
void buildBSplineTable(int dim, int npoints, TableHost& t);

struct Grid {
  double lo[3] = {0, 0, 0};
  double cell = 1.0;        // cell edge (half the search radius on the staged path)
  double R2 = 0.0;          // squared search radius bound
  int nc[3] = {1, 1, 1};    // cells per axis
  int nb[3] = {1, 1, 1};    // bricks per axis
  int64_t ncells = 0;
  int span = 1;             // stencil half-width in cells (1: cell = R, 2: cell = R/2)
};

struct Timing {
  float neighbors_ms = 0, density_ms = 0, derivs_ms = 0, pack_ms = 0, ghosts_ms = 0, cull_ms = 0;
};

// One planar sub-boundary: the reference's PlanarBoundary with its enter and
// exit planes (a ReflectingBoundary is one with enter = exit; a
// PeriodicBoundary is two, plane1 -> plane2 then plane2 -> plane1).
struct SubBoundary {
  double enterP[3], enterN[3], exitP[3], exitN[3];   // unit normals
  int reflect;                                       // 1: fields mapped by R = I - 2 n n, 0: copied
  int periodicSecond;                                // second half of a periodic pair
  double R[9];                                       // row major, D x D used
};

// Per-member integrator state of an ensemble (device array, host mirror).
struct SegState {
  double t;                 // member time
  double lastDt;            // Integrator::mLastDt
  double tEnd;              // the member runs to this time
  double gamma;             // GammaLawGas gamma
  double dt, hdt;           // this step's dt and dt/2
  unsigned long long voteBits;   // GenericHydro::dt minimum (bits of a positive double)
  double cycle;
  unsigned int fail;        // SPHGPU_MEMBER_* of this step (0: healthy)
  unsigned int failPad;
};

// One member of an ensemble: a contiguous range of internal caller rows.
struct Member {
  int64_t id;
  int64_t first, rows;
};

struct Sim {
  int dim = 3;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string lastError;
  int64_t nInt = 0, nTot = 0;

  TableHost table;
  int64_t tableVersion = 0;   // counts table installs (kernel arguments derive from the table)
  DBuf dTable;              // 6*n interleaved [Ak Bk Ck Ag Bg Cg] per point
  DBuf dTable32;            // FP32 copy, 8 floats per point [Ak Bk Ck 0 Ag Bg Cg 0]
  int precision = SPHGPU_PRECISION_FP64;   // of sum density (evaluate_derivatives takes its own)
  DBuf dNperh, dWsum;

  // node state, caller order
  DBuf pos, vel, mass, rho, eps, P, cs, H, omega, fCl, fCq, fshear;
  bool hasMultipliers = false;

  // neighbour structure (nodes in sorted order; rows in transposed ELL:
  // entry w of slot k at nbr[w * nTot + k], rowCount[k] entries)
  Grid grid;
  double kernelExtent = 2.0;
  double rMax = 1.0;         // largest search radius of the last neighbour build
  double posAbsMax = 0.0;    // largest |position component| of the last neighbour build
  int rowOrder = SPHGPU_ROW_ORDER_REFERENCE;
  bool neighborsValid = false;
  bool isotropic = false;   // every H is h*I (off-diagonals exactly zero)
  DBuf cellOf, cellCount, cellStart, cellFill;
  DBuf perm;                // sorted slot -> caller index
  DBuf iperm;               // caller index -> sorted slot
  DBuf xs, Hs;              // positions / H in sorted order
  DBuf rowCount, nbr;
  int ellCap = 0;           // capacity of the blocked-transposed ELL (multiple of 4)
  int64_t nEntries = 0;
  int maxRow = 0;
  int stagedFallbacks = 0;  // bricks whose candidates exceeded the LDS stage
  DBuf scanTmp, reduceTmp, flags;

  // packed per-node records (sorted order) and outputs (sorted order)
  DBuf rec;
  bool recValid = false;    // records match the caller-order state
  DBuf recM;                // mixed-precision neighbour records (32 / 48 / 48 B per slot in 1 / 2 / 3-D)
  DBuf recD;                // FP64 records of the collapsed isotropic pair loop (96 B per slot, two planes)
  DBuf recS;                // FP64 density-sum neighbour records, isotropic H (48 B per slot)
  int pairRec = 0;           // pair-kernel record current with rec: 1 mixed, 2 FP64 isotropic, 0 none
  int buildRecHint = -1;     // record kinds (pack sec bits, isotropic H) the neighbour build's pack also
                             // writes: -1 the density sum's, else what the caller will read first
  bool recSValid = false;    // recS matches rec (written by the record pack; valid through a cull)
  DBuf recSM;               // mixed density-sum neighbour records (32 B per slot)
  bool recSMValid = false;   // recSM matches rec (written by the record pack; the cull keeps it)
  DBuf pairDv;              // compatible energy: own-frame pair term per row entry (D planes)
  DBuf compDe;              // compatible energy: DepsDt of the policy, per slot
  DBuf compAcc, compDe0;    // compatible energy: DvDt and DepsDt by caller index, ghost rows from the boundaries
  DBuf compRec;             // compatible energy: per slot [v + a hdt (3), |DepsDt0| + epsilon, m, caller index]
  DBuf out;                 // all derivative fields, field-major, sorted order
  int64_t outStride = 0;
  bool derivsValid = false;
  int64_t ncalc = 0;        // unordered pairs evaluated by the last pass

  // planar ghost boundaries generated on the device
  std::vector<SubBoundary> subs;
  bool deviceGhosts = false;     // ghost rows nInt..nTot-1 were made by setGhostNodes
  DBuf ghostControl, ghostSub;   // per ghost: control node (caller index), sub-boundary
  std::vector<int64_t> subStart; // first ghost (offset from nInt) of each sub-boundary, +1 entry
  int64_t ghostsGenerated = 0;   // before culling
  DBuf keep, keepScan, slotKeep, ghostControlSpare, ghostSubSpare, ipermSpare, hmaxOf;
  DBuf genDev;                   // device-held counts of a ghost generation
  DBuf spare[13];                // one per caller-order state field (culling compacts into it)
  bool markGhostUse = false;     // the next build marks the ghost slots rows hold (for culling)
  bool rowsMapOrder = false;     // several NodeLists, reference order: rows in ConnectivityMap order
                                 // (iterateIdealH) instead of the owner loop's delivery order

  // Slots of the current build.  Culling renumbers the caller-order rows but
  // keeps the slot space: culled ghosts become dead slots (perm = -1) that no
  // row holds, so rows and their order are untouched.
  int64_t nSlots = 0;
  DBuf work, workFlag, workScan; // internal slots in slot order (n_internal entries)

  // time integration (SynchronousRK2)
  double time = 0.0;
  double lastDt = 1.0e-5;        // Integrator::mLastDt default
  int64_t cycle = 0;
  DBuf state0[5];                // pos, vel, eps, rho, H of the internal rows at the step start
  DBuf omega0;                   // grad-h: every row's step-start omega (state.assign restores it)
  DBuf dtIn;                     // per internal node: maxViscousPressure, tr DvDx, DvDt (caller order)
  bool haveDtIn = false;

  // Ensemble: independent simulations packed in one node set.  Each member's
  // internal rows are contiguous and in member order; every row (ghosts too)
  // carries its member index in `seg`.  Neighbour grids, ghost selection, the
  // time step and the EOS are per member; the density and pair passes run
  // over all members in one launch.  nSeg = 0: a plain simulation.
  bool ensemble = false;
  int nSeg = 0;
  // Several NodeLists in one node set (not an ensemble): NodeList L owns the
  // internal caller rows [listStart[L], listStart[L+1]); `seg` holds every
  // row's NodeList (ghosts inherit their control's).  Pairs form across
  // NodeLists; each NodeList makes its own ghosts; rhoSum and normalization
  // take same-NodeList pairs; the density sum follows sumAllLists
  // (sumMassDensityOverAllNodeLists); listGamma gives each its gamma-law EOS.
  int nLists = 0;                // 0: one NodeList
  std::vector<int64_t> listStart;
  std::vector<double> listGammaHost;
  DBuf listGamma;                // per NodeList gamma (device), empty: the integrator's gamma
  bool sumAllLists = true;
  DBuf seg;                      // member index per caller row (as a double)
  DBuf slotList;                 // several NodeLists, mixed precision: NodeList of every slot (-1: dead)
  DBuf segGrid;                  // neighbour grid list (levels, members), in one buffer:
  void* gridList = nullptr;      //   the grids (DGrid, neighbors.hip),
  int64_t* gridCellBase = nullptr;  // their first global cells (+1 entry),
  int32_t* gridFirst = nullptr;  //   the first grid of each member (+1 entry)
  DBuf segBoundsBits;            // per member: 9 ordered-bit bounds of the neighbour build
  DBuf segHmaxBits;              // per member: ghost-selection hmax bits
  DBuf segState;                 // SegState per member
  DBuf plainSeg;                 // SegState of a plain simulation's step (device-side dt)
  std::vector<SegState> segHost;
  std::vector<Member> members;
  int64_t nextMemberId = 0;
  double dtThreshold = 0.0;      // a member whose dt vote falls below this fails (0: never)
  std::vector<int64_t> failedIds;   // members evicted by the last step, with their reasons
  std::vector<int32_t> failedReasons;
  std::vector<int64_t> failedCycles;   // steps each had taken
  DBuf segSpare, dtInSpare;

  // HIP graph of the device part of a time step (dt vote, the two passes and
  // updates, finalize, enforcement, clock advance): captured once its launch
  // key (extents, buffers, parameters, host flags) repeats, then replayed.
  bool useGraphs = false;
  bool capturing = false;          // the step is being captured: no stage events
  struct StepFlags {
    bool recValid, recSValid, recSMValid, derivsValid, pairsValid, haveDtIn;
    int pairRec;
    int64_t outStride;
    unsigned timingPending;
  };
  struct StepGraph {
    std::vector<uint64_t> key;
    hipGraphExec_t exec = nullptr;
    StepFlags flags{};             // the host flags the step leaves
  };
  std::vector<StepGraph> stepGraphs;              // most recent last (a few: culling alternates buffers)
  std::vector<std::vector<uint64_t>> recentKeys;  // keys of the last eager steps
  int64_t graphCaptures = 0, graphReplays = 0;
  hipEvent_t ev[14] = {};          // 0-11 stage timings, 12-13 the copy probe
  Timing timing;
  // Stage timings and the pair count are read back lazily (no host wait
  // inside a step): bit 0 neighbours, 1 density, 2 derivatives, 3 ghosts.
  unsigned timingPending = 0;
  unsigned long long* ncalcHost = nullptr;   // pinned; written by the stream after each pass
  // Pinned host staging: small readbacks (copies into pageable memory would
  // each wait for the stream) and uploads the stream reads after the call
  // returns (the grid lists of a neighbour build).  Grown only after a sync.
  void* pinnedOut = nullptr;
  size_t pinnedOutBytes = 0;
  void* pinnedIn = nullptr;
  size_t pinnedInBytes = 0;
  hipEvent_t pinnedInDone = nullptr;   // recorded after the copies that read pinnedIn
  DBuf pick;                     // device side of gathered readback values
  std::vector<int32_t> cullStarts;  // new ghost starts read with the last build's row readback
  bool cullStartsReady = false;
  bool pairsValid = false;                   // a pass ran on the current neighbour rows
};
void resolveTimings(Sim& s);

// neighbors.hip
void buildNeighbors(Sim& s, double kernelExtent);
// hydro.hip
// sec: pack sec bits of the records written with the FP64 record (isotropic H)
void packRecords(Sim& s, int sec = 0);
void sumDensity(Sim& s);
// updateOnly: store only DxDt, DvDt, DepsDt, DrhoDt and Hideal (the fields an
// integrator update reads); the other fields keep what they held
void evalDerivs(Sim& s, const sphgpu_hydro_params& p, bool updateOnly = false);
// The pack sec bits of the records the pair kernel and the density sum read.
int pairRecordSec(const Sim& s, const sphgpu_hydro_params& hp);
int densityRecordSec(const Sim& s);
// probe.hip
double copyProbe(Sim& s, int64_t bytes, int repeats);
double gatherProbe(Sim& s, int loads, int repeats);
// ghosts.hip
void setBoundaries(Sim& s, int n, const sphgpu_boundary* b);
void setGhostNodes(Sim& s, double kernelExtent, bool cull);
void applyGhostBoundaries(Sim& s);
void applyGhostDensity(Sim& s);   // after sum density: ghost rho and their records
void applyGhostScalar(Sim& s, DBuf& field);   // ghost rows of one scalar field from their controls
void applyGhostVector(Sim& s, DBuf& field);   // ghost rows of one Vector field (R.a or a copy)
void ensureStateRows(Sim& s, int64_t rows);   // room for `rows` rows in every caller-order field (keeps nTot)
bool stateFieldUsed(const Sim& s, int which); // caller-order fields, in culling order
DBuf& stateField(Sim& s, int which);
int stateFieldComps(const Sim& s, int which);
constexpr int kStateFields = 13;
// hydro.hip
void omegaGradh(Sim& s);          // computeSPHOmegaGradhCorrection into the omega field
// one iterateIdealH pass over the current rows; returns the largest change
double idealHPass(Sim& s, const sphgpu_hydro_params& hp, double tol, double* deltaH, unsigned long long* maxBits);
void conservationTotals(Sim& s, int64_t r0, int64_t r1, double* out);   // integrate.hip, 9 values
// The same totals into device memory, asynchronously, with the partial sums in
// the library's static scratch array (no allocation); resolveScratch(device)
// must have run once on the device (sphgpu_create does it).
void conservationTotalsNoAlloc(Sim& s, int64_t r0, int64_t r1, double* outDevice);
void resolveScratch(int device);
// integrate.hip: iterateIdealH; returns the iterations done, *maxDelta the last largest change
int iterateIdealH(Sim& s, const sphgpu_hydro_params& hp, int maxIterations, double tol, double kernelExtent,
                  bool cull, double* maxDelta);
void enforceBoundaries(Sim& s);   // internal nodes behind a plane are mapped back
// integrate.hip
void eosGammaLaw(Sim& s, double gamma, double Pmin, double Pmax, int floorType);
double hydroDt(Sim& s, double nPerh, double cfl, int useVelocityMagnitude);
double rk2Step(Sim& s, const sphgpu_hydro_params& hp, const sphgpu_integrator_params& ip, double maxTime);
// ensemble.hip
int64_t ensembleAdd(Sim& s, const sphgpu_member& m, const sphgpu_integrator_params& ip);
void ensembleRemove(Sim& s, int nIds, const int64_t* ids);
int64_t ensembleBytesPerNode(int dim);
// serve.cpp
void serve(Sim& s, sphgpu_request* req, int n, const sphgpu_hydro_params& hp, const sphgpu_integrator_params& ip,
           const sphgpu_serve_params& sp);
int ensembleStep(Sim& s, const sphgpu_hydro_params& hp, const sphgpu_integrator_params& ip);
// ensemble update / dt (integrate.hip)
void ensembleRk2(Sim& s, const sphgpu_hydro_params& hp, const sphgpu_integrator_params& ip);
void eosRows(Sim& s, int64_t first, int64_t rows, double gamma, double Pmin, double Pmax, int floorType);
// scans (neighbors.hip)
// Up to four int32 buffers zeroed by one kernel (a graph-safe memset that
// also saves dispatches).
// Pinned staging of at least `bytes`: which = 0 readbacks, 1 uploads.
void* hostStaging(Sim& s, int which, size_t bytes);
// Gathers up to 16 int32 values of a device array into s.pick and copies them
// to pinned host memory, one round trip: returns the host copy.
const int32_t* readInts(Sim& s, const int32_t* src, const int64_t* at, int m);
// The same gather with the copy to pinned `hostOut` left in flight.
void pickInts(Sim& s, const int32_t* src, const int64_t* at, int m, int32_t* hostOut);
// First half of the ghost cull (keep flags, scan, new ghost starts to pinned
// hostOut); the cull takes the starts from s.cullStarts when the neighbour
// build has read them.
void cullLaunch(Sim& s, int32_t* hostOut);

struct ZeroSet {
  int32_t* p[4];
  int64_t n[4];
  int m;
  void add(int32_t* ptr, int64_t count) {
    if (m < 4 && count > 0) { p[m] = ptr; n[m] = count; ++m; }
  }
};
void zeroMany(const ZeroSet& z, hipStream_t st);
void exclusiveScanI32toI64(const int32_t* in, int64_t* out, int64_t n, Sim& s);
void exclusiveScanI32(const int32_t* in, int32_t* out, int64_t n, Sim& s);

// tune.cpp: launch-shape autotuner.  A workload class of one kernel family;
// tunedVariant returns the variant to launch, timing every variant once (on
// the live inputs, launched by `launch`) when tuning is on and the class is
// new; otherwise the recorded choice, or `fallback`.
struct TuneKey {
  int kernel, dim, precision, rowBucket, sizeBucket;
};
bool operator<(const TuneKey& a, const TuneKey& b);
enum { kTunePairIso = 0 };
TuneKey tuneKey(int kernel, int dim, int precision, double rowLength, int64_t nodes);
void tuneEnable(bool on);
int tunedVariant(Sim& s, const TuneKey& key, int nVariants, int fallback, const std::function<void(int)>& launch,
                 bool* timed = nullptr);
std::string tuneExport();
int tuneImport(const std::string& text);

inline int nsymOf(int dim) { return dim * (dim + 1) / 2; }
// Row groups with their own ghost selection: ensemble members, or NodeLists.
inline int groupCount(const Sim& s) { return s.nSeg > 0 ? s.nSeg : (s.nLists > 1 ? s.nLists : 0); }
void setNodeLists(Sim& s, int n, const int64_t* counts, const double* gamma, int sumAll);   // ghosts.hip

// Number of doubles per node of each derivative output field, in the order of
// the SPHGPU_D_* ids (see sphgpu.h).
int derivComponents(int dim, int field);
int derivOffsetDoubles(int dim, int field);   // offset of a field's block, per node
int derivTotalDoubles(int dim);

// Record layout shared by the record packer and the pair kernels: one 128-byte
// line per node for the isotropic-H path (H = h I), two lines otherwise.
// Fields: [x y z | h^-1 or H | m Hdet orig rho] [v | cs Prho m/rho P som fCl fCq fsh]
template <int D> struct RecLayout;

}  // namespace sphgpu
