/*
 * sphgpu.h -- C-ABI of the MI355X-native SPH derivative pass (libsphgpu.so).
 *
 * The reference exposes this path only as C++ virtuals on Physics<Dim>
 * (src/Physics/Physics.hh:51-75) implemented by SPHHydroBase<Dim>
 * (src/SPH/SPHHydroBase.hh:53-71), fed by ConnectivityMap / NestedGridNeighbor
 * (src/Neighbor/ConnectivityMap.hh:46-121, src/Neighbor/NestedGridNeighbor.hh:45-51)
 * and TableKernel (src/Kernel/TableKernel.hh:29-67).  This header is the plain-C
 * layer under a GPU hydro package with the same signatures: plain pointers and
 * sizes, int status codes, no C++ types.
 *
 * Conventions
 *  - One sphgpu_sim per NodeList on one GPU.  Calls on distinct sims are
 *    thread-safe; calls on one sim are not.  Every sim owns one HIP stream;
 *    compute calls are asynchronous on it unless stated, downloads synchronise.
 *  - Node order is the caller's (the reference's Field order): internal nodes
 *    0..n_internal-1, then ghost nodes up to n_total-1.
 *  - Per-node values use Spheral's in-memory element layout: Vector = D
 *    doubles, SymTensor = D(D+1)/2 doubles (xx,xy,xz,yy,yz,zz), Tensor = D*D
 *    doubles (row major).  Host buffers are caller-owned; device state is owned
 *    by the sim and stays resident between calls.
 *  - Every function returns 0 (SPHGPU_OK) or a negative SPHGPU_ERR_* code; the
 *    message of the last failure is sphgpu_last_error(sim).  No exception crosses
 *    the ABI.
 */
#ifndef SPHGPU_H
#define SPHGPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  SPHGPU_OK = 0,
  SPHGPU_ERR_ARG = -1,       /* bad argument (null, size, dimension)          */
  SPHGPU_ERR_STATE = -2,     /* call out of order (e.g. no neighbours yet)    */
  SPHGPU_ERR_HIP = -3,       /* HIP runtime failure                           */
  SPHGPU_ERR_UNSUPPORTED = -4,/* option not implemented on the device path     */
  SPHGPU_ERR_CAPACITY = -5   /* a row or cell exceeded a device capacity      */
};

/* Node state fields, named as in src/Hydro/HydroFieldNames.cc:11-63. */
enum {
  SPHGPU_F_POSITION = 0,               /* "position"                Vector    */
  SPHGPU_F_VELOCITY = 1,               /* "velocity"                Vector    */
  SPHGPU_F_MASS = 2,                   /* "mass"                    Scalar    */
  SPHGPU_F_MASS_DENSITY = 3,           /* "mass density"            Scalar    */
  SPHGPU_F_SPECIFIC_THERMAL_ENERGY = 4,/* "specific thermal energy" Scalar    */
  SPHGPU_F_H = 5,                      /* "H"                       SymTensor */
  SPHGPU_F_PRESSURE = 6,               /* "pressure"                Scalar    */
  SPHGPU_F_SOUND_SPEED = 7,            /* "sound speed"             Scalar    */
  SPHGPU_F_OMEGA_GRADH = 8,            /* "grad h corrections"      Scalar    */
  SPHGPU_F_CL_MULTIPLIER = 9,          /* Q Cl multiplier           Scalar    */
  SPHGPU_F_CQ_MULTIPLIER = 10,         /* Q Cq multiplier           Scalar    */
  SPHGPU_F_SHEAR_CORRECTION = 11,      /* Balsara shear correction  Scalar    */
  SPHGPU_NUM_STATE_FIELDS = 12
};

/* Derivative fields written by sphgpu_evaluate_derivatives, in the order
 * SPHHydroBase::evaluateDerivatives fetches them (src/SPH/SPHHydroBase.cc:521-540). */
enum {
  SPHGPU_D_RHO_SUM = 0,                /* "new mass density"              Scalar    */
  SPHGPU_D_NORMALIZATION = 1,          /* "normalization"                 Scalar    */
  SPHGPU_D_DXDT = 2,                   /* "delta position"                Vector    */
  SPHGPU_D_DRHODT = 3,                 /* "delta mass density"            Scalar    */
  SPHGPU_D_DVDT = 4,                   /* "delta velocity hydro"          Vector    */
  SPHGPU_D_DEPSDT = 5,                 /* "delta specific thermal energy" Scalar    */
  SPHGPU_D_DVDX = 6,                   /* "velocity gradient"             Tensor    */
  SPHGPU_D_LOCAL_DVDX = 7,             /* "internal velocity gradient"    Tensor    */
  SPHGPU_D_M = 8,                      /* "M SPH gradient correction"     Tensor    */
  SPHGPU_D_LOCAL_M = 9,                /* "local M ..."                   Tensor    */
  SPHGPU_D_DHDT = 10,                  /* "delta H"                       SymTensor */
  SPHGPU_D_HIDEAL = 11,                /* "new H"                         SymTensor */
  SPHGPU_D_MAX_VISCOUS_PRESSURE = 12,  /* "max viscous pressure"          Scalar    */
  SPHGPU_D_EFF_VISCOUS_PRESSURE = 13,  /* "effective viscous pressure"    Scalar    */
  SPHGPU_D_VISCOUS_WORK = 14,          /* "viscous work rate"             Scalar    */
  SPHGPU_D_XSPH_WEIGHT_SUM = 15,       /* "XSPH weight sum"               Scalar    */
  SPHGPU_D_XSPH_DELTA_V = 16,          /* "XSPH delta vi"                 Vector    */
  SPHGPU_D_WEIGHTED_NEIGHBOR_SUM = 17, /* "weighted neighbor sum"         Scalar    */
  SPHGPU_D_MASS_SECOND_MOMENT = 18,    /* "mass second moment"            SymTensor */
  SPHGPU_NUM_DERIV_FIELDS = 19
};

/* Arithmetic of the pair pass. */
enum {
  SPHGPU_PRECISION_FP64 = 0,  /* FP64 throughout, reference summation order  */
  SPHGPU_PRECISION_MIXED = 1  /* FP32 pair kinematics and kernel lookups, FP64
                                 positions and per-node accumulators; needs
                                 H = h I, no Q multipliers, no tensile term */
};

/* Options of SPHHydroBase (src/SPH/SPHHydroBase.hh:53-71, defaults of the SPH()
 * factory src/SPH/SPHHydros.py:317-339) and MonaghanGingoldViscosity
 * (src/ArtificialViscosity/MonaghanGingoldViscosity.hh:28-31,
 *  src/ArtificialViscosity/ArtificialViscosity.cc:64-82) that the pass reads,
 * plus the FluidNodeList smoothing-scale limits (src/NodeList/FluidNodeLists.py:12-31). */
typedef struct sphgpu_hydro_params {
  double Cl;                   /* Clinear (1 for the B-spline SPH() default)   */
  double Cq;                   /* Cquadratic                                   */
  double eps2;                 /* Q epsilon2 (1e-2)                            */
  double epsTensile;           /* tensile correction (0)                       */
  double nPerh;                /* nodes per smoothing scale                    */
  double hmin, hmax, hminratio;
  int XSPH;                    /* 1 */
  int correctVelocityGradient; /* 1 */
  int evolveTotalEnergy;       /* 0 */
  int linearInExpansion;       /* 0 */
  int quadraticInExpansion;    /* 0 */
  int compatibleEnergyEvolution; /* 0; 1: the pass keeps every row entry's
                                    pair term for the RK2 step's
                                    SpecificThermalEnergyPolicy (FP64, or
                                    mixed precision with the pair terms
                                    computed in FP32, exact negatives at the
                                    two ends of a pair) */
  int precision;               /* SPHGPU_PRECISION_* */
} sphgpu_hydro_params;

typedef struct sphgpu_sim sphgpu_sim;

/* Create a simulation context for a Dim<dim> NodeList on HIP device `device`.
 * Replaces constructing SPHHydroBase<Dim> (src/SPH/SPHHydroBase.cc:68-160). */
int sphgpu_create(int dim, int device, sphgpu_sim** out);
/* Release all device memory and the stream. */
void sphgpu_destroy(sphgpu_sim* sim);
/* Message of the last failed call on this sim (never null). */
const char* sphgpu_last_error(const sphgpu_sim* sim);

/* Install a tabulated kernel from the host TableKernel: `coeffs` holds 8 arrays
 * of npoints doubles, [a b c] of W, [a b c] of dW/deta, the nperh values and the
 * Wsum values (src/Kernel/TableKernel.hh:44-67, TableKernel.cc:469-548). */
int sphgpu_set_table(sphgpu_sim* sim, int npoints, double extent, const double* coeffs);
/* Build TableKernel(BSplineKernel<dim>(), npoints) on the host and install it
 * (src/Kernel/TableKernel.cc:211-305). */
int sphgpu_set_table_bspline(sphgpu_sim* sim, int npoints);
/* Read back the installed table in the sphgpu_set_table layout. */
int sphgpu_get_table(sphgpu_sim* sim, int npoints, double* coeffs);
/* W(eta, Hdet) and dW/deta(eta, Hdet) from the installed table on the host
 * (src/Kernel/TableKernelInline.hh:11-38). */
int sphgpu_table_values(sphgpu_sim* sim, double eta, double Hdet, double* W, double* gradW);

/* Size the node arrays (NodeList::numInternalNodes / numGhostNodes,
 * src/NodeList/NodeListInline.hh:28-42).  Invalidates neighbours. */
int sphgpu_set_nodes(sphgpu_sim* sim, int64_t n_internal, int64_t n_total);
/* Several NodeLists in one sim (a DataBase of fluid NodeLists,
 * src/DataBase/DataBaseInline.hh:230-258): NodeList L owns the next
 * internal_counts[L] internal rows, NodeLists in order (the counts add up to
 * n_internal; call after sphgpu_set_nodes with n_total = n_internal, the
 * device boundaries make each NodeList's ghosts with its own hmax,
 * PlanarBoundary::setGhostNodes(nodeList), src/Boundary/PlanarBoundary.cc:
 * 112-182).  Pairs form across NodeLists as within one, each owned as
 * ConnectivityMap::calculatePairInteraction decides
 * (src/Neighbor/ConnectivityMapInline.hh:169-186); REFERENCE row order walks
 * each row in the order the reference's loops deliver the pair terms, so the
 * FP64 pass is bitwise the reference's.  evaluateDerivatives sums rhoSum and
 * normalization over same-NodeList pairs only (src/SPH/SPHHydroBase.cc:746-751);
 * the density sum counts a pair across NodeLists only with
 * sum_over_all_node_lists (the SPH() default, src/SPH/SPHHydros.py:332), each
 * end then taking its own mass times the other end's kernel
 * (src/SPH/computeSPHSumMassDensity.cc:77-100).  gamma (n values, may be
 * null) gives each NodeList its gamma-law EOS for sphgpu_eos_gamma_law and the
 * time step.  The grad-h correction sums over every NodeList alike
 * (src/SPH/computeSPHOmegaGradhCorrection.cc:55-130).  n = 1 returns to one
 * NodeList.  Compatible energy replays each node's pair terms in the order
 * the owner loop delivers them (reference row order); sphgpu_iterate_ideal_h
 * weights a neighbour of another NodeList by m_j rho_i / (m_i rho_j) and walks
 * rows in ConnectivityMap order (src/Utilities/iterateIdealH.cc:164-205).
 * Mixed precision applies the same NodeList rules in its FP32 kernels.  Not
 * in ensembles (SPHGPU_ERR_STATE). */
int sphgpu_set_node_lists(sphgpu_sim* sim, int n, const int64_t* internal_counts, const double* gamma,
                          int sum_over_all_node_lists);
/* Copy one state field host -> device: n_total rows, or, when the ghosts were
 * made by sphgpu_set_ghost_nodes, n_internal rows (ghost rows then keep their
 * values until sphgpu_apply_ghost_boundaries). */
int sphgpu_upload(sphgpu_sim* sim, int field, const double* host, int64_t n_rows);
/* Copy one state field (n_total rows) device -> host; synchronises. */
int sphgpu_download(sphgpu_sim* sim, int field, double* host, int64_t n_rows);

/* Planar ghost boundaries generated on the device.
 * SPHGPU_BOUNDARY_REFLECTING: ReflectingBoundary(plane1)
 *   (src/Boundary/ReflectingBoundary.cc:40-175): ghosts mirrored through the
 *   plane; vectors and tensors mapped by R = I - 2 n n
 *   (src/Utilities/planarReflectingOperator.hh), scalars copied.
 * SPHGPU_BOUNDARY_PERIODIC: PeriodicBoundary(plane1, plane2)
 *   (src/Boundary/PeriodicBoundary.cc:66-135): two planar halves, plane1 ->
 *   plane2 then plane2 -> plane1, values copied.
 * Normals point into the domain and are normalised as GeomPlane does
 * (src/Geometry/GeomPlane.cc:41-46).  Components beyond dim are ignored. */
enum { SPHGPU_BOUNDARY_REFLECTING = 0, SPHGPU_BOUNDARY_PERIODIC = 1 };
typedef struct sphgpu_plane { double point[3]; double normal[3]; } sphgpu_plane;
typedef struct sphgpu_boundary { int kind; sphgpu_plane plane1; sphgpu_plane plane2; } sphgpu_boundary;

/* Register the boundaries in the order the physics package holds them
 * (Physics::boundaryBegin/End, src/Physics/Physics.hh:98-110); n = 0 clears. */
int sphgpu_set_boundaries(sphgpu_sim* sim, int n, const sphgpu_boundary* boundaries);
/* Integrator::setGhostNodes (src/Integrator/Integrator.cc:443-595): drop old
 * ghosts, let each boundary in order select its controls among all current
 * nodes (PlanarBoundary::setGhostNodes, src/Boundary/PlanarBoundary.cc:112-182)
 * and append their ghosts, fill every ghost state field
 * (applyGhostBoundaries), build the neighbours of the internal nodes
 * (as sphgpu_build_neighbors), and, if `cull`, remove the ghosts that no
 * internal node sees and that control no other ghost, keeping the order of the
 * rest (src/Boundary/Boundary.cc:182-233).  The internal rows of every state
 * field must be uploaded first (n_total = n_internal). */
int sphgpu_set_ghost_nodes(sphgpu_sim* sim, double kernel_extent, int cull);
/* Integrator::applyGhostBoundaries (src/Integrator/Integrator.cc:600-650):
 * refresh positions, H and every state field of the ghosts from their control
 * nodes, boundary by boundary, with the current ghost set. */
int sphgpu_apply_ghost_boundaries(sphgpu_sim* sim);
/* Current node counts (NodeList::numInternalNodes / numNodes). */
int sphgpu_node_counts(sphgpu_sim* sim, int64_t* n_internal, int64_t* n_total);
/* Control node (caller index) and planar sub-boundary of every ghost
 * (n_total - n_internal entries each; Boundary::controlNodes, ghostNodes). */
int sphgpu_download_ghost_controls(sphgpu_sim* sim, int32_t* control, int32_t* sub_boundary);

/* Order of each node's neighbour row inside the device passes.
 *  REFERENCE: ascending node index, the order ConnectivityMap keeps
 *    (src/Neighbor/ConnectivityMap.cc:735-738); the FP64 pass then sums in the
 *    reference's order and reproduces its results bit for bit.
 *  SEARCH: the deterministic order of the device search (no per-row sort);
 *    FP64 results differ from the reference only by summation order.
 * sphgpu_download_neighbors returns rows in ascending order in both modes. */
enum { SPHGPU_ROW_ORDER_REFERENCE = 0, SPHGPU_ROW_ORDER_SEARCH = 1 };
/* Select the row order of the next neighbour build (default REFERENCE).  Like
 * ConnectivityMap's domainDecompositionIndependent ordering switch
 * (src/Neighbor/ConnectivityMap.cc:729-740); invalidates the current build. */
int sphgpu_set_row_order(sphgpu_sim* sim, int order);

/* ConnectivityMap::rebuild (src/Neighbor/ConnectivityMapInline.hh:39-74,
 * computeConnectivity src/Neighbor/ConnectivityMap.cc:521-890): rows for the
 * internal nodes, j in N(i) iff |H_i r_ij|^2 <= ext^2 or |H_j r_ij|^2 <= ext^2,
 * j != i, rows ascending in node index.  Built from resident positions and H. */
int sphgpu_build_neighbors(sphgpu_sim* sim, double kernel_extent);
/* Total number of (directed) neighbour entries of the last build
 * (sum of ConnectivityMap::numNeighborsForNode, ConnectivityMap.hh:86-90). */
int sphgpu_neighbor_count(sphgpu_sim* sim, int64_t* n_entries);
/* CSR download in caller indices: offsets[n_internal+1], neighbors[n_entries]
 * (ConnectivityMap::connectivityForNode, ConnectivityMap.hh:67-79). */
int sphgpu_download_neighbors(sphgpu_sim* sim, int64_t* offsets, int32_t* neighbors);

/* Arithmetic of sphgpu_sum_density (SPHGPU_PRECISION_*, default FP64).  No
 * counterpart in the reference, which is FP64 only; evaluate_derivatives takes
 * its precision from sphgpu_hydro_params. */
int sphgpu_set_precision(sphgpu_sim* sim, int precision);

/* computeSPHSumMassDensity (src/SPH/computeSPHSumMassDensity.cc:25-107) over the
 * current neighbours; overwrites the resident "mass density" of internal nodes.
 * MIXED: FP32 kernel lookups, FP64 positions and sums (isotropic H only). */
int sphgpu_sum_density(sphgpu_sim* sim);

/* SPHHydroBase::evaluateDerivatives (src/SPH/SPHHydroBase.cc:471-928) with
 * MonaghanGingoldViscosity::Piij fused (MonaghanGingoldViscosity.cc:63-133) and
 * the SPHSmoothingScale epilogue (src/NodeList/SPHSmoothingScale.cc:102-227).
 * Results stay resident; fetch them with sphgpu_download_derivative. */
int sphgpu_evaluate_derivatives(sphgpu_sim* sim, const sphgpu_hydro_params* params);
/* Copy one derivative field for the internal nodes (n_internal rows). */
int sphgpu_download_derivative(sphgpu_sim* sim, int field, double* host, int64_t n_rows);
/* Number of unordered pairs of the last evaluateDerivatives (the reference's
 * per-node ncalc summed, src/SPH/SPHHydroBase.cc:599, 674). */
int sphgpu_pair_count(sphgpu_sim* sim, int64_t* n_pairs);

/* Time integration.  Options of SynchronousRK2 / Integrator
 * (src/Integrator/Integrator.cc:44-62: dtMin 0, dtMax FLT_MAX, dtGrowth 2),
 * GenericHydro (cfl 0.25, useVelocityMagnitudeForDt 0; src/SPH/SPHHydros.py:17-47),
 * GammaLawGas (src/Material/GammaLawGas.cc:163-223, pressure limits -1e200 /
 * 1e200, src/Material/MaterialEquationsOfState.py:20-30) and the FluidNodeList
 * density limits (rhoMin 1e-10, rhoMax 1e10, src/NodeList/FluidNodeLists.py:12-31). */
typedef struct sphgpu_integrator_params {
  double cfl;
  double dtMin, dtMax, dtGrowth;
  double gamma;                /* GammaLawGas gamma                                   */
  double Pmin, Pmax;           /* applyPressureLimits bounds                         */
  int pressureFloor;           /* 1: P < Pmin -> Pmin (PressureFloor), 0: -> 0       */
  int cullGhostNodes;          /* Integrator::cullGhostNodes (1)                     */
  double rhoMin, rhoMax;       /* IncrementBoundedState limits of the mass density   */
  double kernelExtent;         /* TableKernel extent (2)                             */
  int useVelocityMagnitudeForDt;
  int gradhCorrection;         /* SPHHydroBase gradhCorrection: omega before each derivative pass */
} sphgpu_integrator_params;

/* computeSPHOmegaGradhCorrection (src/SPH/computeSPHOmegaGradhCorrection.cc:26-133)
 * over the current neighbours into the resident "grad h corrections" field,
 * ghosts refreshed from their controls (SPHHydroBase::initialize,
 * src/SPH/SPHHydroBase.cc:433-455). */
int sphgpu_omega_gradh(sphgpu_sim* sim);

/* GammaLawGas::setPressure / setSoundSpeed on the internal rows from the
 * resident mass density and specific thermal energy
 * (src/Material/GammaLawGas.cc:163-223, EquationOfStateInline.hh:71-78). */
int sphgpu_eos_gamma_law(sphgpu_sim* sim, double gamma, double Pmin, double Pmax, int pressureFloor);
/* GenericHydro::dt (src/Physics/GenericHydro.cc:114-331): cfl times the
 * smallest sound-speed, viscous, divergence, pairwise-velocity and
 * acceleration limit over the internal nodes, with the current neighbours and
 * the derivatives of the last step (zero before the first). */
int sphgpu_hydro_dt(sphgpu_sim* sim, double nPerh, double cfl, int useVelocityMagnitudeForDt, double* dt);
/* SynchronousRK2::step (src/Integrator/SynchronousRK2.cc:78-144) with the
 * SPH state policies (src/SPH/SPHHydroBase.cc:238-336): ghosts and
 * neighbours at the step start, dt selection, derivatives at t, update to
 * t + dt/2, ghosts, derivatives at t + dt/2, update from the saved state,
 * ghosts, sum density (SPHHydroBase::finalize), boundary enforcement.
 * hydro->compatibleEnergyEvolution selects SpecificThermalEnergyPolicy for
 * eps (src/Hydro/SpecificThermalEnergyPolicy.cc:64-189).
 * Advances the sim's time; returns the step taken in *dt. */
int sphgpu_rk2_step(sphgpu_sim* sim, const sphgpu_hydro_params* hydro, const sphgpu_integrator_params* integ,
                    double max_time, double* dt);
/* iterateIdealH (src/Utilities/iterateIdealH.cc:30-310) with the SPH smoothing
 * scale (SPHSmoothingScale::idealSmoothingScale, src/NodeList/SPHSmoothingScale.cc:120-197;
 * nPerh, hmin, hmax from `hydro`), nPerhForIteration = 0, no spherical start, no
 * fixed determinant: each iteration makes the ghosts of the sim's boundaries and
 * the connectivity from the current H, then replaces H of every internal node
 * whose last change exceeds `tolerance`; ends after max_iterations or when the
 * largest change (max |phi - 1| over the eigenvalues of H1^1/2 H^-1 H1^1/2) is
 * within the tolerance (reference defaults: 100, 1e-10).  The ghosts are made
 * from the final H.  *iterations and *max_delta report the loop (may be null). */
int sphgpu_iterate_ideal_h(sphgpu_sim* sim, const sphgpu_hydro_params* hydro, int max_iterations, double tolerance,
                           double kernel_extent, int cull, int* iterations, double* max_delta);
/* Conservation totals of the internal nodes, as SpheralConservation takes them
 * (src/SimulationControl/SpheralConservation.py:66-150): out[9] = total mass,
 * linear momentum (x, y, z), angular momentum about the origin (x, y, z; z only
 * in 2-D, zero in 1-D), kinetic energy sum m v^2 / 2, thermal energy sum m eps.
 * Summed on the device in a fixed order (deterministic). */
int sphgpu_conservation(sphgpu_sim* sim, double* out);
/* The same totals over one ensemble member's rows. */
int sphgpu_ensemble_conservation(sphgpu_sim* sim, int64_t id, double* out);
/* No counterpart in the reference: for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops).  The totals of
 * sphgpu_conservation over internal rows [first, first + rows) (all of them,
 * or one ensemble member's), bitwise the same, written to out_device[9] in
 * device memory asynchronously on the sim's stream (sphgpu_synchronize before
 * reading).  Allocates nothing: the partial sums live in a static device array
 * of 256 x 9 doubles per GPU, enough for any row count.  Every call on a GPU
 * shares that array, so calls that overlap on one GPU (sims on other streams
 * of the same device) race; a row range outside the internal rows is rejected
 * with SPHGPU_ERR_ARG. */
int sphgpu_conservation_noalloc(sphgpu_sim* sim, int64_t first, int64_t rows, double* out_device);
/* The previous step's derivatives the dt vote reads (per internal row:
 * maxViscousPressure, tr DvDx, DvDt; 2 + dim doubles), rows [first, first +
 * rows): zero before the first step.  With the state and the clock they let a
 * restored simulation continue on the same trajectory (the reference dumps its
 * derivative fields with the restart state, src/SPH/SPHHydroBase.cc:1283-1318). */
int sphgpu_get_dt_inputs(sphgpu_sim* sim, int64_t first, int64_t rows, double* out);
int sphgpu_set_dt_inputs(sphgpu_sim* sim, int64_t first, int64_t rows, const double* in);
/* Integrator clock: current time, last step, cycle count (and setter). */
int sphgpu_get_time(sphgpu_sim* sim, double* time, double* last_dt, int64_t* cycle);
int sphgpu_set_time(sphgpu_sim* sim, double time, double last_dt, int64_t cycle);

/* Run the device part of each time step (sphgpu_rk2_step, sphgpu_ensemble_step:
 * dt vote and selection, both derivative passes, updates, ghost refreshes,
 * the finalize density sum, boundary enforcement, clock advance) as a HIP
 * graph: the first step with a new launch key runs eagerly, the second is
 * captured, later ones replay it (the key: node and slot counts, device
 * buffers, parameters, record state).  Results are bitwise those of the eager
 * step.  Stage timings are not updated by graph steps.  No counterpart in the
 * reference; a launch-overhead measure for small node sets (default off). */
int sphgpu_set_graphs(sphgpu_sim* sim, int on);
/* Steps captured as graphs and steps replayed from one, since creation. */
int sphgpu_graph_counts(sphgpu_sim* sim, int64_t* captures, int64_t* replays);

/* Launch-shape autotuner (process-wide; no counterpart in the reference).
 * With tuning on, the first pass of a workload class (kernel, dimension,
 * arithmetic, log2 buckets of row length and node count) times every launch
 * variant on the live inputs and keeps the fastest; variants compute the same
 * values in the same order, so results do not change.  Tuned now: the 1-D
 * and 2-D isotropic pair kernel (workgroups of 64-512 threads, table in LDS
 * or read through L1).  Decisions export as text lines and import again, so
 * that measured choices are reused without timing. */
int sphgpu_autotune(int on);
/* Copy the decision table into buf (NUL-terminated, at most cap bytes);
 * *len = its full length. */
int sphgpu_autotune_export(char* buf, int64_t cap, int64_t* len);
/* Load decisions (the export format); *entries = lines taken. */
int sphgpu_autotune_import(const char* text, int* entries);

/* Device times in milliseconds of the last call of each stage, measured with
 * HIP events on the sim's stream: [neighbours, sum density, derivatives, pack,
 * ghost generation, ghost culling]. */
enum { SPHGPU_NUM_TIMINGS = 6 };
int sphgpu_last_timings(sphgpu_sim* sim, double* ms);
/* Block until all work queued on the sim's stream has finished. */
int sphgpu_synchronize(sphgpu_sim* sim);

/* ---- Ensembles: many independent simulations in one node set --------------
 * No counterpart in the reference, which runs one simulation per process
 * (src/SimulationControl/SpheralController.py:303-336 drives one
 * SynchronousRK2).  Each member is advanced by exactly that step with its own
 * dt, time and gamma; the members share the kernel table, the hydro and
 * integrator parameters, the row order and the boundaries (planes in each
 * member's own coordinates).  A member's rows are contiguous, in member order,
 * and its results do not depend on which other members share the set.
 *
 * Usage: sphgpu_create, table, rows order, sphgpu_set_boundaries, then
 * sphgpu_ensemble_begin; add members; sphgpu_ensemble_step until a member
 * reports its end time; download and remove it; add more at any time. */
typedef struct sphgpu_member {
  int64_t n;                   /* internal nodes (> 0)                                   */
  const double* pos;           /* n*dim  (caller layout of sphgpu_upload)                */
  const double* vel;           /* n*dim                                                  */
  const double* mass;          /* n                                                      */
  const double* rho;           /* n                                                      */
  const double* eps;           /* n     specific thermal energy                          */
  const double* H;             /* n*dim(dim+1)/2                                         */
  double gamma;                /* GammaLawGas gamma (P and c are set from rho, eps)      */
  double tEnd;                 /* the member runs to this time                           */
} sphgpu_member;

/* Switch a sim with no nodes to ensemble mode. */
int sphgpu_ensemble_begin(sphgpu_sim* sim);
/* Append a member (t = 0, last dt 1e-5, cycle 0).  integ supplies the pressure
 * limits of its first EOS evaluation.  *id identifies it from now on. */
int sphgpu_ensemble_add(sphgpu_sim* sim, const sphgpu_member* member, const sphgpu_integrator_params* integ,
                        int64_t* id);
/* One SynchronousRK2 step of every member (each with its own dt, clamped to
 * its end time).  *finished = number of members that reached their end time;
 * they must be downloaded and removed before the next step.
 * A member that fails is evicted by the step itself and the others are
 * stepped as if it had never been there (their results are unchanged): a
 * NaN or infinity in its position or H at the step start, or in its state
 * after the step's final update, or a dt vote below the dt threshold.
 * sphgpu_ensemble_failed lists the members the last step evicted.  The
 * reference has no per-simulation failure: Process::haltAll ends the run
 * (src/Utilities/Process.cc:74-87) and a step below
 * Integrator::dtThreshold is only reported (src/Integrator/Integrator.cc:190-197). */
int sphgpu_ensemble_step(sphgpu_sim* sim, const sphgpu_hydro_params* hydro, const sphgpu_integrator_params* integ,
                         int* finished);
/* Reasons a member fails (bits). */
enum {
  SPHGPU_MEMBER_OK = 0,
  SPHGPU_MEMBER_NONFINITE = 1,    /* NaN or infinity in its state                   */
  SPHGPU_MEMBER_BAD_H = 2,        /* an H without a positive lower eigenvalue bound */
  SPHGPU_MEMBER_DT_COLLAPSE = 4   /* its dt vote fell below the dt threshold        */
};
/* Members evicted by the last sphgpu_ensemble_step: id, reason
 * (SPHGPU_MEMBER_*), steps taken (arrays of length cap, any may be null);
 * *count = number evicted. */
int sphgpu_ensemble_failed(sphgpu_sim* sim, int cap, int64_t* ids, int32_t* reasons, int64_t* cycles, int* count);
/* A member whose dt vote (cfl times GenericHydro::dt's minimum) falls below
 * dtThreshold fails with SPHGPU_MEMBER_DT_COLLAPSE; 0 (the default) never.
 * Like Integrator::dtThreshold (default 1e-10, src/Integrator/Integrator.cc:56),
 * which only reports such a step. */
int sphgpu_ensemble_set_dt_threshold(sphgpu_sim* sim, double dtThreshold);
/* Members in row order: id, time, last dt, cycle, first row, rows (arrays of
 * length cap, any may be null); *count = number of members. */
int sphgpu_ensemble_members(sphgpu_sim* sim, int cap, int64_t* ids, double* time, double* last_dt, int64_t* cycle,
                            int64_t* first_row, int64_t* rows, int* count);
/* A member's internal rows of one state field (SPHGPU_F_*), read or written. */
int sphgpu_ensemble_download(sphgpu_sim* sim, int64_t id, int field, double* host);
int sphgpu_ensemble_upload(sphgpu_sim* sim, int64_t id, int field, const double* host);
/* Set a member's clock (time, last dt, cycle), e.g. when it is restored from a
 * snapshot (the reference restores the integrator's clock with its restart
 * state, src/Integrator/Integrator.cc:794-823). */
int sphgpu_ensemble_set_clock(sphgpu_sim* sim, int64_t id, double time, double last_dt, int64_t cycle);
/* Remove members; the others keep their state and results. */
int sphgpu_ensemble_remove(sphgpu_sim* sim, int n_ids, const int64_t* ids);
/* Device bytes per node the library holds for a member of this dimension
 * (state, neighbour rows, records and derivative fields; ghosts not included):
 * the unit of admission by HBM footprint. */
int64_t sphgpu_ensemble_bytes_per_node(int dim);

/* ---- Serving: continuous batching of arriving simulations -----------------
 * No counterpart in the reference.  Requests become visible at `arrival`
 * seconds after the call starts; they are admitted first come first served
 * while the resident footprint (sphgpu_ensemble_bytes_per_node * n *
 * ghostFactor per member) fits hbmBudget and fewer than maxMembers are
 * resident; all resident members advance together (sphgpu_ensemble_step) and
 * each leaves when it reaches tEnd, with its final state copied to the out
 * arrays, or when it fails (status set, the others unaffected).  Returns when
 * every request has finished or failed.  The sim must be an empty ensemble
 * (sphgpu_ensemble_begin). */
typedef struct sphgpu_request {
  sphgpu_member member;        /* the simulation                                          */
  double arrival;              /* seconds after the start of the call                     */
  double* outPos;              /* n*dim, or null  (final state)                           */
  double* outVel;              /* n*dim, or null                                          */
  double* outRho;              /* n, or null                                              */
  double* outEps;              /* n, or null                                              */
  double* outH;                /* n*dim(dim+1)/2, or null                                 */
  double admitted;             /* out: seconds after the start when it joined             */
  double finished;             /* out: seconds after the start when it reached tEnd or
                                       failed                                             */
  int64_t steps;               /* out: RK2 steps taken                                    */
  int32_t status;              /* out: SPHGPU_MEMBER_OK, or the reason it failed (its out
                                       arrays are then left untouched)                    */
} sphgpu_request;
typedef struct sphgpu_serve_params {
  int64_t hbmBudget;           /* bytes; <= 0: 80% of the device memory free at the start */
  int maxMembers;              /* resident members at most                                */
  double ghostFactor;          /* footprint rows per internal node (>= 1)                 */
  double dtThreshold;          /* sphgpu_ensemble_set_dt_threshold for the call (0: off)  */
} sphgpu_serve_params;
int sphgpu_serve(sphgpu_sim* sim, sphgpu_request* requests, int n_requests, const sphgpu_hydro_params* hydro,
                 const sphgpu_integrator_params* integ, const sphgpu_serve_params* serve);

/* Measurement helper with no counterpart in the reference: copy `bytes` bytes
 * between two device buffers with a plain streaming kernel `repeats` times and
 * return the average kernel time (ms).  Used to calibrate HBM traffic counters
 * and to state the achievable copy bandwidth beside the roofline figures. */
int sphgpu_copy_probe(sphgpu_sim* sim, int64_t bytes, int repeats, double* ms);
/* Measurement helper with no counterpart in the reference: walk the rows of
 * the last neighbour build as the pair and density kernels do (one lane per
 * internal node, their work order) and only load loads_per_neighbour (2, 3 or
 * 6) 16-B pieces of a record per neighbour; return the average kernel time
 * (ms) over `repeats` launches: the gather ceiling of those kernels on these
 * rows, stated beside their times. */
int sphgpu_gather_probe(sphgpu_sim* sim, int loads_per_neighbour, int repeats, double* ms);

#ifdef __cplusplus
}
#endif
#endif /* SPHGPU_H */
