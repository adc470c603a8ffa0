// Drives the C++ host layer (SpheralGPU.hh) the way a Spheral script drives
// SPHHydroBase and SynchronousRK2, for the GPU tests of that layer.
//
//   spheral_host_check derivs <in> <out>   initializeProblemStartup, registerState,
//                                          registerDerivatives, initialize,
//                                          evaluateDerivatives, dt, finalize
//   spheral_host_check rk2 <in> <out>      SynchronousRK2::step, `steps` times
//
// <in> and <out> are sequences of named double arrays:
//   int32 name length, name bytes, int64 count, count doubles.
// Inputs: dim, n_internal, n_total, gamma, pos, vel, mass, rho, eps, H, optional
// P, cs (derivs: replace the equation of state's values in the State),
// optional boundaries (13 doubles each: kind, plane1 point(3), normal(3),
// plane2 point(3), normal(3)), options epsTensile, gradh, compatible, steps.
// Several NodeLists: list_counts (internal nodes of each, in row order), gammas
// (one each), optional nperh (one each, default 1.51) and sum_all
// (sumMassDensityOverAllNodeLists, default 1).
#include <cstdio>
#include <cstring>
#include <deque>
#include <fstream>
#include <iostream>
#include <map>
#include <string>
#include <vector>

#include "SpheralGPU.hh"

using namespace Spheral;
using Records = std::map<std::string, std::vector<double>>;

static Records readRecords(const std::string& path) {
  std::ifstream f(path, std::ios::binary);
  VERIFY2(bool(f), "cannot open " + path);
  Records r;
  while (true) {
    int32_t len = 0;
    if (!f.read(reinterpret_cast<char*>(&len), sizeof(len))) break;
    VERIFY2(len > 0 && len < 256, "bad record name length");
    std::string name(size_t(len), '\0');
    f.read(&name[0], len);
    int64_t count = 0;
    f.read(reinterpret_cast<char*>(&count), sizeof(count));
    VERIFY2(bool(f) && count >= 0 && count < (int64_t(1) << 32), "bad record count");
    std::vector<double> v(static_cast<size_t>(count));
    f.read(reinterpret_cast<char*>(v.data()), std::streamsize(sizeof(double) * size_t(count)));
    VERIFY2(bool(f), "truncated record " + name);
    r[name] = std::move(v);
  }
  return r;
}

static void writeRecords(const std::string& path, const Records& r) {
  std::ofstream f(path, std::ios::binary);
  VERIFY2(bool(f), "cannot write " + path);
  for (const auto& kv : r) {
    const int32_t len = int32_t(kv.first.size());
    const int64_t count = int64_t(kv.second.size());
    f.write(reinterpret_cast<const char*>(&len), sizeof(len));
    f.write(kv.first.data(), len);
    f.write(reinterpret_cast<const char*>(&count), sizeof(count));
    f.write(reinterpret_cast<const char*>(kv.second.data()), std::streamsize(sizeof(double) * kv.second.size()));
  }
  VERIFY2(bool(f), "write failed: " + path);
}

static double scalar(const Records& r, const std::string& k, double dflt) {
  auto it = r.find(k);
  return it == r.end() || it->second.empty() ? dflt : it->second[0];
}
static const std::vector<double>& array(const Records& r, const std::string& k) {
  auto it = r.find(k);
  VERIFY2(it != r.end(), "missing input " + k);
  return it->second;
}

template <typename Dimension>
static typename Dimension::Vector vec(const double* p) {
  typename Dimension::Vector v{};
  for (int k = 0; k < Dimension::nDim; ++k) v[size_t(k)] = p[k];
  return v;
}

template <typename Dimension>
struct Problem {
  std::deque<NodeList<Dimension>> lists;   // stable addresses for the DataBase
  DataBase<Dimension> db;
  TableKernel<Dimension> W;
  MonaghanGingoldViscosity<Dimension> Q;
  SPHHydroBase<Dimension> hydro;

  // The SPH() defaults for the B-spline kernel (src/SPH/SPHHydros.py:317-384):
  // cfl 0.25, XSPH, correctVelocityGradient, RigorousSumDensity, IdealH.
  Problem(const Records& in, int64_t nGhost)
      : W(BSplineKernel<Dimension>(), 1000),
        Q(1.0, 1.0),
        hydro(SPHSmoothingScale<Dimension>(), Q, W, W, 0.0, 0.25, false, scalar(in, "compatible", 0) != 0.0, false,
              scalar(in, "gradh", 0) != 0.0, true, true, scalar(in, "sum_all", 1) != 0.0,
              MassDensityType::RigorousSumDensity, HEvolutionType::IdealH, scalar(in, "epsTensile", 0.0), 4.0,
              typename Dimension::Vector{}, typename Dimension::Vector{}) {
    const std::vector<double> counts =
        in.count("list_counts") ? array(in, "list_counts") : std::vector<double>{scalar(in, "n_internal", 0)};
    const std::vector<double> gammas =
        in.count("gammas") ? array(in, "gammas") : std::vector<double>{scalar(in, "gamma", 5.0 / 3.0)};
    const std::vector<double> nPerh =
        in.count("nperh") ? array(in, "nperh") : std::vector<double>(counts.size(), 1.51);
    VERIFY2(counts.size() == gammas.size() && counts.size() == nPerh.size(), "one gamma and nPerh per NodeList");
    const int D = Dimension::nDim, S = Dimension::nSym;
    size_t at = 0;   // first row of the NodeList in the input arrays
    for (size_t k = 0; k < counts.size(); ++k) {
      // the caller's ghost rows (no boundaries) belong to the last NodeList
      lists.emplace_back("nodes" + std::to_string(k), GammaLawGas<Dimension>(gammas[k], 1.0), int64_t(counts[k]),
                         k + 1 == counts.size() ? nGhost : 0, 1.0e-20, 1.0e20, 0.1, nPerh[k]);
      NodeList<Dimension>& nl = lists.back();
      const size_t n = size_t(nl.numNodes());
      auto take = [&](const char* key, std::vector<double>& dst, int comps) {
        const auto& src = array(in, key);
        VERIFY2(src.size() >= (at + n) * size_t(comps), std::string("input too short: ") + key);
        std::copy(src.begin() + std::ptrdiff_t(at * size_t(comps)),
                  src.begin() + std::ptrdiff_t((at + n) * size_t(comps)), dst.begin());
      };
      take("pos", nl.positions(), D);
      take("vel", nl.velocity(), D);
      take("mass", nl.mass(), 1);
      take("rho", nl.massDensity(), 1);
      take("eps", nl.specificThermalEnergy(), 1);
      take("H", nl.Hfield(), S);
      db.appendNodeList(nl);
      at += n;
    }
    auto b = in.find("boundaries");
    if (b != in.end()) {
      VERIFY2(b->second.size() % 13 == 0, "boundaries: 13 doubles each");
      for (size_t q = 0; q < b->second.size(); q += 13) {
        const double* r = &b->second[q];
        GeomPlane<Dimension> p1(vec<Dimension>(r + 1), vec<Dimension>(r + 4));
        GeomPlane<Dimension> p2(vec<Dimension>(r + 7), vec<Dimension>(r + 10));
        if (r[0] == 0.0) hydro.appendBoundary(ReflectingBoundary<Dimension>(p1));
        else hydro.appendBoundary(PeriodicBoundary<Dimension>(p1, p2));
      }
    }
  }
};

template <typename Dimension>
static Records runDerivs(const Records& in) {
  const int64_t nInt = int64_t(scalar(in, "n_internal", 0));
  const int64_t nTot = int64_t(scalar(in, "n_total", double(nInt)));
  const bool boundaries = in.count("boundaries") != 0;
  Problem<Dimension> pb(in, boundaries ? 0 : nTot - nInt);
  Records out;
  pb.hydro.initializeProblemStartup(pb.db);
  out["eos P"] = pb.hydro.pressure();
  out["eos cs"] = pb.hydro.soundSpeed();
  State<Dimension> state;
  StateDerivatives<Dimension> derivs;
  pb.hydro.registerState(pb.db, state);
  pb.hydro.registerDerivatives(pb.db, derivs);
  const std::vector<std::string> registered = derivs.keys();
  // the test's own pressure and sound speed replace the equation of state's
  if (in.count("P")) state.enroll(HydroFieldNames::pressure, array(in, "P"));
  if (in.count("cs")) state.enroll(HydroFieldNames::soundSpeed, array(in, "cs"));
  pb.hydro.initialize(0.0, 0.0, pb.db, state, derivs);
  pb.hydro.evaluateDerivatives(0.0, 0.0, pb.db, state, derivs);
  VERIFY2(derivs.keys() == registered, "evaluateDerivatives wrote a field registerDerivatives did not make");
  for (const auto& k : derivs.keys()) out["derivs/" + k] = derivs.field(k);
  out["pair count"] = {double(pb.hydro.pairCount())};
  out["omega"] = std::vector<double>(state.field(HydroFieldNames::omegaGradh).begin(),
                                     state.field(HydroFieldNames::omegaGradh).begin() + nInt);
  const auto vote = pb.hydro.dt(pb.db, state, derivs, 0.0);
  out["dt"] = {vote.first};
  pb.hydro.finalize(0.0, 0.0, pb.db, state, derivs);
  out["finalized rho"] = std::vector<double>(state.field(HydroFieldNames::massDensity).begin(),
                                             state.field(HydroFieldNames::massDensity).begin() + nInt);
  return out;
}

template <typename Dimension>
static Records runRK2(const Records& in) {
  Problem<Dimension> pb(in, 0);
  pb.hydro.initializeProblemStartup(pb.db);
  SynchronousRK2<Dimension> integrator(pb.db, {&pb.hydro});
  const int steps = int(scalar(in, "steps", 1));
  Records out;
  std::vector<double> dts;
  for (int k = 0; k < steps; ++k) {
    integrator.step(3.40282347e+38);
    dts.push_back(integrator.lastDt());
  }
  VERIFY2(integrator.currentCycle() == steps, "cycle count");
  out["dt"] = dts;
  out["time"] = {integrator.currentTime()};
  using NL = NodeList<Dimension>;
  out["pos"] = pb.db.gather([](NL& l) -> auto& { return l.positions(); }, Dimension::nDim);
  out["vel"] = pb.db.gather([](NL& l) -> auto& { return l.velocity(); }, Dimension::nDim);
  out["rho"] = pb.db.gather([](NL& l) -> auto& { return l.massDensity(); }, 1);
  out["eps"] = pb.db.gather([](NL& l) -> auto& { return l.specificThermalEnergy(); }, 1);
  out["H"] = pb.db.gather([](NL& l) -> auto& { return l.Hfield(); }, Dimension::nSym);
  out["P"] = pb.hydro.pressure();
  out["cs"] = pb.hydro.soundSpeed();
  return out;
}

template <typename Dimension>
static Records run(const std::string& mode, const Records& in) {
  if (mode == "derivs") return runDerivs<Dimension>(in);
  if (mode == "rk2") return runRK2<Dimension>(in);
  throw VERIFYError("unknown mode " + mode);
}

int main(int argc, char** argv) {
  if (argc != 4) {
    std::fprintf(stderr, "usage: %s derivs|rk2 <in> <out>\n", argv[0]);
    return 2;
  }
  try {
    const Records in = readRecords(argv[2]);
    const int dim = int(scalar(in, "dim", 0));
    std::printf("[host check] %s dim=%d n_internal=%g\n", argv[1], dim, scalar(in, "n_internal", 0));
    std::fflush(stdout);
    Records out;
    if (dim == 1) out = run<Dim1>(argv[1], in);
    else if (dim == 2) out = run<Dim2>(argv[1], in);
    else if (dim == 3) out = run<Dim3>(argv[1], in);
    else throw VERIFYError("dim must be 1, 2 or 3");
    writeRecords(argv[3], out);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "error: %s\n", e.what());
    return 1;
  }
  return 0;
}
