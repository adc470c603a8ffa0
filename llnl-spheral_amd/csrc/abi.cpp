// extern "C" entry points of libsphgpu.so (declared in include/sphgpu.h).
// Each entry point converts exceptions to status codes and records the message
// on the sim; nothing C++ crosses the ABI.
#include <cstring>
#include <algorithm>
#include "internal.hh"
#include "ell.hh"

using namespace sphgpu;

struct sphgpu_sim : Sim {};

namespace {

int fail(sphgpu_sim* s, int code, const std::string& msg) {
  if (s) s->lastError = msg;
  return code;
}

template <class F> int guarded(sphgpu_sim* s, F&& f) {
  if (!s) return SPHGPU_ERR_ARG;
  try {
    SG_CHECK_HIP(hipSetDevice(s->device));
    f();
    return SPHGPU_OK;
  } catch (const Error& e) {
    return fail(s, e.code, e.what());
  } catch (const MemberFailure& e) {
    // only sphgpu_ensemble_step evicts failing members; other calls report them
    return fail(s, SPHGPU_ERR_ARG, "ensemble member " + std::to_string(e.members.front()) +
                                       " (row order) has a non-finite position or H, or H without a positive bound");
  } catch (const std::exception& e) {
    return fail(s, SPHGPU_ERR_HIP, e.what());
  }
}

int stateComponents(int dim, int field) {
  switch (field) {
    case SPHGPU_F_POSITION: case SPHGPU_F_VELOCITY: return dim;
    case SPHGPU_F_H: return nsymOf(dim);
    default: return 1;
  }
}

DBuf& stateBuf(Sim& s, int field) {
  switch (field) {
    case SPHGPU_F_POSITION: return s.pos;
    case SPHGPU_F_VELOCITY: return s.vel;
    case SPHGPU_F_MASS: return s.mass;
    case SPHGPU_F_MASS_DENSITY: return s.rho;
    case SPHGPU_F_SPECIFIC_THERMAL_ENERGY: return s.eps;
    case SPHGPU_F_H: return s.H;
    case SPHGPU_F_PRESSURE: return s.P;
    case SPHGPU_F_SOUND_SPEED: return s.cs;
    case SPHGPU_F_OMEGA_GRADH: return s.omega;
    case SPHGPU_F_CL_MULTIPLIER: return s.fCl;
    case SPHGPU_F_CQ_MULTIPLIER: return s.fCq;
    case SPHGPU_F_SHEAR_CORRECTION: return s.fshear;
  }
  throw Error(SPHGPU_ERR_ARG, "unknown state field id");
}

void installTable(Sim& s) {
  const int n = s.table.n;
  ++s.tableVersion;
  std::vector<double> inter(size_t(6) * n);
  for (int i = 0; i < n; ++i) {
    inter[6 * i + 0] = s.table.Ak[i]; inter[6 * i + 1] = s.table.Bk[i]; inter[6 * i + 2] = s.table.Ck[i];
    inter[6 * i + 3] = s.table.Ag[i]; inter[6 * i + 4] = s.table.Bg[i]; inter[6 * i + 5] = s.table.Cg[i];
  }
  s.dTable.ensure(sizeof(double) * inter.size());
  s.dNperh.ensure(sizeof(double) * n);
  s.dWsum.ensure(sizeof(double) * n);
  SG_CHECK_HIP(hipMemcpyAsync(s.dTable.p, inter.data(), sizeof(double) * inter.size(), hipMemcpyHostToDevice, s.stream));
  std::vector<float> t32(size_t(8) * n, 0.0f);
  for (int i = 0; i < n; ++i) {
    t32[8 * i + 0] = float(s.table.Ak[i]); t32[8 * i + 1] = float(s.table.Bk[i]); t32[8 * i + 2] = float(s.table.Ck[i]);
    t32[8 * i + 4] = float(s.table.Ag[i]); t32[8 * i + 5] = float(s.table.Bg[i]); t32[8 * i + 6] = float(s.table.Cg[i]);
  }
  s.dTable32.ensure(sizeof(float) * t32.size());
  SG_CHECK_HIP(hipMemcpyAsync(s.dTable32.p, t32.data(), sizeof(float) * t32.size(), hipMemcpyHostToDevice, s.stream));
  SG_CHECK_HIP(hipMemcpyAsync(s.dNperh.p, s.table.nperh.data(), sizeof(double) * n, hipMemcpyHostToDevice, s.stream));
  SG_CHECK_HIP(hipMemcpyAsync(s.dWsum.p, s.table.wsum.data(), sizeof(double) * n, hipMemcpyHostToDevice, s.stream));
  SG_CHECK_HIP(hipStreamSynchronize(s.stream));   // the host vectors die on return
}

}  // namespace

namespace sphgpu {
// Stage times of the last call of each stage (events recorded on the stream).
void resolveTimings(Sim& s) {
  auto el = [&](int a, int b) {
    float ms = 0.f;
    SG_CHECK_HIP(hipEventSynchronize(s.ev[b]));
    SG_CHECK_HIP(hipEventElapsedTime(&ms, s.ev[a], s.ev[b]));
    return ms;
  };
  if (s.timingPending & 1u) s.timing.neighbors_ms = el(0, 1);
  if (s.timingPending & 2u) s.timing.density_ms = el(2, 3);
  if (s.timingPending & 4u) { s.timing.derivs_ms = el(5, 6); s.timing.pack_ms = el(4, 5); }
  if (s.timingPending & 8u) { s.timing.ghosts_ms = el(8, 9); s.timing.cull_ms = el(10, 11); }
  s.timingPending = 0;
}
}  // namespace sphgpu

extern "C" {

int sphgpu_create(int dim, int device, sphgpu_sim** out) {
  if (!out || dim < 1 || dim > 3) return SPHGPU_ERR_ARG;
  *out = nullptr;
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) return SPHGPU_ERR_HIP;
  if (device < 0 || device >= count) return SPHGPU_ERR_ARG;
  sphgpu_sim* s = new sphgpu_sim();
  s->dim = dim;
  s->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess) {
    delete s;
    return SPHGPU_ERR_HIP;
  }
  for (auto& e : s->ev) {
    if (hipEventCreate(&e) != hipSuccess) { sphgpu_destroy(s); return SPHGPU_ERR_HIP; }
  }
  // the allocation-free entry point's scratch array on this device (once)
  try {
    resolveScratch(device);
  } catch (const Error& e) {
    sphgpu_destroy(s);
    return e.code;
  }
  void* pinned = nullptr;
  if (hipHostMalloc(&pinned, sizeof(unsigned long long), hipHostMallocDefault) != hipSuccess) {
    sphgpu_destroy(s);
    return SPHGPU_ERR_HIP;
  }
  s->ncalcHost = static_cast<unsigned long long*>(pinned);
  *s->ncalcHost = 0;
  *out = s;
  return SPHGPU_OK;
}

void sphgpu_destroy(sphgpu_sim* s) {
  if (!s) return;
  (void)hipSetDevice(s->device);
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  DBuf* bufs[] = {&s->dTable, &s->dNperh, &s->dWsum, &s->pos, &s->vel, &s->mass, &s->rho, &s->eps, &s->P, &s->cs,
                  &s->H, &s->omega, &s->fCl, &s->fCq, &s->fshear, &s->cellOf, &s->cellCount, &s->cellStart,
                  &s->cellFill, &s->perm, &s->iperm, &s->xs, &s->Hs, &s->rowCount, &s->nbr,
                  &s->scanTmp, &s->reduceTmp, &s->flags, &s->rec, &s->out};
  for (DBuf* b : bufs) b->release();
  for (auto& e : s->ev) if (e) (void)hipEventDestroy(e);
  for (auto& g : s->stepGraphs) if (g.exec) (void)hipGraphExecDestroy(g.exec);
  if (s->ncalcHost) (void)hipHostFree(s->ncalcHost);
  if (s->pinnedOut) (void)hipHostFree(s->pinnedOut);
  if (s->pinnedIn) (void)hipHostFree(s->pinnedIn);
  if (s->pinnedInDone) (void)hipEventDestroy(s->pinnedInDone);
  s->pick.release();
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

const char* sphgpu_last_error(const sphgpu_sim* s) { return s ? s->lastError.c_str() : "null sim"; }

int sphgpu_set_table(sphgpu_sim* s, int npoints, double extent, const double* c) {
  return guarded(s, [&] {
    SG_REQUIRE(npoints >= 4 && extent > 0.0 && c, SPHGPU_ERR_ARG, "bad table arguments");
    TableHost& t = s->table;
    t.n = npoints;
    t.extent = extent;
    t.step = extent / (npoints - 1);
    auto take = [&](int a, std::vector<double>& v) { v.assign(c + size_t(a) * npoints, c + size_t(a + 1) * npoints); };
    take(0, t.Ak); take(1, t.Bk); take(2, t.Ck); take(3, t.Ag); take(4, t.Bg); take(5, t.Cg);
    take(6, t.nperh); take(7, t.wsum);
    installTable(*s);
  });
}

int sphgpu_set_table_bspline(sphgpu_sim* s, int npoints) {
  return guarded(s, [&] {
    SG_REQUIRE(npoints >= 4, SPHGPU_ERR_ARG, "table needs at least 4 points");
    buildBSplineTable(s->dim, npoints, s->table);
    installTable(*s);
  });
}

int sphgpu_get_table(sphgpu_sim* s, int npoints, double* c) {
  return guarded(s, [&] {
    SG_REQUIRE(s->table.n > 0, SPHGPU_ERR_STATE, "no table installed");
    SG_REQUIRE(npoints == s->table.n && c, SPHGPU_ERR_ARG, "npoints does not match the installed table");
    const std::vector<double>* arr[8] = {&s->table.Ak, &s->table.Bk, &s->table.Ck, &s->table.Ag,
                                         &s->table.Bg, &s->table.Cg, &s->table.nperh, &s->table.wsum};
    for (int a = 0; a < 8; ++a) std::memcpy(c + size_t(a) * npoints, arr[a]->data(), sizeof(double) * npoints);
  });
}

int sphgpu_table_values(sphgpu_sim* s, double eta, double Hdet, double* W, double* gW) {
  return guarded(s, [&] {
    SG_REQUIRE(s->table.n > 0, SPHGPU_ERR_STATE, "no table installed");
    SG_REQUIRE(eta >= 0.0 && Hdet >= 0.0, SPHGPU_ERR_ARG, "eta and Hdet must be non-negative");
    if (W) *W = s->table.kernelValue(eta, Hdet);
    if (gW) *gW = s->table.gradValue(eta, Hdet);
  });
}

int sphgpu_set_nodes(sphgpu_sim* s, int64_t nInt, int64_t nTot) {
  return guarded(s, [&] {
    SG_REQUIRE(nInt >= 0 && nTot >= nInt && nTot < (int64_t(1) << 31) - 1, SPHGPU_ERR_ARG, "bad node counts");
    SG_REQUIRE(!s->ensemble, SPHGPU_ERR_STATE, "an ensemble takes its nodes with sphgpu_ensemble_add");
    // the previous step's dt inputs are rows of the old node set
    if (nInt != s->nInt) s->haveDtIn = false;
    s->nInt = nInt;
    s->nTot = nTot;
    s->deviceGhosts = false;
    s->neighborsValid = false;
    s->derivsValid = false;
    s->hasMultipliers = false;
    s->nLists = 0;             // one NodeList until sphgpu_set_node_lists
    s->listStart.clear();
    s->listGammaHost.clear();
    const size_t n = size_t(std::max<int64_t>(nTot, 1));
    const int D = s->dim, NS = nsymOf(D);
    s->pos.ensure(sizeof(double) * n * D);
    s->vel.ensure(sizeof(double) * n * D);
    s->H.ensure(sizeof(double) * n * NS);
    DBuf* scal[] = {&s->mass, &s->rho, &s->eps, &s->P, &s->cs, &s->omega};
    for (DBuf* b : scal) b->ensure(sizeof(double) * n);
    // omega defaults to 1 (no grad-h correction)
    std::vector<double> ones(n, 1.0);
    SG_CHECK_HIP(hipMemcpyAsync(s->omega.p, ones.data(), sizeof(double) * n, hipMemcpyHostToDevice, s->stream));
    SG_CHECK_HIP(hipStreamSynchronize(s->stream));   // `ones` dies on return
  });
}

int sphgpu_upload(sphgpu_sim* s, int field, const double* host, int64_t nRows) {
  return guarded(s, [&] {
    SG_REQUIRE(host && (nRows == s->nTot || (s->deviceGhosts && nRows == s->nInt)), SPHGPU_ERR_ARG,
               "upload needs n_total rows (or n_internal rows once the device made the ghosts)");
    SG_REQUIRE(!s->ensemble, SPHGPU_ERR_STATE, "an ensemble takes its nodes with sphgpu_ensemble_add");
    SG_REQUIRE(field >= 0 && field < SPHGPU_NUM_STATE_FIELDS, SPHGPU_ERR_ARG, "unknown state field id");
    const size_t bytes = sizeof(double) * size_t(nRows) * stateComponents(s->dim, field);
    DBuf& b = stateBuf(*s, field);
    if (field == SPHGPU_F_CL_MULTIPLIER || field == SPHGPU_F_CQ_MULTIPLIER || field == SPHGPU_F_SHEAR_CORRECTION) {
      // first multiplier upload materialises all three with 1.0
      if (!s->hasMultipliers) {
        std::vector<double> ones(std::max<size_t>(s->pos.bytes / (sizeof(double) * size_t(s->dim)), 1), 1.0);
        for (DBuf* m : {&s->fCl, &s->fCq, &s->fshear}) {
          m->ensure(sizeof(double) * ones.size());
          SG_CHECK_HIP(hipMemcpyAsync(m->p, ones.data(), sizeof(double) * ones.size(), hipMemcpyHostToDevice, s->stream));
        }
        s->hasMultipliers = true;
      }
    }
    b.ensure(std::max<size_t>(bytes, 8));
    if (bytes) SG_CHECK_HIP(hipMemcpyAsync(b.p, host, bytes, hipMemcpyHostToDevice, s->stream));
    SG_CHECK_HIP(hipStreamSynchronize(s->stream));
    // Neighbours stay those of the last build (the reference's finalize sums density
    // with the connectivity of the step start); the packed records are rebuilt lazily.
    // Records use the compact isotropic layout only while every H is h*I and no Q
    // multipliers are present.
    if (field == SPHGPU_F_H && s->isotropic) {
      const int D = s->dim, NS = nsymOf(D);
      for (int64_t i = 0; i < nRows && s->isotropic; ++i) {
        const double* h = host + i * NS;
        bool iso = true;
        int k = 0;
        for (int a = 0; a < D; ++a)
          for (int b = a; b < D; ++b, ++k) iso = iso && (a == b ? h[k] == h[0] : h[k] == 0.0);
        if (!iso) s->isotropic = false;
      }
    }
    if (field == SPHGPU_F_CL_MULTIPLIER || field == SPHGPU_F_CQ_MULTIPLIER || field == SPHGPU_F_SHEAR_CORRECTION)
      s->isotropic = false;
    s->recValid = false;
    s->derivsValid = false;
  });
}

int sphgpu_download(sphgpu_sim* s, int field, double* host, int64_t nRows) {
  return guarded(s, [&] {
    SG_REQUIRE(host && (nRows == s->nTot || nRows == s->nInt), SPHGPU_ERR_ARG,
               "download needs n_total or n_internal rows");
    SG_REQUIRE(field >= 0 && field < SPHGPU_NUM_STATE_FIELDS, SPHGPU_ERR_ARG, "unknown state field id");
    const size_t bytes = sizeof(double) * size_t(nRows) * stateComponents(s->dim, field);
    DBuf& b = stateBuf(*s, field);
    if ((field == SPHGPU_F_CL_MULTIPLIER || field == SPHGPU_F_CQ_MULTIPLIER || field == SPHGPU_F_SHEAR_CORRECTION) &&
        !s->hasMultipliers) {
      std::fill(host, host + nRows, 1.0);
      return;
    }
    if (bytes) SG_CHECK_HIP(hipMemcpyAsync(host, b.p, bytes, hipMemcpyDeviceToHost, s->stream));
    SG_CHECK_HIP(hipStreamSynchronize(s->stream));
  });
}

int sphgpu_build_neighbors(sphgpu_sim* s, double ext) {
  return guarded(s, [&] { buildNeighbors(*s, ext); });
}

int sphgpu_set_boundaries(sphgpu_sim* s, int n, const sphgpu_boundary* b) {
  return guarded(s, [&] {
    setBoundaries(*s, n, b);
    if (s->deviceGhosts) { s->nTot = s->nInt; s->deviceGhosts = false; }
    s->neighborsValid = false;
    s->recValid = false;
    s->derivsValid = false;
  });
}

int sphgpu_set_ghost_nodes(sphgpu_sim* s, double ext, int cull) {
  return guarded(s, [&] { setGhostNodes(*s, ext, cull != 0); });
}

int sphgpu_apply_ghost_boundaries(sphgpu_sim* s) {
  return guarded(s, [&] { applyGhostBoundaries(*s); });
}

int sphgpu_node_counts(sphgpu_sim* s, int64_t* nInt, int64_t* nTot) {
  return guarded(s, [&] {
    SG_REQUIRE(nInt && nTot, SPHGPU_ERR_ARG, "null output");
    *nInt = s->nInt;
    *nTot = s->nTot;
  });
}

int sphgpu_download_ghost_controls(sphgpu_sim* s, int32_t* control, int32_t* sub) {
  return guarded(s, [&] {
    SG_REQUIRE(control && sub, SPHGPU_ERR_ARG, "null output");
    const int64_t ng = s->deviceGhosts ? s->nTot - s->nInt : 0;
    if (ng == 0) return;
    SG_CHECK_HIP(hipMemcpyAsync(control, s->ghostControl.p, sizeof(int32_t) * size_t(ng), hipMemcpyDeviceToHost,
                                s->stream));
    SG_CHECK_HIP(hipMemcpyAsync(sub, s->ghostSub.p, sizeof(int32_t) * size_t(ng), hipMemcpyDeviceToHost, s->stream));
    SG_CHECK_HIP(hipStreamSynchronize(s->stream));
  });
}

int sphgpu_omega_gradh(sphgpu_sim* s) {
  return guarded(s, [&] { omegaGradh(*s); });
}

int sphgpu_eos_gamma_law(sphgpu_sim* s, double gamma, double Pmin, double Pmax, int floorType) {
  return guarded(s, [&] { eosGammaLaw(*s, gamma, Pmin, Pmax, floorType); });
}

int sphgpu_hydro_dt(sphgpu_sim* s, double nPerh, double cfl, int useVelMag, double* dt) {
  return guarded(s, [&] {
    SG_REQUIRE(dt, SPHGPU_ERR_ARG, "null output");
    *dt = hydroDt(*s, nPerh, cfl, useVelMag);
  });
}

int sphgpu_rk2_step(sphgpu_sim* s, const sphgpu_hydro_params* hp, const sphgpu_integrator_params* ip, double maxTime,
                    double* dt) {
  return guarded(s, [&] {
    SG_REQUIRE(hp && ip && dt, SPHGPU_ERR_ARG, "null argument");
    *dt = rk2Step(*s, *hp, *ip, maxTime);
  });
}

int sphgpu_get_time(sphgpu_sim* s, double* t, double* lastDt, int64_t* cycle) {
  return guarded(s, [&] {
    SG_REQUIRE(t && lastDt && cycle, SPHGPU_ERR_ARG, "null output");
    *t = s->time; *lastDt = s->lastDt; *cycle = s->cycle;
  });
}

int sphgpu_set_time(sphgpu_sim* s, double t, double lastDt, int64_t cycle) {
  return guarded(s, [&] {
    SG_REQUIRE(lastDt > 0.0 && cycle >= 0, SPHGPU_ERR_ARG, "bad clock");
    s->time = t; s->lastDt = lastDt; s->cycle = cycle;
  });
}

int sphgpu_set_row_order(sphgpu_sim* s, int order) {
  return guarded(s, [&] {
    SG_REQUIRE(order == SPHGPU_ROW_ORDER_REFERENCE || order == SPHGPU_ROW_ORDER_SEARCH, SPHGPU_ERR_ARG,
               "unknown row order");
    if (order != s->rowOrder) s->neighborsValid = false;
    s->rowOrder = order;
  });
}

int sphgpu_neighbor_count(sphgpu_sim* s, int64_t* n) {
  return guarded(s, [&] {
    SG_REQUIRE(n, SPHGPU_ERR_ARG, "null output");
    SG_REQUIRE(s->neighborsValid, SPHGPU_ERR_STATE, "no current neighbour build");
    *n = s->nEntries;
  });
}

int sphgpu_download_neighbors(sphgpu_sim* s, int64_t* offsets, int32_t* nbr) {
  return guarded(s, [&] {
    SG_REQUIRE(offsets, SPHGPU_ERR_ARG, "null offsets");
    SG_REQUIRE(s->neighborsValid, SPHGPU_ERR_STATE, "no current neighbour build");
    // Rows are held as a transposed ELL over sorted slots; the ConnectivityMap
    // contract is rows of caller indices in ascending order.
    const int64_t n = s->nSlots;
    const size_t cap = size_t(s->ellCap);
    std::vector<int32_t> cnt(static_cast<size_t>(n)), perm(static_cast<size_t>(n));
    std::vector<int32_t> iperm(static_cast<size_t>(s->nTot));
    std::vector<int32_t> ell(static_cast<size_t>(n) * cap);
    SG_CHECK_HIP(hipMemcpyAsync(cnt.data(), s->rowCount.p, sizeof(int32_t) * size_t(n), hipMemcpyDeviceToHost, s->stream));
    SG_CHECK_HIP(hipMemcpyAsync(perm.data(), s->perm.p, sizeof(int32_t) * size_t(n), hipMemcpyDeviceToHost, s->stream));
    SG_CHECK_HIP(hipMemcpyAsync(iperm.data(), s->iperm.p, sizeof(int32_t) * iperm.size(), hipMemcpyDeviceToHost, s->stream));
    if (!ell.empty()) SG_CHECK_HIP(hipMemcpyAsync(ell.data(), s->nbr.p, sizeof(int32_t) * ell.size(), hipMemcpyDeviceToHost, s->stream));
    SG_CHECK_HIP(hipStreamSynchronize(s->stream));
    int64_t w = 0;
    for (int64_t i = 0; i < s->nInt; ++i) {
      offsets[i] = w;
      const int64_t k = iperm[size_t(i)];
      const int len = cnt[size_t(k)];
      if (nbr) {
        for (int q = 0; q < len; ++q) nbr[w + q] = perm[size_t(ell[size_t(ellIndex(q, n, k))])];
        std::sort(nbr + w, nbr + w + len);
      }
      w += len;
    }
    offsets[s->nInt] = w;
  });
}

int sphgpu_set_precision(sphgpu_sim* s, int precision) {
  return guarded(s, [&] {
    SG_REQUIRE(precision == SPHGPU_PRECISION_FP64 || precision == SPHGPU_PRECISION_MIXED, SPHGPU_ERR_ARG,
               "unknown precision");
    s->precision = precision;
  });
}

int sphgpu_sum_density(sphgpu_sim* s) {
  return guarded(s, [&] { sumDensity(*s); });
}

int sphgpu_evaluate_derivatives(sphgpu_sim* s, const sphgpu_hydro_params* p) {
  return guarded(s, [&] {
    SG_REQUIRE(p, SPHGPU_ERR_ARG, "null params");
    evalDerivs(*s, *p);
  });
}

int sphgpu_download_derivative(sphgpu_sim* s, int field, double* host, int64_t nRows) {
  return guarded(s, [&] {
    SG_REQUIRE(host && nRows == s->nInt, SPHGPU_ERR_ARG, "derivative download needs n_internal rows");
    SG_REQUIRE(field >= 0 && field < SPHGPU_NUM_DERIV_FIELDS, SPHGPU_ERR_ARG, "unknown derivative field id");
    SG_REQUIRE(s->derivsValid, SPHGPU_ERR_STATE, "no current derivatives");
    const int comps = derivComponents(s->dim, field);
    const int64_t n = s->outStride;
    std::vector<double> block(static_cast<size_t>(n) * comps);
    std::vector<int32_t> iperm(static_cast<size_t>(s->nTot));
    const double* src = s->out.as<double>() + int64_t(derivOffsetDoubles(s->dim, field)) * n;
    SG_CHECK_HIP(hipMemcpyAsync(block.data(), src, sizeof(double) * block.size(), hipMemcpyDeviceToHost, s->stream));
    SG_CHECK_HIP(hipMemcpyAsync(iperm.data(), s->iperm.p, sizeof(int32_t) * iperm.size(), hipMemcpyDeviceToHost, s->stream));
    SG_CHECK_HIP(hipStreamSynchronize(s->stream));
    for (int64_t i = 0; i < nRows; ++i) {
      const int64_t k = iperm[size_t(i)];
      std::memcpy(host + i * comps, block.data() + k * comps, sizeof(double) * comps);
    }
  });
}

int sphgpu_pair_count(sphgpu_sim* s, int64_t* n) {
  return guarded(s, [&] {
    SG_REQUIRE(n, SPHGPU_ERR_ARG, "null output");
    SG_REQUIRE(s->pairsValid, SPHGPU_ERR_STATE, "no derivative pass on the current neighbour rows");
    SG_CHECK_HIP(hipStreamSynchronize(s->stream));
    *n = int64_t(*s->ncalcHost);
  });
}

int sphgpu_last_timings(sphgpu_sim* s, double* ms) {
  return guarded(s, [&] {
    SG_REQUIRE(ms, SPHGPU_ERR_ARG, "null output");
    resolveTimings(*s);
    ms[0] = s->timing.neighbors_ms; ms[1] = s->timing.density_ms; ms[2] = s->timing.derivs_ms; ms[3] = s->timing.pack_ms;
    ms[4] = s->timing.ghosts_ms; ms[5] = s->timing.cull_ms;
  });
}

int sphgpu_synchronize(sphgpu_sim* s) {
  return guarded(s, [&] { SG_CHECK_HIP(hipStreamSynchronize(s->stream)); });
}

int sphgpu_iterate_ideal_h(sphgpu_sim* s, const sphgpu_hydro_params* hp, int maxIterations, double tolerance,
                           double kernelExtent, int cull, int* iterations, double* maxDelta) {
  return guarded(s, [&] {
    SG_REQUIRE(hp, SPHGPU_ERR_ARG, "null parameters");
    const int it = iterateIdealH(*s, *hp, maxIterations, tolerance, kernelExtent, cull != 0, maxDelta);
    if (iterations) *iterations = it;
  });
}

int sphgpu_conservation(sphgpu_sim* s, double* out) {
  return guarded(s, [&] {
    SG_REQUIRE(out, SPHGPU_ERR_ARG, "null output");
    SG_REQUIRE(s->nInt > 0, SPHGPU_ERR_STATE, "no nodes");
    conservationTotals(*s, 0, s->nInt, out);
  });
}

int sphgpu_ensemble_conservation(sphgpu_sim* s, int64_t id, double* out) {
  return guarded(s, [&] {
    SG_REQUIRE(out, SPHGPU_ERR_ARG, "null output");
    SG_REQUIRE(s->ensemble, SPHGPU_ERR_STATE, "not an ensemble");
    auto it = std::find_if(s->members.begin(), s->members.end(), [&](const Member& mb) { return mb.id == id; });
    SG_REQUIRE(it != s->members.end(), SPHGPU_ERR_ARG, "no member with that id");
    conservationTotals(*s, it->first, it->first + it->rows, out);
  });
}

int sphgpu_set_node_lists(sphgpu_sim* s, int n, const int64_t* internal_counts, const double* gamma,
                          int sum_over_all_node_lists) {
  return guarded(s, [&] { setNodeLists(*s, n, internal_counts, gamma, sum_over_all_node_lists); });
}

int sphgpu_set_graphs(sphgpu_sim* s, int on) {
  return guarded(s, [&] { s->useGraphs = on != 0; });
}

int sphgpu_graph_counts(sphgpu_sim* s, int64_t* captures, int64_t* replays) {
  return guarded(s, [&] {
    if (captures) *captures = s->graphCaptures;
    if (replays) *replays = s->graphReplays;
  });
}

int sphgpu_conservation_noalloc(sphgpu_sim* s, int64_t first, int64_t rows, double* out_device) {
  return guarded(s, [&] {
    SG_REQUIRE(out_device, SPHGPU_ERR_ARG, "null output");
    SG_REQUIRE(first >= 0 && rows > 0 && first + rows <= s->nInt, SPHGPU_ERR_ARG,
               "row range outside the internal rows");
    conservationTotalsNoAlloc(*s, first, first + rows, out_device);
  });
}

int sphgpu_get_dt_inputs(sphgpu_sim* s, int64_t first, int64_t rows, double* out) {
  return guarded(s, [&] {
    SG_REQUIRE(out && first >= 0 && rows >= 0 && first + rows <= s->nInt, SPHGPU_ERR_ARG, "bad row range");
    const size_t per = size_t(2 + s->dim);
    if (!s->haveDtIn) {
      std::fill(out, out + per * size_t(rows), 0.0);
      return;
    }
    SG_CHECK_HIP(hipMemcpyAsync(out, s->dtIn.as<double>() + size_t(first) * per, sizeof(double) * per * size_t(rows),
                                hipMemcpyDeviceToHost, s->stream));
    SG_CHECK_HIP(hipStreamSynchronize(s->stream));
  });
}

int sphgpu_set_dt_inputs(sphgpu_sim* s, int64_t first, int64_t rows, const double* in) {
  return guarded(s, [&] {
    SG_REQUIRE(in && first >= 0 && rows >= 0 && first + rows <= s->nInt, SPHGPU_ERR_ARG, "bad row range");
    const size_t per = size_t(2 + s->dim);
    if (!s->haveDtIn) {
      s->dtIn.ensure(sizeof(double) * per * size_t(std::max<int64_t>(s->nInt, 1)));
      SG_CHECK_HIP(hipMemsetAsync(s->dtIn.p, 0, sizeof(double) * per * size_t(s->nInt), s->stream));
      s->haveDtIn = true;
    }
    SG_CHECK_HIP(hipMemcpyAsync(s->dtIn.as<double>() + size_t(first) * per, in, sizeof(double) * per * size_t(rows),
                                hipMemcpyHostToDevice, s->stream));
    SG_CHECK_HIP(hipStreamSynchronize(s->stream));
  });
}

int sphgpu_ensemble_begin(sphgpu_sim* s) {
  return guarded(s, [&] {
    SG_REQUIRE(s->nTot == 0 && !s->ensemble, SPHGPU_ERR_STATE, "ensemble mode needs a sim with no nodes");
    SG_REQUIRE(s->nLists <= 1, SPHGPU_ERR_STATE, "an ensemble holds independent simulations, not NodeLists");
    SG_REQUIRE(!s->hasMultipliers, SPHGPU_ERR_UNSUPPORTED, "ensembles take no Q multiplier fields");
    s->ensemble = true;
    s->nSeg = 0;
    s->deviceGhosts = true;    // every ghost row is made by the device boundaries
  });
}

int sphgpu_ensemble_add(sphgpu_sim* s, const sphgpu_member* m, const sphgpu_integrator_params* ip, int64_t* id) {
  return guarded(s, [&] {
    SG_REQUIRE(m && ip && id, SPHGPU_ERR_ARG, "null argument");
    *id = ensembleAdd(*s, *m, *ip);
  });
}

int sphgpu_ensemble_step(sphgpu_sim* s, const sphgpu_hydro_params* hp, const sphgpu_integrator_params* ip,
                         int* finished) {
  return guarded(s, [&] {
    SG_REQUIRE(hp && ip, SPHGPU_ERR_ARG, "null parameters");
    const int done = ensembleStep(*s, *hp, *ip);
    if (finished) *finished = done;
  });
}

int sphgpu_ensemble_set_dt_threshold(sphgpu_sim* s, double dtThreshold) {
  return guarded(s, [&] {
    SG_REQUIRE(s->ensemble, SPHGPU_ERR_STATE, "not an ensemble");
    SG_REQUIRE(dtThreshold >= 0.0, SPHGPU_ERR_ARG, "dt threshold must be >= 0");
    s->dtThreshold = dtThreshold;
  });
}

int sphgpu_ensemble_failed(sphgpu_sim* s, int cap, int64_t* ids, int32_t* reasons, int64_t* cycles, int* count) {
  return guarded(s, [&] {
    SG_REQUIRE(s->ensemble, SPHGPU_ERR_STATE, "not an ensemble");
    SG_REQUIRE(cap >= 0, SPHGPU_ERR_ARG, "negative capacity");
    const int nf = int(s->failedIds.size());
    if (count) *count = nf;
    for (int q = 0; q < std::min(cap, nf); ++q) {
      if (ids) ids[q] = s->failedIds[size_t(q)];
      if (reasons) reasons[q] = s->failedReasons[size_t(q)];
      if (cycles) cycles[q] = s->failedCycles[size_t(q)];
    }
  });
}

int sphgpu_ensemble_members(sphgpu_sim* s, int cap, int64_t* ids, double* time, double* lastDt, int64_t* cycle,
                            int64_t* first, int64_t* rows, int* count) {
  return guarded(s, [&] {
    SG_REQUIRE(s->ensemble, SPHGPU_ERR_STATE, "not an ensemble");
    SG_REQUIRE(cap >= 0, SPHGPU_ERR_ARG, "negative capacity");
    const int nm = int(s->members.size());
    if (count) *count = nm;
    for (int q = 0; q < std::min(cap, nm); ++q) {
      const Member& mb = s->members[size_t(q)];
      const SegState& g = s->segHost[size_t(q)];
      if (ids) ids[q] = mb.id;
      if (time) time[q] = g.t;
      if (lastDt) lastDt[q] = g.lastDt;
      if (cycle) cycle[q] = int64_t(g.cycle);
      if (first) first[q] = mb.first;
      if (rows) rows[q] = mb.rows;
    }
  });
}

int sphgpu_ensemble_download(sphgpu_sim* s, int64_t id, int field, double* host) {
  return guarded(s, [&] {
    SG_REQUIRE(s->ensemble, SPHGPU_ERR_STATE, "not an ensemble");
    SG_REQUIRE(host, SPHGPU_ERR_ARG, "null buffer");
    auto it = std::find_if(s->members.begin(), s->members.end(), [&](const Member& mb) { return mb.id == id; });
    SG_REQUIRE(it != s->members.end(), SPHGPU_ERR_ARG, "no member with that id");
    SG_REQUIRE(field >= 0 && field <= SPHGPU_F_OMEGA_GRADH, SPHGPU_ERR_ARG, "unknown state field id");
    const int comps = stateComponents(s->dim, field);
    DBuf& b = stateBuf(*s, field);
    SG_CHECK_HIP(hipMemcpyAsync(host, b.as<double>() + it->first * comps, sizeof(double) * size_t(it->rows) * comps,
                                hipMemcpyDeviceToHost, s->stream));
    SG_CHECK_HIP(hipStreamSynchronize(s->stream));
  });
}

int sphgpu_ensemble_set_clock(sphgpu_sim* s, int64_t id, double time, double lastDt, int64_t cycle) {
  return guarded(s, [&] {
    SG_REQUIRE(s->ensemble, SPHGPU_ERR_STATE, "not an ensemble");
    auto it = std::find_if(s->members.begin(), s->members.end(), [&](const Member& mb) { return mb.id == id; });
    SG_REQUIRE(it != s->members.end(), SPHGPU_ERR_ARG, "no member with that id");
    SG_REQUIRE(lastDt > 0.0 && cycle >= 0, SPHGPU_ERR_ARG, "bad clock");
    const size_t q = size_t(it - s->members.begin());
    SG_CHECK_HIP(hipMemcpyAsync(s->segHost.data(), s->segState.p, sizeof(SegState) * s->segHost.size(),
                                hipMemcpyDeviceToHost, s->stream));
    SG_CHECK_HIP(hipStreamSynchronize(s->stream));
    s->segHost[q].t = time;
    s->segHost[q].lastDt = lastDt;
    s->segHost[q].cycle = double(cycle);
    SG_CHECK_HIP(hipMemcpyAsync(static_cast<SegState*>(s->segState.p) + q, &s->segHost[q], sizeof(SegState),
                                hipMemcpyHostToDevice, s->stream));
    SG_CHECK_HIP(hipStreamSynchronize(s->stream));
  });
}

int sphgpu_ensemble_upload(sphgpu_sim* s, int64_t id, int field, const double* host) {
  return guarded(s, [&] {
    SG_REQUIRE(s->ensemble, SPHGPU_ERR_STATE, "not an ensemble");
    SG_REQUIRE(host, SPHGPU_ERR_ARG, "null buffer");
    auto it = std::find_if(s->members.begin(), s->members.end(), [&](const Member& mb) { return mb.id == id; });
    SG_REQUIRE(it != s->members.end(), SPHGPU_ERR_ARG, "no member with that id");
    SG_REQUIRE(field >= 0 && field <= SPHGPU_F_OMEGA_GRADH, SPHGPU_ERR_ARG, "unknown state field id");
    const int comps = stateComponents(s->dim, field);
    DBuf& b = stateBuf(*s, field);
    SG_CHECK_HIP(hipMemcpyAsync(b.as<double>() + it->first * comps, host, sizeof(double) * size_t(it->rows) * comps,
                                hipMemcpyHostToDevice, s->stream));
    SG_CHECK_HIP(hipStreamSynchronize(s->stream));
    s->recValid = false;
    s->derivsValid = false;
  });
}

int sphgpu_ensemble_remove(sphgpu_sim* s, int nIds, const int64_t* ids) {
  return guarded(s, [&] { ensembleRemove(*s, nIds, ids); });
}

int sphgpu_serve(sphgpu_sim* s, sphgpu_request* req, int n, const sphgpu_hydro_params* hp,
                 const sphgpu_integrator_params* ip, const sphgpu_serve_params* sp) {
  return guarded(s, [&] {
    SG_REQUIRE(hp && ip && sp, SPHGPU_ERR_ARG, "null parameters");
    serve(*s, req, n, *hp, *ip, *sp);
  });
}

int sphgpu_autotune(int on) {
  tuneEnable(on != 0);
  return SPHGPU_OK;
}

int sphgpu_autotune_export(char* buf, int64_t cap, int64_t* len) {
  try {
    const std::string t = tuneExport();
    if (len) *len = int64_t(t.size());
    if (buf && cap > 0) {
      const size_t n = std::min<size_t>(t.size(), size_t(cap - 1));
      std::memcpy(buf, t.data(), n);
      buf[n] = '\0';
    }
    return SPHGPU_OK;
  } catch (...) {
    return SPHGPU_ERR_ARG;
  }
}

int sphgpu_autotune_import(const char* text, int* entries) {
  if (!text) return SPHGPU_ERR_ARG;
  try {
    const int n = tuneImport(text);
    if (entries) *entries = n;
    return SPHGPU_OK;
  } catch (...) {
    return SPHGPU_ERR_ARG;
  }
}

int64_t sphgpu_ensemble_bytes_per_node(int dim) {
  if (dim < 1 || dim > 3) return -1;
  return ensembleBytesPerNode(dim);
}

int sphgpu_copy_probe(sphgpu_sim* s, int64_t bytes, int repeats, double* ms) {
  return guarded(s, [&] {
    SG_REQUIRE(ms, SPHGPU_ERR_ARG, "null output");
    *ms = copyProbe(*s, bytes, repeats);
  });
}

int sphgpu_gather_probe(sphgpu_sim* s, int loads_per_neighbour, int repeats, double* ms) {
  return guarded(s, [&] {
    SG_REQUIRE(ms, SPHGPU_ERR_ARG, "null output");
    *ms = gatherProbe(*s, loads_per_neighbour, repeats);
  });
}

}  // extern "C"
