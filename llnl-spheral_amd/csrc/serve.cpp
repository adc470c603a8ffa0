// Request loop over an ensemble: continuous batching of independent
// simulations.  Requests become visible at their arrival time (seconds after
// the loop starts, on the host's steady clock), are admitted first come first
// served while the ensemble's footprint stays within the HBM budget and a
// member slot is free, are stepped together with every other resident
// member, and leave as soon as they reach their end time (final state copied
// out, rows compacted away).  A request that alone exceeds the budget is
// admitted when the ensemble is empty.
#include <algorithm>
#include <chrono>
#include <thread>
#include <vector>
#include "internal.hh"

namespace sphgpu {

int64_t memberFootprint(const Sim& s, const sphgpu_member& m, const sphgpu_serve_params& sp) {
  return int64_t(double(ensembleBytesPerNode(s.dim)) * double(m.n) * std::max(1.0, sp.ghostFactor));
}

void serve(Sim& s, sphgpu_request* req, int n, const sphgpu_hydro_params& hp, const sphgpu_integrator_params& ip,
           const sphgpu_serve_params& sp) {
  SG_REQUIRE(s.ensemble && s.nSeg == 0, SPHGPU_ERR_STATE, "serving needs an empty ensemble");
  SG_REQUIRE(n >= 0 && (n == 0 || req), SPHGPU_ERR_ARG, "bad request list");
  SG_REQUIRE(sp.maxMembers > 0, SPHGPU_ERR_ARG, "maxMembers must be positive");
  int64_t budget = sp.hbmBudget;
  if (budget <= 0) {
    size_t freeB = 0, totalB = 0;
    SG_CHECK_HIP(hipMemGetInfo(&freeB, &totalB));
    budget = int64_t(0.8 * double(freeB));
  }
  SG_REQUIRE(sp.dtThreshold >= 0.0, SPHGPU_ERR_ARG, "dt threshold must be >= 0");
  std::vector<int> order(static_cast<size_t>(n));
  for (int q = 0; q < n; ++q) {
    order[size_t(q)] = q;
    req[q].admitted = req[q].finished = -1.0;
    req[q].steps = 0;
    req[q].status = SPHGPU_MEMBER_OK;
  }
  const double thresholdBefore = s.dtThreshold;
  struct Restore {      // the ensemble's own threshold again, also when a step throws
    Sim& s;
    double v;
    ~Restore() { s.dtThreshold = v; }
  } restore{s, thresholdBefore};
  s.dtThreshold = sp.dtThreshold;
  std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return req[a].arrival < req[b].arrival; });
  using clock = std::chrono::steady_clock;
  const auto t0 = clock::now();
  auto elapsed = [&] { return std::chrono::duration<double>(clock::now() - t0).count(); };

  struct Active { int64_t id; int q; int64_t bytes; };
  std::vector<Active> active;
  int64_t used = 0;
  size_t next = 0;
  int done = 0;
  while (done < n) {
    // admission, first come first served
    while (next < order.size()) {
      sphgpu_request& r = req[order[next]];
      if (r.arrival > elapsed()) break;
      const int64_t fp = memberFootprint(s, r.member, sp);
      const bool room = int(active.size()) < sp.maxMembers && used + fp <= budget;
      if (!room && !active.empty()) break;
      const int64_t id = ensembleAdd(s, r.member, ip);
      r.admitted = elapsed();
      active.push_back(Active{id, order[next], fp});
      used += fp;
      ++next;
    }
    if (active.empty()) {
      // idle until the next arrival
      const double wait = req[order[next]].arrival - elapsed();
      if (wait > 0.0) std::this_thread::sleep_for(std::chrono::duration<double>(wait));
      continue;
    }
    const int finished = ensembleStep(s, hp, ip);
    // members the step evicted: failed, the others go on unaffected
    for (size_t q = 0; q < s.failedIds.size(); ++q) {
      auto it = std::find_if(active.begin(), active.end(), [&](const Active& x) { return x.id == s.failedIds[q]; });
      SG_REQUIRE(it != active.end(), SPHGPU_ERR_STATE, "failed member not tracked");
      sphgpu_request& r = req[it->q];
      r.status = s.failedReasons[q];
      r.finished = elapsed();
      r.steps = s.failedCycles[q];
      used -= it->bytes;
      active.erase(it);
      ++done;
    }
    if (finished == 0) continue;
    // completions: every finished member's copies queued, one wait for all of
    // them, then the members removed together
    std::vector<int64_t> gone;
    std::vector<std::pair<int, int64_t>> ended;   // (request, cycles)
    for (size_t a = 0; a < s.members.size(); ++a) {
      const SegState& g = s.segHost[a];
      if (g.t < g.tEnd) continue;
      const Member& mb = s.members[a];
      auto it = std::find_if(active.begin(), active.end(), [&](const Active& x) { return x.id == mb.id; });
      SG_REQUIRE(it != active.end(), SPHGPU_ERR_STATE, "finished member not tracked");
      sphgpu_request& r = req[it->q];
      const int D = s.dim, NS = nsymOf(D);
      auto get = [&](DBuf& b, double* dst, int comps) {
        if (!dst) return;
        SG_CHECK_HIP(hipMemcpyAsync(dst, b.as<double>() + mb.first * comps, sizeof(double) * size_t(mb.rows) * comps,
                                    hipMemcpyDeviceToHost, s.stream));
      };
      get(s.pos, r.outPos, D);
      get(s.vel, r.outVel, D);
      get(s.rho, r.outRho, 1);
      get(s.eps, r.outEps, 1);
      get(s.H, r.outH, NS);
      ended.emplace_back(it->q, int64_t(g.cycle));
      used -= it->bytes;
      gone.push_back(mb.id);
      active.erase(it);
    }
    SG_CHECK_HIP(hipStreamSynchronize(s.stream));
    const double tEnd = elapsed();
    for (const auto& e : ended) {
      req[e.first].finished = tEnd;
      req[e.first].steps = e.second;
      ++done;
    }
    ensembleRemove(s, int(gone.size()), gone.data());
  }
}

}  // namespace sphgpu
