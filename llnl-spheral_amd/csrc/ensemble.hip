// Ensembles: many independent simulations in one node set.
//
// Layout: member m's internal rows are the caller rows [first_m, first_m +
// rows_m), members in order; every row carries its member index in the `seg`
// field, and the ghost rows the device boundaries make inherit their
// control's.  The neighbour build gives each member the grid it would get
// alone (cells numbered member after member), the ghost selection uses each
// member's own hmax, and the integrator keeps a SegState (time, last dt, end
// time, gamma, this step's dt) per member.  The density and pair passes need
// nothing of this: they run over all slots of all members in one launch.
//
// A member's arithmetic does not depend on the others: its rows, their
// order, its ghosts, its dt and every sum are those of the same member alone
// (relative caller order is kept by additions and removals).
#include "internal.hh"
#include "dgeom.hh"
#include <cstring>
#include <algorithm>

namespace sphgpu {

namespace {

constexpr int kBlock = 256;
inline int nblocks(int64_t n, int bs = kBlock) { return int(std::max<int64_t>(1, (n + bs - 1) / bs)); }

__global__ void k_fill_rows(double* __restrict__ a, int64_t first, int64_t rows, double v) {
  const int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (r < rows) a[first + r] = v;
}

// Removal: new row r belongs to surviving member m (newFirst[m] <= r <
// newFirst[m+1]) and comes from old row oldFirst[m] + (r - newFirst[m]).
// setSeg: write the new member index instead of copying.
__global__ void k_member_gather(const double* __restrict__ src, double* __restrict__ dst, int comps,
                                const int64_t* __restrict__ newFirst, const int64_t* __restrict__ oldFirst, int nm,
                                int64_t rows, int setSeg) {
  const int64_t r = blockIdx.x * int64_t(blockDim.x) + threadIdx.x;
  if (r >= rows) return;
  int lo = 0, hi = nm - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (newFirst[mid] <= r) lo = mid; else hi = mid - 1;
  }
  if (setSeg) { dst[r] = double(lo); return; }
  const int64_t o = oldFirst[lo] + (r - newFirst[lo]);
  for (int c = 0; c < comps; ++c) dst[r * comps + c] = src[o * comps + c];
}

void pullSegState(Sim& s) {
  s.segHost.resize(size_t(s.nSeg));
  if (s.nSeg == 0) return;
  SG_CHECK_HIP(hipMemcpyAsync(s.segHost.data(), s.segState.p, sizeof(SegState) * size_t(s.nSeg),
                              hipMemcpyDeviceToHost, s.stream));
  SG_CHECK_HIP(hipStreamSynchronize(s.stream));
}

void pushSegState(Sim& s) {
  if (s.nSeg == 0) return;
  s.segState.ensure(sizeof(SegState) * size_t(s.nSeg + s.nSeg / 2 + 8));
  SG_CHECK_HIP(hipMemcpyAsync(s.segState.p, s.segHost.data(), sizeof(SegState) * size_t(s.nSeg),
                              hipMemcpyHostToDevice, s.stream));
  SG_CHECK_HIP(hipStreamSynchronize(s.stream));
}

void invalidate(Sim& s) {
  s.neighborsValid = false;
  s.recValid = false;
  s.derivsValid = false;
}

}  // namespace

int64_t ensembleAdd(Sim& s, const sphgpu_member& m, const sphgpu_integrator_params& ip) {
  SG_REQUIRE(s.ensemble, SPHGPU_ERR_STATE, "not an ensemble: call sphgpu_ensemble_begin first");
  SG_REQUIRE(m.n > 0 && m.pos && m.vel && m.mass && m.rho && m.eps && m.H, SPHGPU_ERR_ARG, "incomplete member");
  SG_REQUIRE(m.tEnd > 0.0 && m.gamma > 1.0, SPHGPU_ERR_ARG, "member needs tEnd > 0 and gamma > 1");
  SG_REQUIRE(s.nInt + m.n < (int64_t(1) << 30), SPHGPU_ERR_CAPACITY, "ensemble too large");
  hipStream_t st = s.stream;
  const int D = s.dim, NS = nsymOf(D);
  // ghost rows are remade at the next step
  s.nTot = s.nInt;
  const int64_t first = s.nInt, rows = m.n, nNew = first + rows;
  ensureStateRows(s, nNew);
  auto put = [&](DBuf& b, const double* src, int comps) {
    SG_CHECK_HIP(hipMemcpyAsync(b.as<double>() + first * comps, src, sizeof(double) * size_t(rows) * comps,
                                hipMemcpyHostToDevice, st));
  };
  put(s.pos, m.pos, D);
  put(s.vel, m.vel, D);
  put(s.mass, m.mass, 1);
  put(s.rho, m.rho, 1);
  put(s.eps, m.eps, 1);
  put(s.H, m.H, NS);
  hipLaunchKernelGGL(k_fill_rows, dim3(nblocks(rows)), dim3(kBlock), 0, st, s.omega.as<double>(), first, rows, 1.0);
  hipLaunchKernelGGL(k_fill_rows, dim3(nblocks(rows)), dim3(kBlock), 0, st, s.seg.as<double>(), first, rows,
                     double(s.nSeg));
  SG_CHECK_HIP(hipGetLastError());
  eosRows(s, first, rows, m.gamma, ip.Pmin, ip.Pmax, ip.pressureFloor);
  if (s.haveDtIn) {
    // the last step's dt inputs of a new member are zero (as before a first step)
    const size_t per = sizeof(double) * size_t(2 + D);
    s.dtIn.grow(per * size_t(nNew + nNew / 4), per * size_t(first), st);
    SG_CHECK_HIP(hipMemsetAsync(s.dtIn.as<char>() + per * size_t(first), 0, per * size_t(rows), st));
  }
  pullSegState(s);
  SegState g{};
  g.t = 0.0;
  g.lastDt = 1.0e-5;
  g.tEnd = m.tEnd;
  g.gamma = m.gamma;
  g.cycle = 0.0;
  s.segHost.push_back(g);
  const int64_t id = s.nextMemberId++;
  s.members.push_back(Member{id, first, rows});
  s.nSeg += 1;
  s.nInt = nNew;
  s.nTot = nNew;
  pushSegState(s);   // the whole mirror: growing the device array needs no copy
  invalidate(s);
  return id;
}

void ensembleRemove(Sim& s, int nIds, const int64_t* ids) {
  SG_REQUIRE(s.ensemble, SPHGPU_ERR_STATE, "not an ensemble");
  SG_REQUIRE(nIds >= 0 && (nIds == 0 || ids), SPHGPU_ERR_ARG, "bad id list");
  std::vector<char> drop(s.members.size(), 0);
  for (int q = 0; q < nIds; ++q) {
    auto it = std::find_if(s.members.begin(), s.members.end(), [&](const Member& mb) { return mb.id == ids[q]; });
    SG_REQUIRE(it != s.members.end(), SPHGPU_ERR_ARG, "no member with that id");
    drop[size_t(it - s.members.begin())] = 1;
  }
  if (nIds == 0) return;
  hipStream_t st = s.stream;
  pullSegState(s);
  std::vector<Member> keepM;
  std::vector<SegState> keepS;
  std::vector<int64_t> newFirst, oldFirst;
  int64_t rowsNew = 0;
  for (size_t q = 0; q < s.members.size(); ++q) {
    if (drop[q]) continue;
    Member mb = s.members[q];
    newFirst.push_back(rowsNew);
    oldFirst.push_back(mb.first);
    mb.first = rowsNew;
    rowsNew += mb.rows;
    keepM.push_back(mb);
    keepS.push_back(s.segHost[q]);
  }
  const int nm = int(keepM.size());
  if (nm > 0) {
    DBuf maps;
    maps.ensure(sizeof(int64_t) * size_t(2 * nm));
    SG_CHECK_HIP(hipMemcpyAsync(maps.p, newFirst.data(), sizeof(int64_t) * size_t(nm), hipMemcpyHostToDevice, st));
    SG_CHECK_HIP(hipMemcpyAsync(maps.as<int64_t>() + nm, oldFirst.data(), sizeof(int64_t) * size_t(nm),
                                hipMemcpyHostToDevice, st));
    const int64_t* dNew = maps.as<int64_t>();
    const int64_t* dOld = dNew + nm;
    for (int w = 0; w < kStateFields; ++w) {
      if (!stateFieldUsed(s, w)) continue;
      DBuf& f = stateField(s, w);
      const int comps = stateFieldComps(s, w);
      DBuf& spare = s.spare[w];
      spare.ensure(std::max<size_t>(f.bytes, 8));
      hipLaunchKernelGGL(k_member_gather, dim3(nblocks(rowsNew)), dim3(kBlock), 0, st, f.as<double>(),
                         spare.as<double>(), comps, dNew, dOld, nm, rowsNew, w == 12 ? 1 : 0);
      f.swap(spare);
    }
    if (s.haveDtIn) {
      s.dtInSpare.ensure(std::max<size_t>(s.dtIn.bytes, 8));
      hipLaunchKernelGGL(k_member_gather, dim3(nblocks(rowsNew)), dim3(kBlock), 0, st, s.dtIn.as<double>(),
                         s.dtInSpare.as<double>(), 2 + s.dim, dNew, dOld, nm, rowsNew, 0);
      s.dtIn.swap(s.dtInSpare);
    }
    SG_CHECK_HIP(hipGetLastError());
    SG_CHECK_HIP(hipStreamSynchronize(st));   // `maps` is released on return
  }
  s.members = keepM;
  s.segHost = keepS;
  s.nSeg = nm;
  s.nInt = rowsNew;
  s.nTot = rowsNew;
  pushSegState(s);
  invalidate(s);
}

// Failure isolation.  A member fails alone when its position or H holds a
// NaN or infinity at the step start (found by the neighbour build's
// per-member bounds, before anything of the step is written), when its state
// turns non-finite in the step's final update, or when its dt vote falls
// below the ensemble's dtThreshold.  It is evicted (rows compacted away,
// recorded in failedIds with its reason) and the others are unaffected: their
// rows, grids, ghosts, dt and sums never involve another member.  The
// reference has no per-simulation failure: Process::haltAll ends the run
// (src/Utilities/Process.cc:74-87) and a step below Integrator::dtThreshold is
// only reported (src/Integrator/Integrator.cc:190-197).
namespace {
void evictFailed(Sim& s, const std::vector<int>& rowsOrder, const std::vector<int>& reasons) {
  std::vector<int64_t> ids;
  for (size_t q = 0; q < rowsOrder.size(); ++q) {
    const int64_t id = s.members[size_t(rowsOrder[q])].id;
    ids.push_back(id);
    s.failedIds.push_back(id);
    s.failedReasons.push_back(reasons[q]);
    s.failedCycles.push_back(int64_t(s.segHost[size_t(rowsOrder[q])].cycle));
  }
  ensembleRemove(s, int(ids.size()), ids.data());
}
}  // namespace

int ensembleStep(Sim& s, const sphgpu_hydro_params& hp, const sphgpu_integrator_params& ip) {
  SG_REQUIRE(s.ensemble && s.nSeg > 0, SPHGPU_ERR_STATE, "the ensemble has no members");
  SG_REQUIRE(s.table.n > 0, SPHGPU_ERR_STATE, "no kernel table installed");
  SG_REQUIRE(ip.kernelExtent > 0.0 && ip.cfl > 0.0 && ip.dtGrowth > 0.0, SPHGPU_ERR_ARG, "bad integrator parameters");
  SG_REQUIRE(!s.hasMultipliers, SPHGPU_ERR_UNSUPPORTED, "ensembles take no Q multiplier fields");
  for (size_t q = 0; q < s.segHost.size(); ++q)
    SG_REQUIRE(s.segHost[q].t < s.segHost[q].tEnd, SPHGPU_ERR_STATE,
               "a member reached its end time: download and remove it before the next step");
  s.failedIds.clear();
  s.failedReasons.clear();
  s.failedCycles.clear();
  for (;;) {
    try {
      ensembleRk2(s, hp, ip);
      break;
    } catch (MemberFailure& f) {
      evictFailed(s, f.members, f.reasons);
      if (s.nSeg == 0) return 0;
    }
  }
  MemberFailure late;
  for (size_t q = 0; q < s.segHost.size(); ++q) {
    const unsigned f = s.segHost[q].fail;
    if (f) late.add(int(q), (f & SPHGPU_MEMBER_NONFINITE) ? SPHGPU_MEMBER_NONFINITE : int(f));
  }
  if (!late.members.empty()) evictFailed(s, late.members, late.reasons);
  int done = 0;
  for (const SegState& g : s.segHost) done += (g.t >= g.tEnd) ? 1 : 0;
  return done;
}

int64_t ensembleBytesPerNode(int dim) {
  const int D = dim, NS = nsymOf(D);
  const int64_t state = 8 * (2 * D + NS + 8);          // pos vel H m rho eps P cs omega seg hmax
  const int64_t spare = state;                          // compaction spares
  const int64_t rec = 128 + 64 + 128;                   // FP64, mixed and collapsed records
  const int64_t out = 8 * int64_t(derivTotalDoubles(D));
  const int64_t ell = 4 * (D == 3 ? 160 : (D == 2 ? 48 : 16)) + 4;
  const int64_t sortMaps = 4 * 8;                      // cellOf, perm, iperm, work, flags, scans
  const int64_t step = 8 * (2 * D + 2 + NS) + 8 * (2 + D) + 8;   // state0, dt inputs, grad-h omega0
  return state + spare + rec + out + ell + sortMaps + step;
}

}  // namespace sphgpu
