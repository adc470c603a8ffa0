// Autotuner of launch shapes.
//
// A kernel family with several launch variants (workgroup size, where its
// table lives) asks for a variant per workload class: the kernel, the
// dimension, the arithmetic, the log2 bucket of the row length and of the
// node count.  The first time a class is seen with tuning on, each variant is
// launched twice on the live inputs and timed with HIP events on the sim's
// stream (the second launch counts, so code loading and cold caches do not);
// the fastest is kept for the rest of the process.  The variants of a family
// compute the same values in the same order, so tuning never changes a
// result.  Decisions can be exported as text and imported again, so that
// measured choices are committed and reused without timing.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <mutex>
#include <sstream>
#include "internal.hh"

namespace sphgpu {

namespace {
struct Choice {
  int variant = 0;
  std::vector<float> ms;       // per variant (empty when imported)
};
std::mutex gMu;
std::map<TuneKey, Choice> gTable;
bool gEnabled = false;

int bucket(double x) { return x > 0.0 ? int(std::floor(std::log2(x))) : -1; }
}  // namespace

bool operator<(const TuneKey& a, const TuneKey& b) {
  return std::tie(a.kernel, a.dim, a.precision, a.rowBucket, a.sizeBucket) <
         std::tie(b.kernel, b.dim, b.precision, b.rowBucket, b.sizeBucket);
}

TuneKey tuneKey(int kernel, int dim, int precision, double rowLength, int64_t nodes) {
  return TuneKey{kernel, dim, precision, bucket(rowLength), bucket(double(nodes))};
}

void tuneEnable(bool on) {
  std::lock_guard<std::mutex> lock(gMu);
  gEnabled = on;
}

int tunedVariant(Sim& s, const TuneKey& key, int nVariants, int fallback, const std::function<void(int)>& launch,
                 bool* timed) {
  if (timed) *timed = false;
  {
    std::lock_guard<std::mutex> lock(gMu);
    auto it = gTable.find(key);
    if (it != gTable.end() && it->second.variant < nVariants) return it->second.variant;
    if (!gEnabled || s.capturing || nVariants <= 1) return fallback;
  }
  Choice c;
  c.ms.assign(size_t(nVariants), 0.0f);
  hipEvent_t e0 = s.ev[12], e1 = s.ev[13];
  for (int v = 0; v < nVariants; ++v) {
    for (int rep = 0; rep < 2; ++rep) {
      SG_CHECK_HIP(hipEventRecord(e0, s.stream));
      launch(v);
      SG_CHECK_HIP(hipEventRecord(e1, s.stream));
      SG_CHECK_HIP(hipEventSynchronize(e1));
      SG_CHECK_HIP(hipEventElapsedTime(&c.ms[size_t(v)], e0, e1));
    }
  }
  c.variant = int(std::min_element(c.ms.begin(), c.ms.end()) - c.ms.begin());
  if (timed) *timed = true;
  std::lock_guard<std::mutex> lock(gMu);
  gTable[key] = c;
  return c.variant;
}

std::string tuneExport() {
  std::lock_guard<std::mutex> lock(gMu);
  std::ostringstream o;
  o << "# kernel dim precision rowBucket sizeBucket variant [ms per variant]\n";
  for (const auto& kv : gTable) {
    const TuneKey& k = kv.first;
    o << k.kernel << ' ' << k.dim << ' ' << k.precision << ' ' << k.rowBucket << ' ' << k.sizeBucket << ' '
      << kv.second.variant;
    for (float m : kv.second.ms) {
      char b[32];
      std::snprintf(b, sizeof(b), " %.4f", double(m));
      o << b;
    }
    o << '\n';
  }
  return o.str();
}

int tuneImport(const std::string& text) {
  std::istringstream in(text);
  std::string line;
  int n = 0;
  std::lock_guard<std::mutex> lock(gMu);
  while (std::getline(in, line)) {
    if (line.empty() || line[0] == '#') continue;
    std::istringstream l(line);
    TuneKey k{};
    Choice c;
    if (!(l >> k.kernel >> k.dim >> k.precision >> k.rowBucket >> k.sizeBucket >> c.variant)) continue;
    float m;
    while (l >> m) c.ms.push_back(m);
    gTable[k] = c;
    ++n;
  }
  return n;
}

}  // namespace sphgpu
