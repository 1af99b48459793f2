// Allocation registry and dry-run support (see common.h).
#include <map>
#include <mutex>

#include "common.h"

namespace ccgpu {

namespace {
// Deliberately never destroyed: buffers with static storage may be released
// during process exit, after ordinary statics are gone.
std::mutex& g_mu_ref() {
  static auto* m = new std::mutex;
  return *m;
}
#define g_mu g_mu_ref()
std::map<uintptr_t, size_t>& allocs() {
  static auto* m = new std::map<uintptr_t, size_t>;
  return *m;
}
uintptr_t g_next = uintptr_t(1) << 44;  // fake device addresses for dry runs
}  // namespace

bool capturing(hipStream_t s) {
  if (dryrun()) return false;
  hipStreamCaptureStatus st = hipStreamCaptureStatusNone;
  CCGPU_HIP(hipStreamIsCapturing(s, &st));
  return st == hipStreamCaptureStatusActive;
}

namespace {
std::map<hipStream_t, StreamScratch*>& scratch_map() {
  static auto* m = new std::map<hipStream_t, StreamScratch*>;
  return *m;
}
}  // namespace

StreamScratch& scratch_for(hipStream_t s) {
  std::lock_guard<std::mutex> lk(g_mu);
  StreamScratch*& p = scratch_map()[s];
  if (!p) p = new StreamScratch;
  return *p;
}

void release_scratch(hipStream_t s) {
  StreamScratch* p = nullptr;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = scratch_map().find(s);
    if (it == scratch_map().end()) return;
    p = it->second;
    scratch_map().erase(it);
  }
  delete p;
}

bool& dryrun() {
  static bool flag = false;
  return flag;
}

void* dry_alloc(size_t bytes) {
  std::lock_guard<std::mutex> lk(g_mu);
  const uintptr_t p = g_next;
  g_next += (bytes + (1u << 20)) & ~uintptr_t((1u << 20) - 1);
  return reinterpret_cast<void*>(p);
}

void register_alloc(const void* p, size_t bytes) {
  std::lock_guard<std::mutex> lk(g_mu);
  allocs()[reinterpret_cast<uintptr_t>(p)] = bytes;
}

void unregister_alloc(const void* p) {
  std::lock_guard<std::mutex> lk(g_mu);
  allocs().erase(reinterpret_cast<uintptr_t>(p));
}

namespace {
constexpr int kMaxExternal = 4;
thread_local uintptr_t t_ext_p[kMaxExternal];
thread_local size_t t_ext_n[kMaxExternal];
thread_local int t_ext_cnt = 0;
}  // namespace

ExternalSpans::ExternalSpans(const void* const* p, const size_t* bytes, int n) {
  CCGPU_REQUIRE(t_ext_cnt == 0 && n >= 0 && n <= kMaxExternal, "external spans: nested scope or too many spans");
  for (int i = 0; i < n; ++i) {
    t_ext_p[i] = reinterpret_cast<uintptr_t>(p[i]);
    t_ext_n[i] = bytes[i];
  }
  t_ext_cnt = n;
}

ExternalSpans::~ExternalSpans() { t_ext_cnt = 0; }

void check_span(const void* p, long lo, long hi, size_t esz, const char* what) {
  if (hi < lo) return;  // empty
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
  for (int i = 0; i < t_ext_cnt; ++i) {
    if (a < t_ext_p[i] || a >= t_ext_p[i] + t_ext_n[i]) continue;
    const long long first = (long long)a + (long long)lo * (long long)esz;
    const long long last = (long long)a + ((long long)hi + 1) * (long long)esz;
    if (first < (long long)t_ext_p[i] || last > (long long)(t_ext_p[i] + t_ext_n[i]))
      throw Error(1, std::string("bounds: ") + what + " exceeds the caller's array (span " + std::to_string(lo) +
                         ".." + std::to_string(hi) + " elements)");
    return;
  }
  std::lock_guard<std::mutex> lk(g_mu);
  auto& m = allocs();
  auto it = m.upper_bound(a);
  if (it == m.begin()) throw Error(1, std::string("bounds: ") + what + " not in a registered allocation");
  --it;
  const uintptr_t b0 = it->first, b1 = it->first + it->second;
  const long long first = (long long)a + (long long)lo * (long long)esz;
  const long long last = (long long)a + ((long long)hi + 1) * (long long)esz;
  if (first < (long long)b0 || last > (long long)b1)
    throw Error(1, std::string("bounds: ") + what + " exceeds its allocation (span " +
                       std::to_string(lo) + ".." + std::to_string(hi) + " elements, allocation " +
                       std::to_string(it->second / esz) + " elements, offset " +
                       std::to_string((long long)(a - b0) / (long long)esz) + ")");
}

}  // namespace ccgpu

namespace ccgpu {
namespace {
struct LaunchRec {
  const char* name;
  double flops;
  hipEvent_t e0, e1;
};
struct PhaseRec {
  const char* name;
  hipEvent_t e;
};
struct Prof {
  bool on = false;
  std::vector<LaunchRec> launches;
  std::vector<PhaseRec> phases;
};
Prof& prof() {
  static auto* p = new Prof;
  return *p;
}
hipEvent_t new_event(hipStream_t s) {
  hipEvent_t e;
  CCGPU_HIP(hipEventCreate(&e));
  CCGPU_HIP(hipEventRecord(e, s));
  return e;
}
}  // namespace

bool profiling() { return prof().on && !dryrun(); }

void prof_start() {
  Prof& p = prof();
  for (auto& l : p.launches) {
    (void)hipEventDestroy(l.e0);
    (void)hipEventDestroy(l.e1);
  }
  for (auto& f : p.phases) (void)hipEventDestroy(f.e);
  p.launches.clear();
  p.phases.clear();
  p.on = true;
}

int prof_begin(const char* kernel, double flops, hipStream_t s) {
  if (!profiling() || capturing(s)) return -1;
  Prof& p = prof();
  p.launches.push_back({kernel, flops, new_event(s), nullptr});
  return static_cast<int>(p.launches.size()) - 1;
}

void prof_end(int id, hipStream_t s) {
  if (id < 0 || !profiling()) return;
  prof().launches[id].e1 = new_event(s);
}

void prof_phase(const char* name, hipStream_t s) {
  if (!profiling() || capturing(s)) return;
  prof().phases.push_back({name, new_event(s)});
}

std::string prof_stop_json() {
  Prof& p = prof();
  if (dryrun()) return "{}";
  CCGPU_HIP(hipDeviceSynchronize());
  struct Agg {
    long n = 0;
    double ms = 0, flops = 0;
  };
  std::map<std::string, Agg> kern;
  for (auto& l : p.launches) {
    float ms = 0.f;
    if (l.e1) CCGPU_HIP(hipEventElapsedTime(&ms, l.e0, l.e1));
    Agg& a = kern[l.name];
    a.n += 1;
    a.ms += ms;
    a.flops += l.flops;
  }
  std::map<std::string, double> ph;
  for (size_t i = 1; i < p.phases.size(); ++i) {
    if (std::string(p.phases[i].name) == "start") continue;
    float ms = 0.f;
    CCGPU_HIP(hipEventElapsedTime(&ms, p.phases[i - 1].e, p.phases[i].e));
    ph[p.phases[i].name] += ms;
  }
  std::string js = "{\"kernels\": {";
  bool first = true;
  for (auto& kv : kern) {
    char buf[512];
    snprintf(buf, sizeof buf, "%s\"%s\": {\"count\": %ld, \"ms\": %.6f, \"flops\": %.6e}", first ? "" : ", ",
             kv.first.c_str(), kv.second.n, kv.second.ms, kv.second.flops);
    js += buf;
    first = false;
  }
  js += "}, \"phases\": {";
  first = true;
  for (auto& kv : ph) {
    char buf[256];
    snprintf(buf, sizeof buf, "%s\"%s\": %.6f", first ? "" : ", ", kv.first.c_str(), kv.second);
    js += buf;
    first = false;
  }
  js += "}}";
  prof_start();
  p.on = false;
  return js;
}
}  // namespace ccgpu
