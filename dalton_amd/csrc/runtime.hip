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

void check_span(const void* p, long lo, long hi, size_t esz, const char* what) {
  if (hi < lo) return;  // empty
  std::lock_guard<std::mutex> lk(g_mu);
  const uintptr_t a = reinterpret_cast<uintptr_t>(p);
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
