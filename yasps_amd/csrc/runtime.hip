// Error reporting, version and grow-on-demand device scratch.
#include "common.hpp"

#include <mutex>
#include <string>
#include <vector>

namespace {
thread_local std::string g_last_error;
std::mutex g_scratch_mutex;
struct Slot {
  void* ptr = nullptr;
  size_t bytes = 0;
};
std::vector<Slot> g_slots;
}  // namespace

extern "C" void ysp_set_error(const char* file, int line, const char* msg) {
  char buf[512];
  std::snprintf(buf, sizeof(buf), "%s:%d: %s", file, line, msg);
  g_last_error = buf;
}

extern "C" const char* ysp_last_error(void) { return g_last_error.c_str(); }

extern "C" int ysp_version(void) { return 1; }

// Deferred flag checks: an entry point whose kernels raise a device flag (e.g. a slot
// layout the kernel does not implement) copies the flag word to pinned host memory in
// stream order instead of synchronising; the next ysp_check_deferred / ysp_sync on that
// stream examines every flag it queued.
namespace {
constexpr int kDeferSlots = 256;
struct Deferred {
  hipStream_t stream;
  int slot;
  const char* what;
};
std::mutex g_defer_mutex;
int* g_defer_host = nullptr;          // pinned, kDeferSlots words
int g_defer_next = 0;
std::vector<Deferred> g_deferred;
}  // namespace

namespace ysp {
int defer_flag_check(const int* flag, hipStream_t s, const char* what) {
  {
    std::lock_guard<std::mutex> lock(g_defer_mutex);
    if (!g_defer_host && hipHostMalloc(reinterpret_cast<void**>(&g_defer_host),
                                       kDeferSlots * sizeof(int), hipHostMallocDefault) != hipSuccess)
      g_defer_host = nullptr;
    if (g_defer_host) {
      const int slot = g_defer_next;
      bool busy = false;
      for (const Deferred& d : g_deferred) busy = busy || d.slot == slot;
      if (!busy) {
        g_defer_next = (g_defer_next + 1) % kDeferSlots;
        g_defer_host[slot] = 0;
        YSP_CHECK(hipMemcpyAsync(g_defer_host + slot, flag, sizeof(int), hipMemcpyDeviceToHost, s));
        g_deferred.push_back({s, slot, what});
        return YSP_OK;
      }
    }
  }
  // no free slot: check now
  int h = 0;
  YSP_CHECK(hipMemcpyAsync(&h, flag, sizeof(int), hipMemcpyDeviceToHost, s));
  YSP_CHECK(hipStreamSynchronize(s));
  if (h) {
    ysp_set_error(__FILE__, __LINE__, what);
    return YSP_ERR_UNSUPPORTED;
  }
  return YSP_OK;
}
}  // namespace ysp

extern "C" int ysp_check_deferred(void* stream) {
  hipStream_t s = ysp::as_stream(stream);
  YSP_CHECK(hipStreamSynchronize(s));
  std::lock_guard<std::mutex> lock(g_defer_mutex);
  const char* failed = nullptr;
  std::vector<Deferred> keep;
  for (const Deferred& d : g_deferred) {
    if (d.stream != s) {
      keep.push_back(d);
      continue;
    }
    if (g_defer_host[d.slot] && !failed) failed = d.what;
  }
  g_deferred.swap(keep);
  if (failed) {
    ysp_set_error(__FILE__, __LINE__, failed);
    return YSP_ERR_UNSUPPORTED;
  }
  return YSP_OK;
}

extern "C" int ysp_sync(void* stream) { return ysp_check_deferred(stream); }

namespace ysp {
void* scratch(int slot, size_t bytes) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (slot >= static_cast<int>(g_slots.size())) g_slots.resize(slot + 1);
  Slot& s = g_slots[slot];
  if (bytes == 0) bytes = 256;
  if (s.bytes < bytes) {
    if (s.ptr) {
      // The previous buffer may still be read by queued work on any stream.
      (void)hipDeviceSynchronize();
      (void)hipFree(s.ptr);
      s.ptr = nullptr;
      s.bytes = 0;
    }
    size_t want = bytes + bytes / 4;
    const hipError_t e = hipMalloc(&s.ptr, want);
    if (e != hipSuccess) {
      s.ptr = nullptr;
      ysp_set_error(__FILE__, __LINE__, hipGetErrorString(e));
      return nullptr;
    }
    s.bytes = want;
  }
  return s.ptr;
}
}  // namespace ysp

extern "C" void ysp_release_scratch(void) {
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  (void)hipDeviceSynchronize();
  for (auto& s : g_slots) {
    if (s.ptr) (void)hipFree(s.ptr);
    s.ptr = nullptr;
    s.bytes = 0;
  }
}
