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

extern "C" int ysp_sync(void* stream) {
  YSP_CHECK(hipStreamSynchronize(ysp::as_stream(stream)));
  return YSP_OK;
}

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
