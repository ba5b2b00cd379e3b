// Library runtime: device selection, stream, allocations, scratch, profiling.
#include "common.hpp"

#include <cstring>
#include <map>
#include <mutex>

namespace gre {

namespace {
thread_local std::string g_last_error;
hipStream_t g_stream = nullptr;
hipStream_t g_own_stream = nullptr;
bool g_inited = false;
int g_device = 0;
uint64_t g_bytes = 0;

struct ScratchSlot { void* p = nullptr; size_t bytes = 0; };
ScratchSlot g_scratch[16];

struct ProfEntry { int64_t launches = 0; double ms = 0.0; std::vector<std::pair<hipEvent_t, hipEvent_t>> pending; };
bool g_prof_enabled = false;
std::map<std::string, int> g_prof_ids;
std::vector<std::string> g_prof_names;
std::vector<ProfEntry> g_prof;

void prof_drain(ProfEntry& pe) {
  for (auto& ev : pe.pending) {
    float ms = 0.f;
    GRE_HIP(hipEventSynchronize(ev.second));
    GRE_HIP(hipEventElapsedTime(&ms, ev.first, ev.second));
    pe.ms += ms;
    pe.launches++;
    (void)hipEventDestroy(ev.first);
    (void)hipEventDestroy(ev.second);
  }
  pe.pending.clear();
}
}  // namespace

void set_last_error(const std::string& m) { g_last_error = m; }

void ensure_init() {
  if (g_inited) return;
  int n = 0;
  GRE_HIP(hipGetDeviceCount(&n));
  GRE_CHECK(n > 0, GRE_DEV_FAILURE, "no HIP device visible");
  GRE_HIP(hipSetDevice(g_device));
  hipDeviceProp_t prop;
  GRE_HIP(hipGetDeviceProperties(&prop, g_device));
  GRE_CHECK(std::strncmp(prop.gcnArchName, "gfx950", 6) == 0, GRE_DEV_FAILURE,
            std::string("libgre_dev.so is built for gfx950; device is ") + prop.gcnArchName);
  GRE_HIP(hipStreamCreateWithFlags(&g_own_stream, hipStreamNonBlocking));
  g_stream = g_own_stream;
  g_inited = true;
}

hipStream_t stream() {
  ensure_init();
  return g_stream;
}

void DevBuf::alloc(size_t n, size_t align) {
  release();
  if (n == 0) n = 8;
  ensure_init();
  const size_t pad = align > 256 ? align : 0;
  hipError_t e = hipMalloc(&raw, n + pad);
  if (e != hipSuccess) {
    raw = nullptr;
    (void)hipGetLastError();
    throw Error(GRE_DEV_NOBUFFER, "hipMalloc(" + std::to_string(n + pad) + ") failed: " + hipGetErrorString(e));
  }
  p = pad ? reinterpret_cast<void*>((reinterpret_cast<uintptr_t>(raw) + pad - 1) / pad * pad) : raw;
  bytes = n;
  held = n + pad;
  g_bytes += held;
}

void DevBuf::release() {
  if (raw) {
    (void)hipFree(raw);
    g_bytes -= held;
  }
  raw = nullptr;
  p = nullptr;
  bytes = 0;
  held = 0;
}

void* scratch(int slot, size_t bytes) {
  ScratchSlot& s = g_scratch[slot];
  if (s.bytes < bytes) {
    if (s.p) {
      GRE_HIP(hipStreamSynchronize(stream()));
      (void)hipFree(s.p);
      g_bytes -= s.bytes;
    }
    size_t nb = bytes + bytes / 4 + 256;
    hipError_t e = hipMalloc(&s.p, nb);
    if (e != hipSuccess) {
      s.p = nullptr; s.bytes = 0;
      (void)hipGetLastError();
      throw Error(GRE_DEV_NOBUFFER, "scratch hipMalloc failed");
    }
    s.bytes = nb;
    g_bytes += nb;
  }
  return s.p;
}

bool prof_on() { return g_prof_enabled; }

ProfScope::ProfScope(const char* name) : id(-1) {
  if (!g_prof_enabled) return;
  auto it = g_prof_ids.find(name);
  if (it == g_prof_ids.end()) {
    id = (int)g_prof.size();
    g_prof_ids[name] = id;
    g_prof_names.push_back(name);
    g_prof.emplace_back();
  } else {
    id = it->second;
  }
  GRE_HIP(hipEventCreate(&a));
  GRE_HIP(hipEventCreate(&b));
  GRE_HIP(hipEventRecord(a, stream()));
}

ProfScope::~ProfScope() {
  if (id < 0) return;
  (void)hipEventRecord(b, stream());
  g_prof[id].pending.emplace_back(a, b);
  if (g_prof[id].pending.size() > 4096) {
    try { prof_drain(g_prof[id]); } catch (...) {}
  }
}

}  // namespace gre

using namespace gre;

extern "C" {

const char* gre_last_error(void) { return g_last_error.c_str(); }

int gre_init(int device) {
  return guard([&] {
    if (g_inited && device != g_device)
      throw Error(GRE_DEV_FAILURE, "gre_init: library already bound to another device");
    g_device = device;
    ensure_init();
  });
}

int gre_set_stream(void* s) {
  return guard([&] {
    ensure_init();
    g_stream = s ? static_cast<hipStream_t>(s) : g_own_stream;
  });
}

int gre_get_stream(void** s) {
  return guard([&] {
    GRE_CHECK(s, GRE_DEV_NULL, "null out pointer");
    *s = stream();
  });
}

int gre_sync(void) {
  return guard([&] { GRE_HIP(hipStreamSynchronize(stream())); });
}

int gre_prof_enable(int on) {
  return guard([&] { g_prof_enabled = on != 0; });
}

int gre_prof_reset(void) {
  return guard([&] {
    GRE_HIP(hipStreamSynchronize(stream()));
    for (auto& pe : g_prof) {
      for (auto& ev : pe.pending) { (void)hipEventDestroy(ev.first); (void)hipEventDestroy(ev.second); }
      pe.pending.clear();
      pe.launches = 0;
      pe.ms = 0.0;
    }
  });
}

int gre_prof_get(const char* name, int64_t* launches, double* total_ms) {
  return guard([&] {
    GRE_CHECK(name && launches && total_ms, GRE_DEV_NULL, "null argument");
    auto it = g_prof_ids.find(name);
    if (it == g_prof_ids.end()) { *launches = 0; *total_ms = 0.0; return; }
    ProfEntry& pe = g_prof[it->second];
    prof_drain(pe);
    *launches = pe.launches;
    *total_ms = pe.ms;
  });
}

int gre_mem_in_use(uint64_t* bytes) {
  return guard([&] {
    GRE_CHECK(bytes, GRE_DEV_NULL, "null out pointer");
    *bytes = g_bytes;
  });
}

}  // extern "C"
