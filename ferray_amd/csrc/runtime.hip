// Host runtime of the gfx950 GEMM library: device selection, per-thread
// streams, per-stream grow-only workspaces, device-memory helpers, seeded
// device generators and the thread-local status of drop-in calls.
#include <atomic>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <unordered_map>

#include <sys/syscall.h>
#include <unistd.h>

#include "common.hpp"

namespace fg {

namespace {
std::once_flag g_init_once;
int g_init_status = FERRAY_ERR_NO_DEVICE;
int g_device = 0;

struct ThreadState {
  hipStream_t stream = nullptr;
  // A thread that ends releases its stream and workspaces; the main
  // thread's go with the process (on_main_thread).
  ~ThreadState() {
    if (stream && !on_main_thread()) {
      release_workspace(stream);
      (void)hipStreamDestroy(stream);
    }
  }
};
thread_local ThreadState t_state;

struct Workspace {
  void* ptr = nullptr;
  size_t bytes = 0;
};
struct StreamWorkspaces {
  Workspace slot[WS_SLOTS];
};
std::mutex g_ws_mu;
std::unordered_map<hipStream_t, StreamWorkspaces> g_ws;

void init_once() {
  int count = 0;
  if (hipGetDeviceCount(&count) != hipSuccess || count <= 0) {
    g_init_status = FERRAY_ERR_NO_DEVICE;
    return;
  }
  const char* env = std::getenv("FERRAY_DEVICE");
  int dev = env ? std::atoi(env) : 0;
  if (dev < 0 || dev >= count) dev = 0;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) {
    g_init_status = FERRAY_ERR_NO_DEVICE;
    return;
  }
  // Kernels are built for gfx950 only.
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    g_init_status = FERRAY_ERR_UNSUPPORTED;
    return;
  }
  g_device = dev;
  g_init_status = FERRAY_OK;
}
}  // namespace

bool on_main_thread() { return syscall(SYS_gettid) == getpid(); }

int ensure_device() {
  std::call_once(g_init_once, init_once);
  if (g_init_status != FERRAY_OK) return g_init_status;
  int cur = -1;
  if (hipGetDevice(&cur) != hipSuccess || cur != g_device) FG_CHECK(hipSetDevice(g_device));
  return FERRAY_OK;
}

hipStream_t resolve_stream(void* stream) {
  if (stream) return reinterpret_cast<hipStream_t>(stream);
  if (!t_state.stream) {
    if (hipStreamCreateWithFlags(&t_state.stream, hipStreamNonBlocking) != hipSuccess)
      t_state.stream = nullptr;
  }
  return t_state.stream;
}

void release_workspace(hipStream_t stream) {
  std::lock_guard<std::mutex> lk(g_ws_mu);
  auto it = g_ws.find(stream);
  if (it == g_ws.end()) return;
  bool synced = false;
  for (Workspace& w : it->second.slot) {
    if (!w.ptr) continue;
    if (!synced) (void)hipStreamSynchronize(stream);
    synced = true;
    (void)hipFree(w.ptr);
  }
  g_ws.erase(it);
}

int workspace(hipStream_t stream, size_t bytes, void** out, int slot) {
  if (slot < 0 || slot >= WS_SLOTS) return FERRAY_ERR_INVALID;
  std::lock_guard<std::mutex> lk(g_ws_mu);
  Workspace& w = g_ws[stream].slot[slot];
  if (w.bytes < bytes) {
    if (w.ptr) {
      FG_CHECK(hipStreamSynchronize(stream));
      FG_CHECK(hipFree(w.ptr));
      w.ptr = nullptr;
      w.bytes = 0;
    }
    size_t want = bytes + bytes / 4 + 4096;
    FG_CHECK(hipMalloc(&w.ptr, want));
    w.bytes = want;
  }
  *out = w.ptr;
  return FERRAY_OK;
}

// ---------------------------------------------------------------------
// Seeded generators
// ---------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}
__device__ __forceinline__ uint64_t splitmix(uint64_t seed, uint64_t i) {
  return mix64(seed + (i + 1) * 0x9E3779B97F4A7C15ULL);
}
__device__ __forceinline__ double unif_pm1(uint64_t r) {
  return (double)(r >> 11) * 0x1.0p-52 - 1.0;
}
__device__ __forceinline__ int64_t unif_int(uint64_t r, int64_t lo, int64_t hi) {
  uint64_t span = (uint64_t)(hi - lo + 1);
  return lo + (int64_t)(((r >> 32) * span) >> 32);
}
__device__ __forceinline__ double normal01(uint64_t seed, uint64_t i) {
  // Box-Muller on two decorrelated draws of element i.
  uint64_t r1 = splitmix(seed, 2 * i), r2 = splitmix(seed, 2 * i + 1);
  double u1 = ((double)(r1 >> 11) + 1.0) * 0x1.0p-53;  // (0, 1]
  double u2 = (double)(r2 >> 11) * 0x1.0p-53;
  return sqrt(-2.0 * log(u1)) * cospi(2.0 * u2);
}

__global__ void fill_kernel(void* dst, size_t count, int kind, uint64_t seed,
                            int64_t lo, int64_t hi) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  size_t stride = (size_t)gridDim.x * blockDim.x;
  for (; i < count; i += stride) {
    switch (kind) {
      case 0: ((double*)dst)[i] = unif_pm1(splitmix(seed, i)); break;
      case 1: ((double*)dst)[i] = normal01(seed, i); break;
      case 2: {
        double u = unif_pm1(splitmix(seed, 2 * i));
        int e = (int)unif_int(splitmix(seed, 2 * i + 1), lo, hi);
        ((double*)dst)[i] = ldexp(u, e);
        break;
      }
      case 3: ((float*)dst)[i] = (float)unif_pm1(splitmix(seed, i)); break;
      case 4: ((float*)dst)[i] = (float)normal01(seed, i); break;
      case 5: ((int8_t*)dst)[i] = (int8_t)unif_int(splitmix(seed, i), lo, hi); break;
      case 6: ((int16_t*)dst)[i] = (int16_t)unif_int(splitmix(seed, i), lo, hi); break;
      case 7: ((uint8_t*)dst)[i] = (uint8_t)unif_int(splitmix(seed, i), lo, hi); break;
      default: break;
    }
  }
}

// Device-to-device copy, 16 bytes per thread where alignment allows.
__global__ void copy_kernel(uint8_t* __restrict__ dst, const uint8_t* __restrict__ src,
                            size_t bytes) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t nvec = (((uintptr_t)dst | (uintptr_t)src) % 16 == 0) ? bytes / 16 : 0;
  typedef unsigned u4_t __attribute__((ext_vector_type(4)));
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nvec; i += stride)
    reinterpret_cast<u4_t*>(dst)[i] = reinterpret_cast<const u4_t*>(src)[i];
  for (size_t i = nvec * 16 + (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < bytes;
       i += stride)
    dst[i] = src[i];
}

}  // namespace fg

using namespace fg;

namespace {
thread_local int t_last_status = FERRAY_OK;
}

extern "C" {

void ferray_set_last_status_internal(int s) { t_last_status = s; }

int ferray_gemm_last_status(void) { return t_last_status; }

const char* ferray_status_string(int status) {
  switch (status) {
    case FERRAY_OK: return "ok";
    case FERRAY_ERR_NO_DEVICE: return "no usable gfx950 device";
    case FERRAY_ERR_OOM: return "device out of memory";
    case FERRAY_ERR_INVALID: return "invalid argument";
    case FERRAY_ERR_HIP: return "HIP runtime error";
    case FERRAY_ERR_UNSUPPORTED: return "unsupported configuration";
    case FERRAY_ERR_SCRATCH: return "problem exceeds the fixed scratch of the allocation-free entry";
    default: return "unknown status";
  }
}

bool ferray_gpu_available(void) { return ensure_device() == FERRAY_OK; }

size_t ferray_gemm_offload_min_dim(int dtype) {
  static const size_t kMeasured[] = {128, 96, 96, 64, 96};  // F64, F32, C64, C32, I8
  if (dtype < FERRAY_DTYPE_F64 || dtype > FERRAY_DTYPE_I8) return 0;
  if (const char* env = std::getenv("FERRAY_OFFLOAD_MIN_DIM")) {
    const long long v = std::atoll(env);
    if (v > 0) return (size_t)v;
  }
  return kMeasured[dtype];
}

int ferray_dev_malloc(size_t bytes, void** out) {
  int s = ensure_device();
  if (s) return s;
  if (!out) return FERRAY_ERR_INVALID;
  *out = nullptr;
  if (bytes == 0) bytes = 1;
  FG_CHECK(hipMalloc(out, bytes));
  return FERRAY_OK;
}

int ferray_dev_free(void* ptr) {
  int s = ensure_device();
  if (s) return s;
  if (ptr) FG_CHECK(hipFree(ptr));
  return FERRAY_OK;
}

int ferray_dev_upload(void* dst, const void* src, size_t bytes, void* stream) {
  int s = ensure_device();
  if (s) return s;
  if (!bytes) return FERRAY_OK;
  hipStream_t st = resolve_stream(stream);
  if (!st) return FERRAY_ERR_HIP;
  if ((s = host_upload(dst, src, bytes, st))) return s;
  FG_CHECK(hipStreamSynchronize(st));
  return FERRAY_OK;
}

int ferray_dev_download(void* dst, const void* src, size_t bytes, void* stream) {
  int s = ensure_device();
  if (s) return s;
  if (!bytes) return FERRAY_OK;
  hipStream_t st = resolve_stream(stream);
  if (!st) return FERRAY_ERR_HIP;
  if ((s = host_download(dst, src, bytes, st))) return s;
  FG_CHECK(hipStreamSynchronize(st));
  return FERRAY_OK;
}

int ferray_dev_copy(void* dst, const void* src, size_t bytes, void* stream) {
  int s = ensure_device();
  if (s) return s;
  if (!bytes) return FERRAY_OK;
  if (!dst || !src) return FERRAY_ERR_INVALID;
  size_t blocks = cdiv(bytes / 16 + 1, 256);
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(copy_kernel, dim3((unsigned)blocks), dim3(256), 0, resolve_stream(stream),
                     reinterpret_cast<uint8_t*>(dst), reinterpret_cast<const uint8_t*>(src), bytes);
  FG_CHECK(hipGetLastError());
  return FERRAY_OK;
}

int ferray_dev_memset(void* dst, int value, size_t bytes, void* stream) {
  int s = ensure_device();
  if (s) return s;
  if (!bytes) return FERRAY_OK;
  FG_CHECK(hipMemsetAsync(dst, value, bytes, resolve_stream(stream)));
  return FERRAY_OK;
}

int ferray_stream_sync(void* stream) {
  int s = ensure_device();
  if (s) return s;
  FG_CHECK(hipStreamSynchronize(resolve_stream(stream)));
  return FERRAY_OK;
}

void* ferray_default_stream(void) {
  if (ensure_device() != FERRAY_OK) return nullptr;
  return resolve_stream(nullptr);
}

int ferray_event_create(void** out) {
  int s = ensure_device();
  if (s) return s;
  if (!out) return FERRAY_ERR_INVALID;
  hipEvent_t ev;
  FG_CHECK(hipEventCreate(&ev));
  *out = ev;
  return FERRAY_OK;
}

int ferray_event_record(void* event, void* stream) {
  int s = ensure_device();
  if (s) return s;
  FG_CHECK(hipEventRecord(reinterpret_cast<hipEvent_t>(event), resolve_stream(stream)));
  return FERRAY_OK;
}

int ferray_event_elapsed_ms(void* start, void* stop, float* ms) {
  int s = ensure_device();
  if (s) return s;
  FG_CHECK(hipEventSynchronize(reinterpret_cast<hipEvent_t>(stop)));
  FG_CHECK(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start),
                               reinterpret_cast<hipEvent_t>(stop)));
  return FERRAY_OK;
}

int ferray_event_destroy(void* event) {
  int s = ensure_device();
  if (s) return s;
  FG_CHECK(hipEventDestroy(reinterpret_cast<hipEvent_t>(event)));
  return FERRAY_OK;
}

int ferray_device_sync(void) {
  int s = ensure_device();
  if (s) return s;
  FG_CHECK(hipDeviceSynchronize());
  return FERRAY_OK;
}

int ferray_dev_fill(void* dst, size_t count, int kind, uint64_t seed, int64_t lo,
                    int64_t hi, void* stream) {
  int s = ensure_device();
  if (s) return s;
  if (kind < 0 || kind > 7 || (!dst && count)) return FERRAY_ERR_INVALID;
  if (!count) return FERRAY_OK;
  if ((kind == 2 || kind >= 5) && hi < lo) return FERRAY_ERR_INVALID;
  size_t blocks = cdiv(count, 256);
  if (blocks > 65536) blocks = 65536;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)blocks), dim3(256), 0, resolve_stream(stream), dst,
                     count, kind, seed, lo, hi);
  FG_CHECK(hipGetLastError());
  return FERRAY_OK;
}

}  // extern "C"
