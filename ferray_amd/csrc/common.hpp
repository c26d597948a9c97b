// Shared definitions for the gfx950 GEMM library: status codes, error
// plumbing and the MFMA operand vector types.
#pragma once
#include <hip/hip_runtime.h>
#include <cstddef>
#include <cstdint>

#include "../../include/ferray_gemm.h"

namespace fg {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef int i4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef _Float16 h8 __attribute__((ext_vector_type(8)));

#define FG_CHECK(expr)                                     \
  do {                                                     \
    hipError_t _e = (expr);                                \
    if (_e != hipSuccess) return ::fg::map_hip_error(_e);  \
  } while (0)

inline int map_hip_error(hipError_t e) {
  if (e == hipErrorOutOfMemory) return FERRAY_ERR_OOM;
  if (e == hipErrorNoDevice || e == hipErrorInvalidDevice) return FERRAY_ERR_NO_DEVICE;
  if (e == hipErrorInvalidValue) return FERRAY_ERR_INVALID;
  return FERRAY_ERR_HIP;
}

// Resolves a caller stream argument: NULL means the calling thread's own
// library stream (created on first use, default priority, non-blocking).
hipStream_t resolve_stream(void* stream);

// Grow-only device scratch owned by the library, WS_SLOTS buffers per
// stream.  Work ordered on `stream` may use the returned memory until the
// next call for the same stream and slot; growing synchronises that stream
// first.  Slot 0 serves the packing and split passes (and TRSM, which keeps
// slot-0 pointers while it calls the real GEMM); slot WS_PAD is used only by
// the real GEMM's padded copies, so a caller holding slot-0 memory may call
// it.
constexpr int WS_SLOTS = 2, WS_PAD = 1;
int workspace(hipStream_t stream, size_t bytes, void** out, int slot = 0);
// Frees the workspace of a stream that is about to be destroyed.
void release_workspace(hipStream_t stream);

// Runtime initialisation on the calling thread (device selection).
int ensure_device();

// True on the process's main thread.  Thread-local resources make no HIP
// calls when the main thread's copies are destroyed: that happens while the
// process tears down, when the runtime (or a tool layered on it) may already
// be gone, so they are left to the process's end.
bool on_main_thread();

// Pageable host memory <-> device, enqueued on `stream`.  Copies of 8 MiB
// and more go through the calling thread's pinned bounce buffers; the upload
// returns once the host buffer may be reused, the download once the host
// buffer holds the data (smaller downloads only enqueue: synchronise the
// stream before reading them).
int host_upload(void* dst, const void* src, size_t bytes, hipStream_t stream);
int host_download(void* dst, const void* src, size_t bytes, hipStream_t stream);

inline size_t cdiv(size_t a, size_t b) { return (a + b - 1) / b; }

// A grid's y extent is at most 65535 blocks.  Kernels whose grid y walks the
// k dimension (k / 64 tiles reach it at k = 4 Mi) take the first y index of
// their launch as an argument y0 and add it to blockIdx.y; for_grid_y calls
// launch(y0, ny) for consecutive slices of at most 65535 covering [0, total).
constexpr size_t kMaxGridY = 65535;
template <class F>
inline void for_grid_y(size_t total, F&& launch) {
  for (size_t y0 = 0; y0 < total; y0 += kMaxGridY)
    launch((int)y0, (unsigned)(total - y0 < kMaxGridY ? total - y0 : kMaxGridY));
}

}  // namespace fg
