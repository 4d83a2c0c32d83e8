// Internal declarations shared by the qfx HIP translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <array>
#include <cstdint>
#include <map>
#include <string>
#include <vector>

#include "qfx.h"

struct QfxProfRec {
  int cls;
  hipEvent_t start, stop;
  double bytes, flops;
};

struct qfx_context {
  int device = 0;
  hipStream_t stream = nullptr;
  // Arena: chunks from hipMalloc, bump-allocated; freed blocks are kept in
  // per-size free lists (sizes rounded to the arena's size classes).
  struct Chunk {
    char* base;
    size_t size, used;
  };
  std::vector<Chunk> chunks;
  size_t chunk_bytes = size_t(1) << 30;
  std::map<size_t, std::vector<void*>> free_lists;
  std::map<void*, size_t> live;  // pointer -> rounded size
  size_t in_use = 0;
  // Device workspace for split reductions (grows on demand).
  void* work = nullptr;
  size_t work_bytes = 0;
  // Profiling.
  bool profiling = false;
  std::vector<QfxProfRec> pending;
  std::vector<hipEvent_t> event_pool;
  uint64_t launches[QFX_K_NUM_CLASSES] = {};
  double ms[QFX_K_NUM_CLASSES] = {};
  double bytes[QFX_K_NUM_CLASSES] = {};
  double flops[QFX_K_NUM_CLASSES] = {};
  hipEvent_t timer_start = nullptr, timer_stop = nullptr;
  // Autotuner of the large-tile GEMM layouts: (precision, operand form, m,
  // n, k) -> layout, measured on first use while enabled.
  bool capturing = false;  // between qfx_graph_begin and qfx_graph_end
  // Pinned staging area of qfx_copy_h2d_async: filled from offset
  // stage_off on, rewound only after the stream has been synchronized.
  // The copies are batched: `uploads` lists the staged copies not yet
  // launched (destination, offset in the staging area, bytes), `upload_at`
  // their destination ranges (start -> end); one scatter launch moves them
  // all before the next call that puts work on the stream.
  unsigned char* stage = nullptr;
  void* stage_dev = nullptr;  // the staging area's device address
  size_t stage_off = 0;
  struct Upload {
    uint64_t dst, src_off, bytes;
  };
  std::vector<Upload> uploads;
  std::map<uintptr_t, uintptr_t> upload_at;
  // Grouped products: host staging of the problem list, the event after its
  // copy to the device, and the compute-unit count (grid of the launch).
  std::vector<unsigned char> group_host;
  hipEvent_t group_ev = nullptr;
  int cus = 0;
  bool autotune = false;
  std::map<std::array<uint64_t, 5>, int> tuned;
  hipEvent_t tune_ev[2] = {nullptr, nullptr};
  // This device's copy of the library's fixed scratch array (the device
  // address hipGetSymbolAddress returned at context creation), and whether
  // the call in progress must take its workspace from it instead of the
  // growable workspace above (qfx_cgemm_noalloc_c64).
  void* fixed_scratch = nullptr;
  bool use_fixed_scratch = false;
};

// Size of the fixed scratch array (one copy per GPU, held until the process
// ends) that the entry point allocating nothing uses as its workspace.
constexpr size_t kFixedScratchBytes = QFX_NOALLOC_SCRATCH_BYTES;

// Error plumbing: every entry point returns a status and records a message.
int qfx_set_error(int status, const std::string& msg);
int qfx_hip_error(hipError_t e, const char* what);

#define QFX_HIP(call)                                      \
  do {                                                     \
    hipError_t e_ = (call);                                \
    if (e_ != hipSuccess) return qfx_hip_error(e_, #call); \
  } while (0)

// Launches the batched staged uploads (qfx_copy_h2d_async) on the stream.
int qfx_flush_uploads(qfx_context* ctx);

// Start of every entry point that puts work on the stream, reads device
// memory from the host or waits for the stream: staged uploads go first.
#define QFX_CHECK_CTX(ctx)                                                \
  do {                                                                    \
    if (!(ctx)) return qfx_set_error(QFX_ERR_INVALID, "null context");    \
    if (!(ctx)->uploads.empty()) {                                        \
      const int flush_status_ = qfx_flush_uploads(ctx);                   \
      if (flush_status_ != QFX_OK) return flush_status_;                  \
    }                                                                     \
  } while (0)

// Entry points that only keep books (arena, profiling records, tables).
#define QFX_CHECK_CTX_NO_FLUSH(ctx)                                       \
  do {                                                                    \
    if (!(ctx)) return qfx_set_error(QFX_ERR_INVALID, "null context");    \
  } while (0)

// Bracket a launch with profiling events (no-op unless profiling is on).
struct QfxLaunchScope {
  qfx_context* ctx;
  QfxProfRec rec;
  bool active;
  QfxLaunchScope(qfx_context* c, int cls, double bytes, double flops);
  ~QfxLaunchScope();
};

// A recorded sequence of stream work (qfx_graph_begin / qfx_graph_end).
struct qfx_graph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  qfx_context* ctx = nullptr;
};

#define QFX_NOT_CAPTURING(ctx, what)                                                        \
  if ((ctx)->capturing)                                                                     \
  return qfx_set_error(QFX_ERR_INVALID, std::string(what) +                                 \
                                            ": not allowed while a graph is being recorded")

// Device workspace of at least `bytes` (reused across calls on the stream).
int qfx_workspace(qfx_context* ctx, size_t bytes, void** out);

// Host-side guard before a launch: when p lies in one of the context's arena
// chunks, [p, p + bytes) must lie inside one live (allocated, not freed)
// block; otherwise QFX_ERR_INVALID with a message naming `what`.  Memory
// outside the arena (allocated by the caller) is not checked.
int qfx_check_extent(qfx_context* ctx, const void* p, double bytes, const char* what);
#define QFX_EXTENT(ctx, p, bytes, what)                        \
  do {                                                         \
    const int extent_status_ = qfx_check_extent(ctx, p, bytes, what); \
    if (extent_status_ != QFX_OK) return extent_status_;       \
  } while (0)

// Offset of matrix index idx under a digit map (see qfx_index_map).  The
// map is a bit permutation, so for i0 a multiple of 2^b and d < 2^b,
// offset(i0 + d) = offset(i0) + offset(d).
__device__ __forceinline__ uint64_t map_offset(const qfx_index_map& d, uint64_t idx) {
  uint64_t off = 0;
  for (int j = 0; j < d.ndigits; ++j) {
    const int b = d.bits[j];
    off += (idx & ((uint64_t(1) << b) - 1)) << d.shift[j];
    idx >>= b;
  }
  return off;
}

// FP8 (E4M3, per-128x128-tile power-of-two scales) complex GEMM of
// qfx_fp8.hip: whether the shape is large enough for its tiles, and the
// launch (maps = {a_rows, a_cols, b_rows, b_cols} or null for row-major).
bool qfx_fp8_applies(uint64_t m, uint64_t n, uint64_t k);
int qfx_cgemm_fp8(qfx_context* ctx, const float2* A, const float2* B, float2* C, uint64_t m,
                  uint64_t n, uint64_t k, const qfx_index_map* maps);
