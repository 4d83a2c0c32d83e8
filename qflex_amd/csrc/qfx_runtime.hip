// Context, device arena, copies, accumulation and timing of the qfx C-ABI.
#include <cstring>
#include <iterator>
#include <mutex>

#include "qfx_internal.h"

static thread_local std::string g_last_error;

int qfx_set_error(int status, const std::string& msg) {
  g_last_error = msg;
  return status;
}

int qfx_hip_error(hipError_t e, const char* what) {
  return qfx_set_error(QFX_ERR_HIP, std::string("HIP error ") + hipGetErrorName(e) + " (" +
                                        hipGetErrorString(e) + ") in " + what);
}

extern "C" const char* qfx_last_error(void) { return g_last_error.c_str(); }

// ---------------------------------------------------------------- context

// Workspace of the entry point that allocates nothing
// (qfx_cgemm_noalloc_c64): a fixed array of which every GPU holds its own
// copy from module load until the process ends.  Host code never uses the
// array's name as an address: it uses what hipGetSymbolAddress returns for
// the device (fixed_scratch_address).
static __device__ __align__(256) unsigned char scratch[kFixedScratchBytes];

// The device address of this device's `scratch`, looked up once per device
// with that device current (the caller has set it).
static int fixed_scratch_address(int device, void** out) {
  static std::mutex mu;
  static std::map<int, void*> per_device;
  std::lock_guard<std::mutex> lock(mu);
  auto it = per_device.find(device);
  if (it == per_device.end()) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return qfx_hip_error(e, "hipGetSymbolAddress(scratch)");
    if (!p) return qfx_set_error(QFX_ERR_HIP, "hipGetSymbolAddress(scratch) returned null");
    it = per_device.emplace(device, p).first;
  }
  *out = it->second;
  return QFX_OK;
}

extern "C" int qfx_context_create(int device, size_t arena_chunk_bytes, qfx_context** out) {
  if (!out) return qfx_set_error(QFX_ERR_INVALID, "null output pointer");
  int n = 0;
  QFX_HIP(hipGetDeviceCount(&n));
  if (device < 0 || device >= n)
    return qfx_set_error(QFX_ERR_INVALID, "device " + std::to_string(device) +
                                              " out of range (devices: " + std::to_string(n) + ")");
  QFX_HIP(hipSetDevice(device));
  void* fixed = nullptr;
  const int fs = fixed_scratch_address(device, &fixed);
  if (fs != QFX_OK) return fs;
  auto* ctx = new qfx_context();
  ctx->device = device;
  ctx->fixed_scratch = fixed;
  if (arena_chunk_bytes) ctx->chunk_bytes = arena_chunk_bytes;
  hipError_t e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete ctx;
    return qfx_hip_error(e, "hipStreamCreateWithFlags");
  }
  QFX_HIP(hipEventCreate(&ctx->timer_start));
  QFX_HIP(hipEventCreate(&ctx->timer_stop));
  *out = ctx;
  return QFX_OK;
}

extern "C" int qfx_context_destroy(qfx_context* ctx) {
  QFX_CHECK_CTX_NO_FLUSH(ctx);
  hipSetDevice(ctx->device);
  hipStreamSynchronize(ctx->stream);
  for (auto& r : ctx->pending) {
    hipEventDestroy(r.start);
    hipEventDestroy(r.stop);
  }
  for (auto ev : ctx->event_pool) hipEventDestroy(ev);
  for (auto& c : ctx->chunks) hipFree(c.base);
  if (ctx->timer_start) hipEventDestroy(ctx->timer_start);
  if (ctx->timer_stop) hipEventDestroy(ctx->timer_stop);
  for (auto ev : ctx->tune_ev)
    if (ev) hipEventDestroy(ev);
  if (ctx->group_ev) hipEventDestroy(ctx->group_ev);
  ctx->uploads.clear();  // their destinations are freed with the arena
  if (ctx->stage) hipHostFree(ctx->stage);
  hipStreamDestroy(ctx->stream);
  delete ctx;
  return QFX_OK;
}

// ---------------------------------------------------------------- arena

// Size classes: 256-byte granules below 64 KiB, otherwise quarter-octave
// steps (<= 19% rounding waste).
static size_t round_class(size_t bytes) {
  if (bytes == 0) bytes = 1;
  if (bytes <= 65536) return (bytes + 255) & ~size_t(255);
  size_t p = size_t(1) << (63 - __builtin_clzll(bytes));
  size_t step = p / 4;
  return (bytes + step - 1) / step * step;
}

extern "C" int qfx_alloc(qfx_context* ctx, size_t bytes, void** dptr) {
  QFX_CHECK_CTX_NO_FLUSH(ctx);
  if (!dptr) return qfx_set_error(QFX_ERR_INVALID, "null output pointer");
  size_t sz = round_class(bytes);
  auto it = ctx->free_lists.find(sz);
  if (it != ctx->free_lists.end() && !it->second.empty()) {
    *dptr = it->second.back();
    it->second.pop_back();
  } else {
    void* p = nullptr;
    for (auto& c : ctx->chunks) {
      if (c.size - c.used >= sz) {
        p = c.base + c.used;
        c.used += sz;
        break;
      }
    }
    if (!p && ctx->capturing)
      return qfx_set_error(QFX_ERR_INVALID,
                           "qfx_alloc: the device arena cannot grow while a graph is recorded");
    if (!p) {
      size_t csz = std::max(ctx->chunk_bytes, sz);
      char* base = nullptr;
      hipError_t e = hipMalloc(&base, csz);
      if (e != hipSuccess) {
        (void)hipGetLastError();
        return qfx_set_error(QFX_ERR_OUT_OF_MEMORY,
                             "device arena: cannot reserve " + std::to_string(csz) +
                                 " bytes (" + hipGetErrorString(e) + ")");
      }
      ctx->chunks.push_back({base, csz, sz});
      p = base;
    }
    *dptr = p;
  }
  ctx->live[*dptr] = sz;
  ctx->in_use += sz;
  return QFX_OK;
}

extern "C" int qfx_free(qfx_context* ctx, void* dptr) {
  QFX_CHECK_CTX_NO_FLUSH(ctx);
  if (!dptr) return QFX_OK;
  auto it = ctx->live.find(dptr);
  if (it == ctx->live.end())
    return qfx_set_error(QFX_ERR_INVALID, "qfx_free: pointer not from this arena");
  // a batched copy into the block is launched before the block can be
  // handed out again
  const uintptr_t b0 = reinterpret_cast<uintptr_t>(dptr), b1 = b0 + it->second;
  auto next = ctx->upload_at.upper_bound(b0);
  if ((next != ctx->upload_at.end() && next->first < b1) ||
      (next != ctx->upload_at.begin() && std::prev(next)->second > b0)) {
    const int st = qfx_flush_uploads(ctx);
    if (st != QFX_OK) return st;
  }
  ctx->free_lists[it->second].push_back(dptr);
  ctx->in_use -= it->second;
  ctx->live.erase(it);
  return QFX_OK;
}

extern "C" int qfx_arena_stats(qfx_context* ctx, size_t* reserved, size_t* in_use) {
  QFX_CHECK_CTX_NO_FLUSH(ctx);
  size_t r = 0;
  for (auto& c : ctx->chunks) r += c.size;
  if (reserved) *reserved = r;
  if (in_use) *in_use = ctx->in_use;
  return QFX_OK;
}

int qfx_check_extent(qfx_context* ctx, const void* p, double bytes, const char* what) {
  if (bytes <= 0) return QFX_OK;
  const char* a = static_cast<const char*>(p);
  bool in_arena = false;
  for (const auto& c : ctx->chunks)
    if (a >= c.base && a < c.base + c.size) in_arena = true;
  if (!in_arena) return QFX_OK;  // memory the caller allocated elsewhere
  auto it = ctx->live.upper_bound(const_cast<void*>(p));
  if (it != ctx->live.begin()) {
    --it;
    const char* b0 = static_cast<const char*>(it->first);
    if (a >= b0 && double(a - b0) + bytes <= double(it->second)) return QFX_OK;
  }
  return qfx_set_error(QFX_ERR_INVALID,
                       std::string(what) + ": " + std::to_string(uint64_t(bytes)) +
                           " bytes from the given address are not inside one live block of "
                           "the context's device arena");
}

int qfx_workspace(qfx_context* ctx, size_t bytes, void** out) {
  if (ctx->use_fixed_scratch) {
    // the no-allocation entry point: the fixed scratch array or nothing
    if (bytes > kFixedScratchBytes)
      return qfx_set_error(QFX_ERR_UNSUPPORTED,
                           "qfx_cgemm_noalloc_c64: this product needs " + std::to_string(bytes) +
                               " bytes of device scratch, more than the fixed " +
                               std::to_string(kFixedScratchBytes) + " bytes");
    *out = ctx->fixed_scratch;
    return QFX_OK;
  }
  if (ctx->work_bytes < bytes && ctx->capturing)
    return qfx_set_error(QFX_ERR_INVALID,
                         "device workspace of " + std::to_string(bytes) +
                             " bytes needed while recording a graph; reserve it with "
                             "qfx_graph_begin's workspace_bytes");
  if (ctx->work_bytes < bytes) {
    if (ctx->work) {
      // The stream may still read the old workspace; it is not freed to HIP,
      // only returned to the arena after the stream drains.
      QFX_HIP(hipStreamSynchronize(ctx->stream));
      qfx_free(ctx, ctx->work);
      ctx->work = nullptr;
      ctx->work_bytes = 0;  // a failed allocation below leaves no workspace
    }
    size_t sz = std::max<size_t>(bytes, 1 << 20);
    int st = qfx_alloc(ctx, sz, &ctx->work);
    if (st != QFX_OK) return st;
    ctx->work_bytes = sz;
  }
  *out = ctx->work;
  return QFX_OK;
}

// ---------------------------------------------------------------- profiling

static hipEvent_t pool_event(qfx_context* ctx) {
  if (!ctx->event_pool.empty()) {
    hipEvent_t e = ctx->event_pool.back();
    ctx->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  hipEventCreate(&e);
  return e;
}

QfxLaunchScope::QfxLaunchScope(qfx_context* c, int cls, double bytes, double flops)
    : ctx(c), active(c->profiling && !c->capturing) {
  if (!active) return;
  rec.cls = cls;
  rec.bytes = bytes;
  rec.flops = flops;
  rec.start = pool_event(ctx);
  rec.stop = pool_event(ctx);
  hipEventRecord(rec.start, ctx->stream);
}

QfxLaunchScope::~QfxLaunchScope() {
  if (!active) return;
  hipEventRecord(rec.stop, ctx->stream);
  ctx->pending.push_back(rec);
  // Bound the number of outstanding event pairs.
  if (ctx->pending.size() >= 4096) {
    hipEventSynchronize(ctx->pending.back().stop);
    for (auto& r : ctx->pending) {
      float ms = 0.f;
      hipEventElapsedTime(&ms, r.start, r.stop);
      ctx->launches[r.cls] += 1;
      ctx->ms[r.cls] += ms;
      ctx->bytes[r.cls] += r.bytes;
      ctx->flops[r.cls] += r.flops;
      ctx->event_pool.push_back(r.start);
      ctx->event_pool.push_back(r.stop);
    }
    ctx->pending.clear();
  }
}

extern "C" int qfx_profile_enable(qfx_context* ctx, int on) {
  QFX_CHECK_CTX(ctx);
  ctx->profiling = on != 0;
  return QFX_OK;
}

extern "C" int qfx_profile_reset(qfx_context* ctx) {
  QFX_CHECK_CTX(ctx);
  QFX_NOT_CAPTURING(ctx, "qfx_profile_reset");
  QFX_HIP(hipStreamSynchronize(ctx->stream));
  for (auto& r : ctx->pending) {
    ctx->event_pool.push_back(r.start);
    ctx->event_pool.push_back(r.stop);
  }
  ctx->pending.clear();
  for (int i = 0; i < QFX_K_NUM_CLASSES; ++i) {
    ctx->launches[i] = 0;
    ctx->ms[i] = ctx->bytes[i] = ctx->flops[i] = 0;
  }
  return QFX_OK;
}

extern "C" int qfx_profile_read(qfx_context* ctx, uint64_t* launches, double* ms,
                                double* bytes, double* flops) {
  QFX_CHECK_CTX(ctx);
  QFX_NOT_CAPTURING(ctx, "qfx_profile_read");
  QFX_HIP(hipStreamSynchronize(ctx->stream));
  for (auto& r : ctx->pending) {
    float t = 0.f;
    QFX_HIP(hipEventElapsedTime(&t, r.start, r.stop));
    ctx->launches[r.cls] += 1;
    ctx->ms[r.cls] += t;
    ctx->bytes[r.cls] += r.bytes;
    ctx->flops[r.cls] += r.flops;
    ctx->event_pool.push_back(r.start);
    ctx->event_pool.push_back(r.stop);
  }
  ctx->pending.clear();
  for (int i = 0; i < QFX_K_NUM_CLASSES; ++i) {
    if (launches) launches[i] = ctx->launches[i];
    if (ms) ms[i] = ctx->ms[i];
    if (bytes) bytes[i] = ctx->bytes[i];
    if (flops) flops[i] = ctx->flops[i];
  }
  return QFX_OK;
}

extern "C" int qfx_timer_start(qfx_context* ctx) {
  QFX_CHECK_CTX(ctx);
  QFX_HIP(hipEventRecord(ctx->timer_start, ctx->stream));
  return QFX_OK;
}

extern "C" int qfx_timer_stop(qfx_context* ctx, float* ms) {
  QFX_CHECK_CTX(ctx);
  QFX_HIP(hipEventRecord(ctx->timer_stop, ctx->stream));
  QFX_HIP(hipEventSynchronize(ctx->timer_stop));
  float t = 0.f;
  QFX_HIP(hipEventElapsedTime(&t, ctx->timer_start, ctx->timer_stop));
  if (ms) *ms = t;
  return QFX_OK;
}

// ---------------------------------------------------------------- copies

extern "C" int qfx_copy_h2d(qfx_context* ctx, void* dst, const void* src, size_t bytes) {
  QFX_CHECK_CTX(ctx);
  QFX_NOT_CAPTURING(ctx, "qfx_copy_h2d");
  if (!bytes) return QFX_OK;
  if (!dst || !src) return qfx_set_error(QFX_ERR_INVALID, "qfx_copy_h2d: null pointer");
  QFX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
  QFX_HIP(hipStreamSynchronize(ctx->stream));
  ctx->stage_off = 0;
  return QFX_OK;
}

constexpr size_t kStageBytes = size_t(16) << 20;

// One block per staged copy: the bytes are read from the pinned staging
// area (host memory the device reads over the bus) and written to the
// destination.
__global__ __launch_bounds__(256) void qfx_scatter_uploads_kernel(
    const unsigned char* __restrict__ stage, const qfx_context::Upload* __restrict__ ups) {
  const qfx_context::Upload u = ups[blockIdx.x];
  unsigned char* dst = reinterpret_cast<unsigned char*>(u.dst);
  const unsigned char* src = stage + u.src_off;
  if (((u.dst | u.src_off | u.bytes) & 15) == 0) {
    for (uint64_t i = threadIdx.x; i < u.bytes / 16; i += 256)
      reinterpret_cast<uint4*>(dst)[i] = reinterpret_cast<const uint4*>(src)[i];
  } else if (((u.dst | u.src_off | u.bytes) & 7) == 0) {
    for (uint64_t i = threadIdx.x; i < u.bytes / 8; i += 256)
      reinterpret_cast<uint2*>(dst)[i] = reinterpret_cast<const uint2*>(src)[i];
  } else {
    for (uint64_t i = threadIdx.x; i < u.bytes; i += 256) dst[i] = src[i];
  }
}

// The upload records go into the staging area behind the data; both stay
// untouched until the stream has been synchronized.
int qfx_flush_uploads(qfx_context* ctx) {
  if (ctx->uploads.empty()) return QFX_OK;
  const size_t n = ctx->uploads.size();
  const size_t off = (ctx->stage_off + 255) & ~size_t(255);
  std::memcpy(ctx->stage + off, ctx->uploads.data(), n * sizeof(qfx_context::Upload));
  ctx->stage_off = off + n * sizeof(qfx_context::Upload);
  ctx->uploads.clear();
  ctx->upload_at.clear();
  double bytes = 0;
  const auto* recs = reinterpret_cast<const qfx_context::Upload*>(ctx->stage + off);
  for (size_t i = 0; i < n; ++i) bytes += 2.0 * double(recs[i].bytes);
  QfxLaunchScope scope(ctx, QFX_K_COPY, bytes, 0);
  const unsigned char* stage_dev = static_cast<const unsigned char*>(ctx->stage_dev);
  qfx_scatter_uploads_kernel<<<unsigned(n), 256, 0, ctx->stream>>>(
      stage_dev, reinterpret_cast<const qfx_context::Upload*>(stage_dev + off));
  QFX_HIP(hipGetLastError());
  return QFX_OK;
}

extern "C" int qfx_copy_h2d_async(qfx_context* ctx, void* dst, const void* src, size_t bytes) {
  QFX_CHECK_CTX_NO_FLUSH(ctx);
  QFX_NOT_CAPTURING(ctx, "qfx_copy_h2d_async");
  if (!bytes) return QFX_OK;
  if (!dst || !src) return qfx_set_error(QFX_ERR_INVALID, "qfx_copy_h2d_async: null pointer");
  if (bytes > kStageBytes / 4) return qfx_copy_h2d(ctx, dst, src, bytes);
  if (!ctx->stage) {
    void* p = nullptr;
    QFX_HIP(hipHostMalloc(&p, kStageBytes, hipHostMallocDefault));
    ctx->stage = static_cast<unsigned char*>(p);
    ctx->stage_off = 0;
    QFX_HIP(hipHostGetDevicePointer(&ctx->stage_dev, p, 0));
  }
  // A second copy into bytes a batched copy still has to write goes into
  // the next batch (the copies of one batch run in no particular order).
  const uintptr_t d0 = reinterpret_cast<uintptr_t>(dst), d1 = d0 + bytes;
  auto next = ctx->upload_at.upper_bound(d0);
  const bool overlap = (next != ctx->upload_at.end() && next->first < d1) ||
                       (next != ctx->upload_at.begin() && std::prev(next)->second > d0);
  if (overlap) {
    const int st = qfx_flush_uploads(ctx);
    if (st != QFX_OK) return st;
  }
  size_t off = (ctx->stage_off + 255) & ~size_t(255);
  // room for the data and, behind it, the records of the batch
  const size_t records = (ctx->uploads.size() + 1) * sizeof(qfx_context::Upload) + 256;
  if (off + bytes + records > kStageBytes) {
    // every earlier staged copy has to be done before the area is reused
    const int st = qfx_flush_uploads(ctx);
    if (st != QFX_OK) return st;
    QFX_HIP(hipStreamSynchronize(ctx->stream));
    off = 0;
  }
  std::memcpy(ctx->stage + off, src, bytes);
  ctx->uploads.push_back({uint64_t(d0), uint64_t(off), uint64_t(bytes)});
  ctx->upload_at[d0] = d1;
  ctx->stage_off = off + bytes;
  return QFX_OK;
}

extern "C" int qfx_copy_d2h(qfx_context* ctx, void* dst, const void* src, size_t bytes) {
  QFX_CHECK_CTX(ctx);
  QFX_NOT_CAPTURING(ctx, "qfx_copy_d2h");
  if (!bytes) return QFX_OK;
  if (!dst || !src) return qfx_set_error(QFX_ERR_INVALID, "qfx_copy_d2h: null pointer");
  QFX_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  QFX_HIP(hipStreamSynchronize(ctx->stream));
  ctx->stage_off = 0;
  return QFX_OK;
}

__global__ void qfx_copy_kernel(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                uint64_t n16) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (; i < n16; i += stride) dst[i] = src[i];
}

__global__ void qfx_copy_bytes_kernel(const unsigned char* __restrict__ src,
                                      unsigned char* __restrict__ dst, uint64_t n) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

static unsigned grid_for(uint64_t n, unsigned block) {
  uint64_t g = (n + block - 1) / block;
  if (g > 65536) g = 65536;
  if (g == 0) g = 1;
  return unsigned(g);
}

extern "C" int qfx_copy_d2d(qfx_context* ctx, void* dst, const void* src, size_t bytes) {
  QFX_CHECK_CTX(ctx);
  if (!bytes || dst == src) return QFX_OK;
  if (!dst || !src) return qfx_set_error(QFX_ERR_INVALID, "qfx_copy_d2d: null pointer");
  QFX_EXTENT(ctx, src, double(bytes), "qfx_copy_d2d: src");
  QFX_EXTENT(ctx, dst, double(bytes), "qfx_copy_d2d: dst");
  QfxLaunchScope scope(ctx, QFX_K_COPY, 2.0 * bytes, 0);
  if ((reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src) | bytes) % 16 == 0) {
    uint64_t n16 = bytes / 16;
    qfx_copy_kernel<<<grid_for(n16, 256), 256, 0, ctx->stream>>>(
        static_cast<const uint4*>(src), static_cast<uint4*>(dst), n16);
  } else {
    qfx_copy_bytes_kernel<<<grid_for(bytes, 256), 256, 0, ctx->stream>>>(
        static_cast<const unsigned char*>(src), static_cast<unsigned char*>(dst), bytes);
  }
  QFX_HIP(hipGetLastError());
  return QFX_OK;
}

struct QfxPicks {
  uint32_t v[64];
};

// dst[a][j0 + j][r] = src[a][picks[j]][r] for a < pre, j < npk, r < post
// (elements of 8 bytes); dst has `nall` entries along j.
__global__ void qfx_project_kernel(const float2* __restrict__ src, float2* __restrict__ dst,
                                   uint64_t pre, uint64_t dim, uint64_t post, uint32_t npk,
                                   uint64_t j0, uint64_t nall, QfxPicks picks) {
  const uint64_t total = pre * npk * post;
  uint64_t e = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  const uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (; e < total; e += stride) {
    const uint64_t r = e % post, q = e / post;
    const uint64_t j = q % npk, a = q / npk;
    dst[(a * nall + j0 + j) * post + r] = src[(a * dim + picks.v[j]) * post + r];
  }
}

__global__ void qfx_zero_kernel(unsigned char* __restrict__ dst, uint64_t n) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (; i < n; i += stride) dst[i] = 0;
}

__global__ void qfx_zero16_kernel(uint4* __restrict__ dst, uint64_t n16) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (; i < n16; i += stride) dst[i] = make_uint4(0, 0, 0, 0);
}

extern "C" int qfx_zero(qfx_context* ctx, void* dst, size_t bytes) {
  QFX_CHECK_CTX(ctx);
  if (!bytes) return QFX_OK;
  if (!dst) return qfx_set_error(QFX_ERR_INVALID, "qfx_zero: null pointer");
  QFX_EXTENT(ctx, dst, double(bytes), "qfx_zero: dst");
  QfxLaunchScope scope(ctx, QFX_K_COPY, double(bytes), 0);
  if ((reinterpret_cast<uintptr_t>(dst) | bytes) % 16 == 0)
    qfx_zero16_kernel<<<grid_for(bytes / 16, 256), 256, 0, ctx->stream>>>(
        static_cast<uint4*>(dst), bytes / 16);
  else
    qfx_zero_kernel<<<grid_for(bytes, 256), 256, 0, ctx->stream>>>(
        static_cast<unsigned char*>(dst), bytes);
  QFX_HIP(hipGetLastError());
  return QFX_OK;
}

extern "C" int qfx_synchronize(qfx_context* ctx) {
  QFX_CHECK_CTX(ctx);
  QFX_NOT_CAPTURING(ctx, "qfx_synchronize");
  QFX_HIP(hipStreamSynchronize(ctx->stream));
  ctx->stage_off = 0;
  return QFX_OK;
}

// ---------------------------------------------------------------- accumulate

__global__ void qfx_accumulate_kernel(const float2* __restrict__ src, uint64_t n,
                                      double2* __restrict__ amps) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) {
    float2 v = src[i];
    double2 a = amps[i];
    a.x += double(v.x);
    a.y += double(v.y);
    amps[i] = a;
  }
}

extern "C" int qfx_accumulate_c128(qfx_context* ctx, const qfx_c64* src, uint64_t n,
                                   qfx_c128* amps, uint64_t offset) {
  QFX_CHECK_CTX(ctx);
  if (!n) return QFX_OK;
  if (!src || !amps) return qfx_set_error(QFX_ERR_INVALID, "qfx_accumulate_c128: null pointer");
  QFX_EXTENT(ctx, src, 8.0 * double(n), "qfx_accumulate_c128: src");
  QFX_EXTENT(ctx, amps + offset, 16.0 * double(n), "qfx_accumulate_c128: amps");
  QfxLaunchScope scope(ctx, QFX_K_ACCUMULATE, 8.0 * n + 32.0 * n, 2.0 * n);
  unsigned g = unsigned((n + 255) / 256);
  qfx_accumulate_kernel<<<g, 256, 0, ctx->stream>>>(reinterpret_cast<const float2*>(src), n,
                                                     reinterpret_cast<double2*>(amps) + offset);
  QFX_HIP(hipGetLastError());
  return QFX_OK;
}

// amps[i] += src[b][i] for b = 0, 1, ..., nbatch - 1, one FP64 addition at a
// time in that order (the order of the per-slice accumulations it replaces).
__global__ void qfx_accumulate_batched_kernel(const float2* __restrict__ src, uint64_t n,
                                              uint64_t nbatch, double2* __restrict__ amps) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 a = amps[i];
  for (uint64_t b = 0; b < nbatch; ++b) {
    const float2 v = src[b * n + i];
    a.x += double(v.x);
    a.y += double(v.y);
  }
  amps[i] = a;
}

extern "C" int qfx_accumulate_batched_c128(qfx_context* ctx, const qfx_c64* src, uint64_t n,
                                           uint64_t nbatch, qfx_c128* amps, uint64_t offset) {
  QFX_CHECK_CTX(ctx);
  if (!n || !nbatch) return QFX_OK;
  if (!src || !amps)
    return qfx_set_error(QFX_ERR_INVALID, "qfx_accumulate_batched_c128: null pointer");
  if (n > UINT64_MAX / nbatch || (n + 255) / 256 > 0x7fffffffULL)
    return qfx_set_error(QFX_ERR_INVALID, "qfx_accumulate_batched_c128: size too large");
  QFX_EXTENT(ctx, src, 8.0 * double(n) * double(nbatch), "qfx_accumulate_batched_c128: src");
  QFX_EXTENT(ctx, amps + offset, 16.0 * double(n), "qfx_accumulate_batched_c128: amps");
  QfxLaunchScope scope(ctx, QFX_K_ACCUMULATE, 8.0 * n * nbatch + 32.0 * n, 2.0 * n * nbatch);
  const unsigned g = unsigned((n + 255) / 256);
  qfx_accumulate_batched_kernel<<<g, 256, 0, ctx->stream>>>(
      reinterpret_cast<const float2*>(src), n, nbatch, reinterpret_cast<double2*>(amps) + offset);
  QFX_HIP(hipGetLastError());
  return QFX_OK;
}

// ---------------------------------------------------------------- graphs

extern "C" int qfx_graph_begin(qfx_context* ctx, size_t workspace_bytes) {
  QFX_CHECK_CTX(ctx);
  QFX_NOT_CAPTURING(ctx, "qfx_graph_begin");
  if (workspace_bytes) {
    void* ws = nullptr;
    int st = qfx_workspace(ctx, workspace_bytes, &ws);
    if (st != QFX_OK) return st;
  }
  QFX_HIP(hipStreamSynchronize(ctx->stream));
  QFX_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  ctx->capturing = true;
  return QFX_OK;
}

extern "C" int qfx_graph_end(qfx_context* ctx, qfx_graph** out) {
  QFX_CHECK_CTX(ctx);
  if (!ctx->capturing) return qfx_set_error(QFX_ERR_INVALID, "qfx_graph_end: not recording");
  ctx->capturing = false;
  hipGraph_t graph = nullptr;
  QFX_HIP(hipStreamEndCapture(ctx->stream, &graph));
  if (!out) {
    hipGraphDestroy(graph);
    return qfx_set_error(QFX_ERR_INVALID, "qfx_graph_end: null output");
  }
  hipGraphExec_t exec = nullptr;
  hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e != hipSuccess) {
    hipGraphDestroy(graph);
    return qfx_hip_error(e, "hipGraphInstantiate");
  }
  qfx_graph* g = new qfx_graph;
  g->graph = graph;
  g->exec = exec;
  g->ctx = ctx;
  *out = g;
  return QFX_OK;
}

extern "C" int qfx_graph_launch(qfx_context* ctx, qfx_graph* graph) {
  QFX_CHECK_CTX(ctx);
  QFX_NOT_CAPTURING(ctx, "qfx_graph_launch");
  if (!graph || graph->ctx != ctx)
    return qfx_set_error(QFX_ERR_INVALID, "qfx_graph_launch: graph of another context");
  QFX_HIP(hipGraphLaunch(graph->exec, ctx->stream));
  return QFX_OK;
}

extern "C" int qfx_graph_destroy(qfx_graph* graph) {
  if (!graph) return QFX_OK;
  if (graph->ctx) hipStreamSynchronize(graph->ctx->stream);
  if (graph->exec) hipGraphExecDestroy(graph->exec);
  if (graph->graph) hipGraphDestroy(graph->graph);
  delete graph;
  return QFX_OK;
}

// ---------------------------------------------------------------- projections

extern "C" int qfx_project_c64(qfx_context* ctx, const qfx_c64* src, qfx_c64* dst, uint64_t pre,
                               uint64_t dim, uint64_t post, uint64_t npicks,
                               const uint64_t* picks) {
  QFX_CHECK_CTX(ctx);
  if (pre == 0 || post == 0 || npicks == 0) return QFX_OK;
  if (!src || !dst || !picks)
    return qfx_set_error(QFX_ERR_INVALID, "qfx_project_c64: null pointer");
  for (uint64_t j = 0; j < npicks; ++j)
    if (picks[j] >= dim)
      return qfx_set_error(QFX_ERR_INVALID, "qfx_project_c64: index value " +
                                                std::to_string(picks[j]) + " outside [0, " +
                                                std::to_string(dim) + ")");
  QFX_EXTENT(ctx, src, 8.0 * double(pre) * double(dim) * double(post), "qfx_project_c64: src");
  QFX_EXTENT(ctx, dst, 8.0 * double(pre) * double(npicks) * double(post), "qfx_project_c64: dst");
  QfxLaunchScope scope(ctx, QFX_K_COPY, 16.0 * pre * npicks * post, 0);
  for (uint64_t j0 = 0; j0 < npicks; j0 += 64) {
    const uint32_t npk = uint32_t(std::min<uint64_t>(64, npicks - j0));
    QfxPicks pk{};
    for (uint32_t j = 0; j < npk; ++j) pk.v[j] = uint32_t(picks[j0 + j]);
    qfx_project_kernel<<<grid_for(pre * npk * post, 256), 256, 0, ctx->stream>>>(
        reinterpret_cast<const float2*>(src), reinterpret_cast<float2*>(dst), pre, dim, post, npk,
        j0, npicks, pk);
  }
  QFX_HIP(hipGetLastError());
  return QFX_OK;
}
