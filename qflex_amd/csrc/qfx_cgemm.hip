// Complex64 GEMM / GEMV / dot for the contraction path (gfx950).
//
// The reference hands every pairwise contraction to OpenBLAS
// (cblas_cgemm / cgemv / cdotu_sub, reference src/tensor.cpp:960-1042).
//
// MFMA form.  A complex A[m x k] (interleaved re,im) is read as a real
// matrix Ahat[m x 2k].  B[k x n] is expanded, while staging from LDS, into
// the real matrix Bhat[2k x 2n] with 2x2 blocks [[Br, Bi], [-Bi, Br]], so
// that Ahat . Bhat = [Cr Ci] interleaved, i.e. exactly C in memory order.
// One v_mfma_f32_32x32x2_f32 then consumes one complex k for 32 rows and
// 16 complex columns.  FLOPs are the textbook 8*m*n*k.
//
// Split-BF16 form.  Each real operand value x is split into BF16 terms
// hi = bf16(x), mid = bf16(x - hi), lo = bf16(x - hi - mid); products of the
// terms are accumulated in FP32 by v_mfma_f32_32x32x16_bf16 (8 complex k per
// instruction).  SPLIT_BF16 keeps hi*hi + hi*mid + mid*hi (3 products),
// SPLIT_BF16X3 adds mid*mid + hi*lo + lo*hi (6 products).
#include <algorithm>

#include "qfx_internal.h"

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// ------------------------------------------------------------ scalar path
// One thread per output element, k in increasing order with fmaf: the
// defined reduction order of QFX_PREC_FP32_EXACT.
__global__ void cgemm_scalar_kernel(const float2* __restrict__ A, const float2* __restrict__ B,
                                    float2* __restrict__ C, uint64_t M, uint64_t N,
                                    uint64_t K) {
  uint64_t idx = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (; idx < M * N; idx += stride) {
    uint64_t m = idx / N, n = idx - m * N;
    const float2* a = A + m * K;
    const float2* b = B + n;
    float re = 0.f, im = 0.f;
    for (uint64_t k = 0; k < K; ++k) {
      float2 x = a[k];
      float2 y = b[k * N];
      re = fmaf(x.x, y.x, re);
      re = fmaf(-x.y, y.y, re);
      im = fmaf(x.x, y.y, im);
      im = fmaf(x.y, y.x, im);
    }
    C[idx] = make_float2(re, im);
  }
}

// ------------------------------------------------------------ FP32 MFMA path
constexpr int kBM = 128;  // complex rows per block
constexpr int kBN = 64;   // complex cols per block
constexpr int kBK = 16;   // complex k per stage
constexpr int kAStride = 2 * kBK + 1;  // floats per LDS row of the A tile

__global__ __launch_bounds__(256) void cgemm_f32_mfma_kernel(const float2* __restrict__ A,
                                                             const float2* __restrict__ B,
                                                             float2* __restrict__ C, int M,
                                                             int N, int K) {
  __shared__ float As[2][kBM * kAStride];
  __shared__ __align__(16) float2 Bs[2][kBK * kBN];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1;  // wave row (0..1): 64 rows each
  const int wn = wave & 1;   // wave col (0..1): 32 complex cols each
  const int m0 = blockIdx.x * kBM;
  const int n0 = blockIdx.y * kBN;

  // Global -> register staging assignment.
  // A: 128 rows x 16 k = 2048 complex; thread t loads rows t/16 + 16*h,
  //    k = t%16 (h = 0..7): 16 consecutive threads read 128 B of one row.
  // B: 16 k x 64 n = 1024 complex; thread t loads k = t/64 + 4*h, n = t%64.
  float2 ra[8], rb[4];
  auto load_tile = [&](int kt) {
    const int k0 = kt * kBK;
#pragma unroll
    for (int h = 0; h < 8; ++h) {
      int r = (tid >> 4) + 16 * h;
      int kk = tid & 15;
      int gm = m0 + r, gk = k0 + kk;
      ra[h] = (gm < M && gk < K) ? A[uint64_t(gm) * K + gk] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      int kk = (tid >> 6) + 4 * h;
      int nn = tid & 63;
      int gk = k0 + kk, gn = n0 + nn;
      rb[h] = (gk < K && gn < N) ? B[uint64_t(gk) * N + gn] : make_float2(0.f, 0.f);
    }
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int h = 0; h < 8; ++h) {
      int r = (tid >> 4) + 16 * h;
      int kk = tid & 15;
      As[buf][r * kAStride + 2 * kk] = ra[h].x;
      As[buf][r * kAStride + 2 * kk + 1] = ra[h].y;
    }
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      int kk = (tid >> 6) + 4 * h;
      int nn = tid & 63;
      Bs[buf][kk * kBN + nn] = rb[h];
    }
  };

  // Per-lane operand selectors.
  const int half = lane >> 5;      // MFMA k index (0: real part row, 1: imag)
  const int col = lane & 31;       // real column within a 32-wide MFMA tile
  const int part = col & 1;        // 0: Cr column, 1: Ci column
  const bool take_x = (half == part);
  const float sgn = half ? -1.f : 1.f;

  floatx16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

  const int nkt = (K + kBK - 1) / kBK;
  load_tile(0);
  store_tile(0);
  __syncthreads();
  for (int kt = 0; kt < nkt; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nkt) load_tile(kt + 1);
    const float* as = As[buf];
    const float2* bs = Bs[buf];
#pragma unroll 4
    for (int kk = 0; kk < kBK; ++kk) {
      float a0 = as[(wm * 64 + col) * kAStride + 2 * kk + half];
      float a1 = as[(wm * 64 + 32 + col) * kAStride + 2 * kk + half];
      float2 b0v = bs[kk * kBN + wn * 32 + (col >> 1)];
      float2 b1v = bs[kk * kBN + wn * 32 + 16 + (col >> 1)];
      float b0 = take_x ? b0v.x : sgn * b0v.y;
      float b1 = take_x ? b1v.x : sgn * b1v.y;
      acc[0][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b0, acc[0][0], 0, 0, 0);
      acc[0][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, b1, acc[0][1], 0, 0, 0);
      acc[1][0] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b0, acc[1][0], 0, 0, 0);
      acc[1][1] = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, b1, acc[1][1], 0, 0, 0);
    }
    if (kt + 1 < nkt) store_tile(buf ^ 1);
    __syncthreads();
  }

  // Output: register r of lane l holds row 8*(r/4) + 4*(l/32) + r%4,
  // real column l%32 of its 32x32 tile.
  float* Cf = reinterpret_cast<float*>(C);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        int row = m0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
        int cn = n0 + wn * 32 + j * 16 + (col >> 1);
        if (row < M && cn < N) Cf[(uint64_t(row) * N + cn) * 2 + part] = acc[i][j][r];
      }
}

// ------------------------------------------------------------ split-BF16 MFMA
// Same block tile; LDS holds FP32 (A rows interleaved, B complex) and the
// split happens when operands are read: v_mfma_f32_32x32x16_bf16 consumes 8
// complex k per instruction.
__device__ __forceinline__ void split3(float x, __bf16& h, __bf16& m, __bf16& l) {
  h = static_cast<__bf16>(x);
  float r = x - static_cast<float>(h);
  m = static_cast<__bf16>(r);
  l = static_cast<__bf16>(r - static_cast<float>(m));
}

constexpr int kAStrideB = 2 * kBK + 4;  // 16-byte aligned rows for b128 reads

template <int TERMS>
__global__ __launch_bounds__(256) void cgemm_bf16_mfma_kernel(const float2* __restrict__ A,
                                                              const float2* __restrict__ B,
                                                              float2* __restrict__ C, int M,
                                                              int N, int K) {
  __shared__ __align__(16) float As[2][kBM * kAStrideB];
  __shared__ __align__(16) float2 Bs[2][kBK * kBN];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1;
  const int wn = wave & 1;
  const int m0 = blockIdx.x * kBM;
  const int n0 = blockIdx.y * kBN;

  float2 ra[8], rb[4];
  auto load_tile = [&](int kt) {
    const int k0 = kt * kBK;
#pragma unroll
    for (int h = 0; h < 8; ++h) {
      int r = (tid >> 4) + 16 * h;
      int kk = tid & 15;
      int gm = m0 + r, gk = k0 + kk;
      ra[h] = (gm < M && gk < K) ? A[uint64_t(gm) * K + gk] : make_float2(0.f, 0.f);
    }
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      int kk = (tid >> 6) + 4 * h;
      int nn = tid & 63;
      int gk = k0 + kk, gn = n0 + nn;
      rb[h] = (gk < K && gn < N) ? B[uint64_t(gk) * N + gn] : make_float2(0.f, 0.f);
    }
  };
  auto store_tile = [&](int buf) {
#pragma unroll
    for (int h = 0; h < 8; ++h) {
      int r = (tid >> 4) + 16 * h;
      int kk = tid & 15;
      *reinterpret_cast<float2*>(&As[buf][r * kAStrideB + 2 * kk]) = ra[h];
    }
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      int kk = (tid >> 6) + 4 * h;
      int nn = tid & 63;
      Bs[buf][kk * kBN + nn] = rb[h];
    }
  };

  const int half = lane >> 5;  // lanes 0-31: complex k 0..3 of the 8; 32-63: k 4..7
  const int col = lane & 31;
  const int part = col & 1;

  floatx16 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

  const int nkt = (K + kBK - 1) / kBK;
  load_tile(0);
  store_tile(0);
  __syncthreads();
  for (int kt = 0; kt < nkt; ++kt) {
    const int buf = kt & 1;
    if (kt + 1 < nkt) load_tile(kt + 1);
    const float* as = As[buf];
    const float2* bs = Bs[buf];
#pragma unroll
    for (int ks = 0; ks < kBK; ks += 8) {
      // A operand: row (lane%32), reals 8*half..8*half+7 of this 8-k step,
      // i.e. complex k = ks + 4*half .. +3 interleaved (re, im).
      bf16x8 ah[2], am[2], al[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) {
        const float* ap = as + (wm * 64 + i * 32 + col) * kAStrideB + 2 * (ks + 4 * half);
        float4 v0 = *reinterpret_cast<const float4*>(ap);
        float4 v1 = *reinterpret_cast<const float4*>(ap + 4);
        float vals[8] = {v0.x, v0.y, v0.z, v0.w, v1.x, v1.y, v1.z, v1.w};
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          __bf16 h, m, l;
          split3(vals[e], h, m, l);
          ah[i][e] = h;
          am[i][e] = m;
          al[i][e] = l;
        }
      }
      // B operand: Bhat rows 8*half..8*half+7 (complex k = ks+4*half+q,
      // q = 0..3, re row then im row), real column col of the 32-wide tile.
      bf16x8 bh[2], bm[2], bl[2];
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int nn = wn * 32 + j * 16 + (col >> 1);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          float2 b = bs[(ks + 4 * half + q) * kBN + nn];
          // re row: (Br, Bi) for (even, odd) columns; im row: (-Bi, Br)
          float v_re = part ? b.y : b.x;
          float v_im = part ? b.x : -b.y;
          __bf16 h, m, l;
          split3(v_re, h, m, l);
          bh[j][2 * q] = h;
          bm[j][2 * q] = m;
          bl[j][2 * q] = l;
          split3(v_im, h, m, l);
          bh[j][2 * q + 1] = h;
          bm[j][2 * q + 1] = m;
          bl[j][2 * q + 1] = l;
        }
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if (TERMS >= 6) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am[i], bm[j], acc[i][j], 0, 0, 0);
          }
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(am[i], bh[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bm[j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
        }
    }
    if (kt + 1 < nkt) store_tile(buf ^ 1);
    __syncthreads();
  }

  float* Cf = reinterpret_cast<float*>(C);
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        int row = m0 + wm * 64 + i * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
        int cn = n0 + wn * 32 + j * 16 + (col >> 1);
        if (row < M && cn < N) Cf[(uint64_t(row) * N + cn) * 2 + part] = acc[i][j][r];
      }
}

// ------------------------------------------------------------ GEMV / dot
__device__ __forceinline__ float2 cfma(float2 a, float2 b, float2 c) {
  c.x = fmaf(a.x, b.x, c.x);
  c.x = fmaf(-a.y, b.y, c.x);
  c.y = fmaf(a.x, b.y, c.y);
  c.y = fmaf(a.y, b.x, c.y);
  return c;
}

__device__ __forceinline__ float2 wave_sum(float2 v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    v.x += __shfl_xor(v.x, off);
    v.y += __shfl_xor(v.y, off);
  }
  return v;
}

// y[m] = A[m x k] . x[k]; one wave per row, fixed lane-strided order.
__global__ __launch_bounds__(256) void gemv_n_kernel(const float2* __restrict__ A,
                                                     const float2* __restrict__ x,
                                                     float2* __restrict__ y, uint64_t M,
                                                     uint64_t K) {
  const int lane = threadIdx.x & 63;
  uint64_t row = uint64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  uint64_t rstride = uint64_t(gridDim.x) * 4;
  for (; row < M; row += rstride) {
    const float2* a = A + row * K;
    float2 acc = make_float2(0.f, 0.f);
    for (uint64_t k = lane; k < K; k += 64) acc = cfma(a[k], x[k], acc);
    acc = wave_sum(acc);
    if (lane == 0) y[row] = acc;
  }
}

// Partial y over a chunk of rows: part[s][c] = sum_{r in chunk s} x[r] A[r][c].
__global__ __launch_bounds__(256) void gemv_t_partial_kernel(const float2* __restrict__ A,
                                                             const float2* __restrict__ x,
                                                             float2* __restrict__ part,
                                                             uint64_t M, uint64_t K,
                                                             uint64_t rows_per_split) {
  uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint64_t s = blockIdx.y;
  if (c >= K) return;
  uint64_t r0 = s * rows_per_split;
  uint64_t r1 = r0 + rows_per_split < M ? r0 + rows_per_split : M;
  float2 acc = make_float2(0.f, 0.f);
  for (uint64_t r = r0; r < r1; ++r) acc = cfma(x[r], A[r * K + c], acc);
  part[s * K + c] = acc;
}

__global__ void sum_partials_kernel(const float2* __restrict__ part, float2* __restrict__ y,
                                    uint64_t nsplit, uint64_t K) {
  uint64_t c = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (c >= K) return;
  float2 acc = make_float2(0.f, 0.f);
  for (uint64_t s = 0; s < nsplit; ++s) {
    float2 v = part[s * K + c];
    acc.x += v.x;
    acc.y += v.y;
  }
  y[c] = acc;
}

// Block partial sums of x[i]*y[i] in a fixed per-lane order, then a tree.
__global__ __launch_bounds__(256) void dot_partial_kernel(const float2* __restrict__ x,
                                                          const float2* __restrict__ y,
                                                          float2* __restrict__ part,
                                                          uint64_t n) {
  __shared__ float2 red[4];
  float2 acc = make_float2(0.f, 0.f);
  for (uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += uint64_t(gridDim.x) * blockDim.x)
    acc = cfma(x[i], y[i], acc);
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    float2 t = red[0];
    for (int w = 1; w < 4; ++w) {
      t.x += red[w].x;
      t.y += red[w].y;
    }
    part[blockIdx.x] = t;
  }
}

unsigned cap_grid(uint64_t g) { return unsigned(std::max<uint64_t>(1, std::min<uint64_t>(g, 65535))); }

}  // namespace

static int launch_gemv_t(qfx_context* ctx, uint64_t m, uint64_t k, const float2* A,
                         const float2* x, float2* y) {
  // Split the m rows so that about 2048 blocks of 256 columns are in flight.
  uint64_t col_blocks = (k + 255) / 256;
  uint64_t nsplit = std::max<uint64_t>(1, std::min<uint64_t>(m, 2048 / col_blocks));
  nsplit = std::min<uint64_t>(nsplit, 65535);
  uint64_t rows_per = (m + nsplit - 1) / nsplit;
  nsplit = (m + rows_per - 1) / rows_per;
  void* ws = nullptr;
  int st = qfx_workspace(ctx, nsplit * k * sizeof(float2), &ws);
  if (st != QFX_OK) return st;
  float2* part = static_cast<float2*>(ws);
  dim3 g{unsigned(col_blocks), unsigned(nsplit), 1};
  gemv_t_partial_kernel<<<g, 256, 0, ctx->stream>>>(A, x, part, m, k, rows_per);
  sum_partials_kernel<<<cap_grid((k + 255) / 256), 256, 0, ctx->stream>>>(part, y, nsplit, k);
  QFX_HIP(hipGetLastError());
  return QFX_OK;
}

extern "C" int qfx_cgemv_c64(qfx_context* ctx, int trans, uint64_t m, uint64_t k,
                             const qfx_c64* A, const qfx_c64* x, qfx_c64* y) {
  QFX_CHECK_CTX(ctx);
  uint64_t ny = trans ? k : m;
  if (ny == 0) return QFX_OK;
  if (!A || !x || !y) return qfx_set_error(QFX_ERR_INVALID, "qfx_cgemv_c64: null pointer");
  if ((trans ? m : k) == 0) return qfx_zero(ctx, y, ny * sizeof(qfx_c64));
  QfxLaunchScope scope(ctx, QFX_K_GEMV, 8.0 * (m * k + m + k), 8.0 * m * k);
  const float2* A2 = reinterpret_cast<const float2*>(A);
  const float2* x2 = reinterpret_cast<const float2*>(x);
  float2* y2 = reinterpret_cast<float2*>(y);
  if (!trans) {
    gemv_n_kernel<<<cap_grid((m + 3) / 4), 256, 0, ctx->stream>>>(A2, x2, y2, m, k);
    QFX_HIP(hipGetLastError());
    return QFX_OK;
  }
  return launch_gemv_t(ctx, m, k, A2, x2, y2);
}

extern "C" int qfx_cdotu_c64(qfx_context* ctx, uint64_t n, const qfx_c64* x, const qfx_c64* y,
                             qfx_c64* out) {
  QFX_CHECK_CTX(ctx);
  if (!out) return qfx_set_error(QFX_ERR_INVALID, "qfx_cdotu_c64: null output");
  if (n == 0) return qfx_zero(ctx, out, sizeof(qfx_c64));
  if (!x || !y) return qfx_set_error(QFX_ERR_INVALID, "qfx_cdotu_c64: null pointer");
  QfxLaunchScope scope(ctx, QFX_K_DOT, 16.0 * n, 8.0 * n);
  uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 1024);
  void* ws = nullptr;
  int st = qfx_workspace(ctx, blocks * sizeof(float2), &ws);
  if (st != QFX_OK) return st;
  float2* part = static_cast<float2*>(ws);
  dot_partial_kernel<<<unsigned(blocks), 256, 0, ctx->stream>>>(
      reinterpret_cast<const float2*>(x), reinterpret_cast<const float2*>(y), part, n);
  sum_partials_kernel<<<1, 256, 0, ctx->stream>>>(part, reinterpret_cast<float2*>(out), blocks,
                                                  1);
  QFX_HIP(hipGetLastError());
  return QFX_OK;
}

extern "C" int qfx_cgemm_c64(qfx_context* ctx, int precision, uint64_t m, uint64_t n,
                             uint64_t k, const qfx_c64* A, const qfx_c64* B, qfx_c64* C) {
  QFX_CHECK_CTX(ctx);
  if (precision < QFX_PREC_FP32 || precision > QFX_PREC_SPLIT_BF16X3)
    return qfx_set_error(QFX_ERR_INVALID, "qfx_cgemm_c64: unknown precision");
  if (m == 0 || n == 0) return QFX_OK;
  if (!C) return qfx_set_error(QFX_ERR_INVALID, "qfx_cgemm_c64: null C");
  if (k == 0) return qfx_zero(ctx, C, m * n * sizeof(qfx_c64));
  if (!A || !B) return qfx_set_error(QFX_ERR_INVALID, "qfx_cgemm_c64: null A or B");
  const float2* A2 = reinterpret_cast<const float2*>(A);
  const float2* B2 = reinterpret_cast<const float2*>(B);
  float2* C2 = reinterpret_cast<float2*>(C);
  const double flops = 8.0 * double(m) * double(n) * double(k);
  const double bytes = 8.0 * (double(m) * k + double(k) * n + double(m) * n);

  const bool big = m >= 64 && n >= 16 && k >= 8 && m <= 0x7fffffffULL &&
                   n <= 0x7fffffffULL && k <= 0x7fffffffULL;
  if (precision == QFX_PREC_FP32_EXACT || !big) {
    if (n == 1 && precision != QFX_PREC_FP32_EXACT) {
      QfxLaunchScope scope(ctx, QFX_K_GEMV, bytes, flops);
      gemv_n_kernel<<<cap_grid((m + 3) / 4), 256, 0, ctx->stream>>>(A2, B2, C2, m, k);
      QFX_HIP(hipGetLastError());
      return QFX_OK;
    }
    if (m == 1 && precision != QFX_PREC_FP32_EXACT) {
      QfxLaunchScope scope(ctx, QFX_K_GEMV, bytes, flops);
      // y[n] = sum_k a[k] B[k][n]: transposed GEMV of B (k x n) with x = A.
      return launch_gemv_t(ctx, k, n, B2, A2, C2);
    }
    QfxLaunchScope scope(ctx, QFX_K_CGEMM, bytes, flops);
    cgemm_scalar_kernel<<<cap_grid((m * n + 255) / 256), 256, 0, ctx->stream>>>(A2, B2, C2, m, n,
                                                                               k);
    QFX_HIP(hipGetLastError());
    return QFX_OK;
  }
  QfxLaunchScope scope(ctx, QFX_K_CGEMM, bytes, flops);
  dim3 grid(unsigned((m + kBM - 1) / kBM), unsigned((n + kBN - 1) / kBN));
  if (grid.y > 65535) return qfx_set_error(QFX_ERR_UNSUPPORTED, "qfx_cgemm_c64: n too large");
  if (precision == QFX_PREC_FP32)
    cgemm_f32_mfma_kernel<<<grid, 256, 0, ctx->stream>>>(A2, B2, C2, int(m), int(n), int(k));
  else if (precision == QFX_PREC_SPLIT_BF16)
    cgemm_bf16_mfma_kernel<3><<<grid, 256, 0, ctx->stream>>>(A2, B2, C2, int(m), int(n), int(k));
  else
    cgemm_bf16_mfma_kernel<6><<<grid, 256, 0, ctx->stream>>>(A2, B2, C2, int(m), int(n), int(k));
  QFX_HIP(hipGetLastError());
  return QFX_OK;
}
