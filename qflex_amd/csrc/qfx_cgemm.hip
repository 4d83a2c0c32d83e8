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
#include <type_traits>

#include "qfx_internal.h"

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
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
// Same block tile as the FP32 kernel.  Operands are split once, while they
// are staged into LDS, into NT BF16 planes (hi, mid[, lo]):
//   As[t][row][32 reals]          A rows, interleaved (re, im) along k;
//   Bs[t][part*64 + n][32 reals]  Bhat columns: part 0 (the Cr column of
//                                 complex column n) holds (Br_k, -Bi_k),
//                                 part 1 (Ci) holds (Bi_k, Br_k).
// v_mfma_f32_32x32x16_bf16 then reads 8 consecutive BF16 of a row per lane
// (one ds_read_b128) and consumes 8 complex k per instruction.
__device__ __forceinline__ void split3(float x, __bf16& h, __bf16& m, __bf16& l) {
  h = static_cast<__bf16>(x);
  float r = x - static_cast<float>(h);
  m = static_cast<__bf16>(r);
  l = static_cast<__bf16>(r - static_cast<float>(m));
}

constexpr int kRowB = 40;  // BF16 per LDS row: 32 + 8 pad (80 B, 16-byte aligned)

template <int TERMS>
__global__ __launch_bounds__(256) void cgemm_bf16_mfma_kernel(const float2* __restrict__ A,
                                                              const float2* __restrict__ B,
                                                              float2* __restrict__ C, int M,
                                                              int N, int K) {
  constexpr int NT = TERMS >= 6 ? 3 : 2;
  __shared__ __align__(16) __bf16 As[NT][kBM * kRowB];
  __shared__ __align__(16) __bf16 Bs[NT][2 * kBN * kRowB];

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
  typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
  auto store_tile = [&]() {
#pragma unroll
    for (int h = 0; h < 8; ++h) {
      int r = (tid >> 4) + 16 * h;
      int kk = tid & 15;
      __bf16 xh, xm, xl, yh, ym, yl;
      split3(ra[h].x, xh, xm, xl);
      split3(ra[h].y, yh, ym, yl);
      __bf16 th[3][2] = {{xh, yh}, {xm, ym}, {xl, yl}};
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        bf16x2 v = {th[t][0], th[t][1]};
        *reinterpret_cast<bf16x2*>(&As[t][r * kRowB + 2 * kk]) = v;
      }
    }
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      int kk = (tid >> 6) + 4 * h;
      int nn = tid & 63;
      __bf16 rh, rm, rl, ih, im, il;
      split3(rb[h].x, rh, rm, rl);
      split3(rb[h].y, ih, im, il);
      __bf16 tr[3] = {rh, rm, rl}, ti[3] = {ih, im, il};
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        bf16x2 c0 = {tr[t], static_cast<__bf16>(-static_cast<float>(ti[t]))};
        bf16x2 c1 = {ti[t], tr[t]};
        *reinterpret_cast<bf16x2*>(&Bs[t][nn * kRowB + 2 * kk]) = c0;
        *reinterpret_cast<bf16x2*>(&Bs[t][(kBN + nn) * kRowB + 2 * kk]) = c1;
      }
    }
  };

  const int half = lane >> 5;
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
  for (int kt = 0; kt < nkt; ++kt) {
    __syncthreads();
    store_tile();
    __syncthreads();
    if (kt + 1 < nkt) load_tile(kt + 1);
#pragma unroll
    for (int ks = 0; ks < kBK; ks += 8) {
      const int off = 2 * ks + 8 * half;
      bf16x8 a[NT][2], b[NT][2];
#pragma unroll
      for (int t = 0; t < NT; ++t) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
          a[t][i] = *reinterpret_cast<const bf16x8*>(&As[t][(wm * 64 + i * 32 + col) * kRowB + off]);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int rc = wn * 64 + j * 32 + col;  // real column of Bhat in this block
          const int row = (rc & 1) * kBN + (rc >> 1);
          b[t][j] = *reinterpret_cast<const bf16x8*>(&Bs[t][row * kRowB + off]);
        }
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          if (TERMS >= 6) {
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[2][i], b[0][j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0][i], b[2][j], acc[i][j], 0, 0, 0);
            acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1][i], b[1][j], acc[i][j], 0, 0, 0);
          }
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[1][i], b[0][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0][i], b[1][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[0][i], b[0][j], acc[i][j], 0, 0, 0);
        }
    }
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

// ------------------------------------------------------------ split-FP16
// Partial maxima of max(|re|, |im|) over the elements of A (blocks
// [0, blocks_a)) and of B (the remaining blocks), one float per block, in a
// single launch; the GEMM blocks take the maximum of the partials.  The
// result does not depend on the order of the reduction.
__device__ __forceinline__ float absmax4(float4 v) {
  return fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w)));
}

__global__ __launch_bounds__(256) void amax_partial_kernel(const float2* __restrict__ A,
                                                           uint64_t na,
                                                           const float2* __restrict__ B,
                                                           uint64_t nb, int blocks_a,
                                                           float* __restrict__ part) {
  __shared__ float red[4];
  const bool is_a = int(blockIdx.x) < blocks_a;
  const float2* x = is_a ? A : B;
  const uint64_t n = is_a ? na : nb;
  const uint64_t nblk = is_a ? uint64_t(blocks_a) : uint64_t(gridDim.x - blocks_a);
  const uint64_t b = is_a ? blockIdx.x : blockIdx.x - blocks_a;
  const uint64_t stride = nblk * 256;
  float m = 0.f;
  if ((reinterpret_cast<uintptr_t>(x) & 15) == 0) {
    // pairs of complex values as float4, four loads in flight per thread
    const float4* x4 = reinterpret_cast<const float4*>(x);
    const uint64_t n4 = n / 2;
    uint64_t i = b * 256 + threadIdx.x;
    for (; i + 3 * stride < n4; i += 4 * stride) {
      const float4 v0 = x4[i], v1 = x4[i + stride], v2 = x4[i + 2 * stride],
                   v3 = x4[i + 3 * stride];
      m = fmaxf(m, fmaxf(fmaxf(absmax4(v0), absmax4(v1)), fmaxf(absmax4(v2), absmax4(v3))));
    }
    for (; i < n4; i += stride) m = fmaxf(m, absmax4(x4[i]));
    if ((n & 1) && b == 0 && threadIdx.x == 0) {
      const float2 v = x[n - 1];
      m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
    }
  } else {
    for (uint64_t i = b * 256 + threadIdx.x; i < n; i += stride) {
      const float2 v = x[i];
      m = fmaxf(m, fmaxf(fabsf(v.x), fabsf(v.y)));
    }
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

// Partial maxima of the operands of one split-FP16 GEMM (from
// amax_partial_kernel, or written by the GEMM that produced the operand).
struct AmaxRef {
  const float* a;  // na partial maxima of |A|
  int na;
  const float* b;  // nb partial maxima of |B|
  int nb;
};

// Power of two s with max|x| * s in [2^14, 2^15) (FP16 max is 65504).
__device__ __forceinline__ float fp16_scale(float amax) {
  if (!(amax > 0.f)) return 1.f;
  int e;
  frexpf(amax, &e);  // amax = f * 2^e, f in [0.5, 1)
  return ldexpf(1.f, 15 - e);
}

// Operand scales from the partial maxima, for every thread of a block of
// THREADS threads; red holds 2 * THREADS / 64 floats of LDS that no other
// thread touches until the block's next barrier.
template <int THREADS>
__device__ __forceinline__ void operand_scales(const AmaxRef& r, float* red, float& sa,
                                               float& sb) {
  constexpr int W = THREADS / 64;
  float ma = 0.f, mb = 0.f;
  for (int i = threadIdx.x; i < r.na; i += THREADS) ma = fmaxf(ma, r.a[i]);
  for (int i = threadIdx.x; i < r.nb; i += THREADS) mb = fmaxf(mb, r.b[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    ma = fmaxf(ma, __shfl_xor(ma, off));
    mb = fmaxf(mb, __shfl_xor(mb, off));
  }
  if ((threadIdx.x & 63) == 0) {
    red[threadIdx.x >> 6] = ma;
    red[W + (threadIdx.x >> 6)] = mb;
  }
  __syncthreads();
  ma = red[0];
  mb = red[W];
#pragma unroll
  for (int w = 1; w < W; ++w) {
    ma = fmaxf(ma, red[w]);
    mb = fmaxf(mb, red[W + w]);
  }
  sa = fp16_scale(ma);
  sb = fp16_scale(mb);
}

// x = hi + lo with hi = fp16(x) and lo = fp16(x - hi): 22 significant bits
// while the residual is an FP16 normal, i.e. for |x| >= 2^-3 after scaling
// (elements down to 2^-18 of the operand's max); smaller elements keep
// fewer bits, which is below 2^-40 of the max in absolute terms.
__device__ __forceinline__ void split_fp16(float x, _Float16& h, _Float16& l) {
  h = static_cast<_Float16>(x);
  l = static_cast<_Float16>(x - static_cast<float>(h));
}

typedef _Float16 halfx8 __attribute__((ext_vector_type(8)));
typedef _Float16 halfx2 __attribute__((ext_vector_type(2)));

// Same tiling as cgemm_bf16_mfma_kernel; two FP16 planes (hi, lo) of the
// scaled operands, products hi*hi + hi*lo + lo*hi in one FP32 accumulator,
// unscaled on output.
__global__ __launch_bounds__(256) void cgemm_fp16x2_mfma_kernel(
    const float2* __restrict__ A, const float2* __restrict__ B, float2* __restrict__ C, int M,
    int N, int K, AmaxRef amax) {
  __shared__ __align__(16) _Float16 As[2][kBM * kRowB];
  __shared__ __align__(16) _Float16 Bs[2][2 * kBN * kRowB];
  __shared__ float red[8];

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave >> 1;
  const int wn = wave & 1;
  const int m0 = blockIdx.x * kBM;
  const int n0 = blockIdx.y * kBN;
  float sa, sb;
  operand_scales<256>(amax, red, sa, sb);
  const float unscale = 1.f / (sa * sb);

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
  auto store_tile = [&]() {
#pragma unroll
    for (int h = 0; h < 8; ++h) {
      int r = (tid >> 4) + 16 * h;
      int kk = tid & 15;
      _Float16 xh, xl, yh, yl;
      split_fp16(ra[h].x * sa, xh, xl);
      split_fp16(ra[h].y * sa, yh, yl);
      *reinterpret_cast<halfx2*>(&As[0][r * kRowB + 2 * kk]) = halfx2{xh, yh};
      *reinterpret_cast<halfx2*>(&As[1][r * kRowB + 2 * kk]) = halfx2{xl, yl};
    }
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      int kk = (tid >> 6) + 4 * h;
      int nn = tid & 63;
      _Float16 rh, rl, ih, il;
      split_fp16(rb[h].x * sb, rh, rl);
      split_fp16(rb[h].y * sb, ih, il);
      *reinterpret_cast<halfx2*>(&Bs[0][nn * kRowB + 2 * kk]) = halfx2{rh, -ih};
      *reinterpret_cast<halfx2*>(&Bs[0][(kBN + nn) * kRowB + 2 * kk]) = halfx2{ih, rh};
      *reinterpret_cast<halfx2*>(&Bs[1][nn * kRowB + 2 * kk]) = halfx2{rl, -il};
      *reinterpret_cast<halfx2*>(&Bs[1][(kBN + nn) * kRowB + 2 * kk]) = halfx2{il, rl};
    }
  };

  const int half = lane >> 5;
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
  for (int kt = 0; kt < nkt; ++kt) {
    __syncthreads();
    store_tile();
    __syncthreads();
    if (kt + 1 < nkt) load_tile(kt + 1);
#pragma unroll
    for (int ks = 0; ks < kBK; ks += 8) {
      const int off = 2 * ks + 8 * half;
      halfx8 a[2][2], b[2][2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
          a[t][i] = *reinterpret_cast<const halfx8*>(&As[t][(wm * 64 + i * 32 + col) * kRowB + off]);
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int rc = wn * 64 + j * 32 + col;
          const int row = (rc & 1) * kBN + (rc >> 1);
          b[t][j] = *reinterpret_cast<const halfx8*>(&Bs[t][row * kRowB + off]);
        }
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[1][i], b[0][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[0][i], b[1][j], acc[i][j], 0, 0, 0);
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(a[0][i], b[0][j], acc[i][j], 0, 0, 0);
        }
    }
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
        if (row < M && cn < N)
          Cf[(uint64_t(row) * N + cn) * 2 + part] = acc[i][j][r] * unscale;
      }
}

// ------------------------------------------------------------ split forms, large tiles
// For large m and n >= 128 (k, n even).  512 threads = 8 waves as WM x WN
// (4 x 2 or 2 x 4); each wave owns 64 rows x 64 complex columns = 2 x 4 MFMA
// tiles of 32 x 32 (128 accumulator registers), so the block tile is
// 64*WM complex rows x 64*WN complex columns.  One stage = 8 complex k = 16
// reals = one 32x32x16 MFMA k-step.  LDS holds two stages of NT planes:
//   As[buf][t][row][16 reals + 8 pad]      (48-byte rows: conflict-free b128)
//   Bs[buf][t][2*n + part][16 + 8 pad]     Bhat columns in Bhat order (the two
//                                          columns of complex n adjacent, so
//                                          that a fragment read is conflict-free)
// Software pipeline, one barrier per stage: while the MFMAs of stage s run
// from LDS buffer s&1, the raw operands of stage s+1 (loaded into registers
// during stage s-1) are split and stored into the other buffer, and the
// global loads of stage s+2 are in flight into the register set just freed.
// The loop body is unrolled by two (static register sets) and branch-free
// (the last stages reload the final tile into a buffer that is not read
// again), so the scheduler can interleave MFMA, VALU and LDS instructions;
// sched_group_barrier fixes that interleave.
// MODE 0: BF16 hi+mid+lo, 6 products; 1: BF16 hi+mid, 3 products;
// 2: FP16 hi+lo of operands scaled by powers of two (amax), 3 products.
constexpr int kBBK = 8;
constexpr int kRowL = 24;  // 16-bit elements per LDS row

template <int MODE>
struct SplitTraits;
template <>
struct SplitTraits<0> {
  typedef __bf16 T;
  typedef bf16x8 V;
  static constexpr int NT = 3;
  static constexpr int PRODUCTS = 6;
};
template <>
struct SplitTraits<1> {
  typedef __bf16 T;
  typedef bf16x8 V;
  static constexpr int NT = 2;
  static constexpr int PRODUCTS = 3;
};
template <>
struct SplitTraits<2> {
  typedef _Float16 T;
  typedef halfx8 V;
  static constexpr int NT = 2;
  static constexpr int PRODUCTS = 3;
};

template <int MODE>
__device__ __forceinline__ void split_terms(float x, typename SplitTraits<MODE>::T* t) {
  if constexpr (MODE == 2) {
    split_fp16(x, t[0], t[1]);
  } else {
    __bf16 h, m, l;
    split3(x, h, m, l);
    t[0] = h;
    t[1] = m;
    if constexpr (SplitTraits<MODE>::NT == 3) t[2] = l;
  }
}

template <int MODE>
__device__ __forceinline__ floatx16 mfma16(typename SplitTraits<MODE>::V a,
                                           typename SplitTraits<MODE>::V b, floatx16 c) {
  if constexpr (MODE == 2)
    return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0);
  else
    return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0);
}

// Operand addressing of the large-tile kernel.
//   AM 0: A row-major m x k, B row-major k x n.
//   AM 1: A and B through digit maps; A read along k (a_cols digit 0 at
//         stride 1), B along n.
//   AM 2: as 1, but A read along m (a_rows digit 0 at stride 1): each lane
//         loads rows m, m+1 at one k and stores them as two LDS rows.
// J = MFMA column tiles per wave (4: 64 complex columns, 2: 32).
// One output tile (rows m0.., complex columns n0..) of the product; the
// block's 512 threads all call it.  cmax_slot: where the tile's max |C|
// goes (or null).
// CHAIN: the block computes the tiles L0, L0 + Lstride, ... below Lend
// (tile L at row tile L / gy, column tile L % gy) one after the other; the
// last two stages of a tile load the first two stages of the next one, and
// its last stage splits the next tile's first stage into LDS, so a tile's
// prologue overlaps the previous tile's last stages and C write-back.
// cmax_slot then points at the per-tile maxima (slot L).  Needs an even
// number of k-stages and whole-operand addressing (not WIDE).
template <int MODE, int WM, int WN, int J, int AM, int I = 2, bool WIDE = false,
          bool CHAIN = false>
__device__ __forceinline__ void split_big_tile(
    const float2* __restrict__ A, const float2* __restrict__ B, float2* __restrict__ C, int M,
    int N, int K, AmaxRef amax, const qfx_index_map& a_rows, const qfx_index_map& a_cols,
    const qfx_index_map& b_rows, const qfx_index_map& b_cols, float* __restrict__ cmax_slot,
    const int m0_first, const int n0_first, const unsigned L0 = 0, const unsigned Lstride = 0,
    const unsigned Lend = 0, const unsigned gy = 1) {
  static_assert(!(CHAIN && WIDE), "chained tiles use whole-operand addressing");
  constexpr int NTH = 64 * WM * WN;  // threads of the block
  static_assert(NTH == 512 || NTH == 256, "8 or 4 waves");
  typedef typename SplitTraits<MODE>::T T;
  typedef typename SplitTraits<MODE>::V V;
  constexpr int NT = SplitTraits<MODE>::NT;
  constexpr int kBM = 32 * I * WM, kBN = 16 * J * WN;
  constexpr int kAPlane = kBM * kRowL;
  constexpr int kBPlane = 2 * kBN * kRowL;
  constexpr int kBuf = NT * (kAPlane + kBPlane);
  constexpr int kAH = kBM * 4 / NTH;            // float4 A loads per thread and stage
  constexpr int kBT = 4 * kBN;                  // float4 B loads per stage
  constexpr int kBH = (kBT + NTH - 1) / NTH;    // ... per thread
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* smem = reinterpret_cast<T*>(smem_raw);

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = tid >> 6;
  const int wm = wave / WN;  // 32 I rows each
  const int wn = wave % WN;  // 16*J complex columns each
  float sa = 1.f, sb = 1.f;
  int m0 = m0_first, n0 = n0_first;  // the tile being computed

  // global -> register assignment (f = tid + 512*h, g = tid + 512*h)
  // A, AM 0/1: row = f >> 2, q = f & 3 (complex k = 2q, 2q+1 of the stage)
  // A, AM 2:   kk = f & 7, row = 2*(f >> 3) (complex rows row, row+1)
  // B: kk = g & 7 (consecutive lanes take consecutive k, so the LDS stores
  //    are conflict-free), nq = g >> 3 (complex n = 2nq, 2nq+1)
  // Each operand is read through a buffer resource with 32-bit offsets.
  // Normally (WIDE false) the resource covers the whole operand, made once.
  // WIDE: it is based at the block's first row (A) or column (B) and at the
  // stage's first k, so only one stage's footprint in the block, not the
  // whole operand, has to fit the offsets (operands above 2 GiB; some scalar
  // work per stage, within the noise of the whole-operand form in time, used
  // only where the whole operand does not fit).  For a
  // digit map, offset(i0 + d) = offset(i0) + offset(d) when i0 is a
  // multiple of the power-of-two tile edge (or of the stage depth) and d is
  // below it.
  uint64_t a_base = 0, b_base = 0, a_span, b_span;  // in elements
  if constexpr (!WIDE) {
    if constexpr (AM == 0) {
      a_span = uint64_t(M) * uint64_t(K);
      b_span = uint64_t(K) * uint64_t(N);
    } else {
      a_span = map_offset(a_rows, uint64_t(M - 1)) + map_offset(a_cols, uint64_t(K - 1)) + 1;
      b_span = map_offset(b_rows, uint64_t(K - 1)) + map_offset(b_cols, uint64_t(N - 1)) + 1;
    }
  } else if constexpr (AM == 0) {
    a_base = uint64_t(m0) * uint64_t(K);
    b_base = uint64_t(n0);
    a_span = uint64_t(kBM - 1) * uint64_t(K) + kBBK;
    b_span = uint64_t(kBBK - 1) * uint64_t(N) + kBN;
  } else {
    a_base = map_offset(a_rows, uint64_t(m0));
    b_base = map_offset(b_cols, uint64_t(n0));
    a_span = map_offset(a_rows, uint64_t(min(kBM, M) - 1)) +
             map_offset(a_cols, uint64_t(min(kBBK, K) - 1)) + 1;
    b_span = map_offset(b_cols, uint64_t(min(kBN, N) - 1)) +
             map_offset(b_rows, uint64_t(min(kBBK, K) - 1)) + 1;
  }
  uint64_t a_off[kAH], a_coff = 0, b_off[kBH], b_roff = 0;
  // the next tile of a chain: its origin and its lanes' row / column offsets
  // (32 bits: chained operands fit the whole-operand resources)
  unsigned Lc = L0;
  bool has_next = false;
  int m0n = 0, n0n = 0;
  // The 128 x 256 tile's digit-mapped forms keep the next tile's offsets in
  // LDS (each lane its own slots, behind the per-tile maxima): in registers
  // they would spill.
  constexpr bool kNextInLds = CHAIN && WM == 2 && WN == 4 && AM != 0;
  unsigned a_offn[kNextInLds ? 1 : kAH], b_offn[kNextInLds ? 1 : kBH];
  auto next_slots = [&]() {
    return reinterpret_cast<unsigned*>(smem_raw + 2 * kBuf * sizeof(T) +
                                       (AM != 0 ? 16 * ((K + kBBK - 1) / kBBK) : 0) + 64);
  };
  auto tile_rows = [&](int tm0, auto* aoff) {
#pragma unroll
    for (int h = 0; h < kAH; ++h) {
      const int f = tid + NTH * h;
      const int row = AM == 1 ? (f >> 2) : 2 * (f >> 3);
      aoff[h] = map_offset(a_rows, uint64_t(WIDE ? row : tm0 + row));
    }
  };
  auto tile_cols = [&](int tn0, auto* boff) {
#pragma unroll
    for (int h = 0; h < kBH; ++h) {
      const int g = tid + NTH * h;
      boff[h] = map_offset(b_cols, uint64_t(WIDE ? 2 * (g >> 3) : tn0 + 2 * (g >> 3)));
    }
  };
  auto set_next = [&]() {  // the chain's tile after Lc, if any
    has_next = Lc + Lstride < Lend;
    if (has_next) {
      const unsigned L = Lc + Lstride;
      m0n = int(L / gy) * kBM;
      n0n = int(L % gy) * kBN;
      if constexpr (kNextInLds) {
        uint64_t ta[kAH], tb[kBH];
        tile_rows(m0n, ta);
        tile_cols(n0n, tb);
        unsigned* nl = next_slots();
#pragma unroll
        for (int h = 0; h < kAH; ++h) nl[h * NTH + tid] = unsigned(ta[h]);
#pragma unroll
        for (int h = 0; h < kBH; ++h) nl[(kAH + h) * NTH + tid] = unsigned(tb[h]);
      } else if constexpr (AM != 0) {
        tile_rows(m0n, a_offn);
        tile_cols(n0n, b_offn);
      }
    }
  };
  if constexpr (AM != 0) {
    tile_rows(m0, a_off);
    a_coff = map_offset(a_cols, AM == 1 ? uint64_t(2 * (tid & 3)) : uint64_t(tid & 7));
    tile_cols(n0, b_off);
    b_roff = map_offset(b_rows, uint64_t(tid & 7));
  }
  if constexpr (CHAIN) set_next();
  // Per-stage k offsets of A and B (uniform), behind the two LDS buffers,
  // so that the pipelined stage has no digit arithmetic.
  uint64_t* ktab = reinterpret_cast<uint64_t*>(smem_raw + 2 * kBuf * sizeof(T));
  const int nkt = (K + kBBK - 1) / kBBK;
  if constexpr (AM != 0) {
    for (int t = tid; t < nkt; t += NTH) {
      ktab[2 * t] = map_offset(a_cols, uint64_t(t) * kBBK);
      ktab[2 * t + 1] = map_offset(b_rows, uint64_t(t) * kBBK);
    }
    __syncthreads();
  }
  // Operands are read through buffer resources: an element outside the
  // tile's range gets an offset past the end of the buffer and reads as 0,
  // so there is no branch and no select near the loads (the waits for a
  // stage's operands then never include the next stage's loads).  The host
  // checks that both spans fit (big_operands_whole, big_mapped_whole; WIDE:
  // big_operands_fit, big_mapped_spans_fit).
  // (uniform values: made scalar so that the resources stay in SGPRs)
  const int a_rec = __builtin_amdgcn_readfirstlane(int(a_span * sizeof(float2)));
  const int b_rec = __builtin_amdgcn_readfirstlane(int(b_span * sizeof(float2)));
  const __amdgpu_buffer_rsrc_t rsrc_a_whole =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(A), 0, a_rec, 0x00020000);
  const __amdgpu_buffer_rsrc_t rsrc_b_whole =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(B), 0, b_rec, 0x00020000);
  constexpr unsigned kOutside = 0xfffffff0u;
  auto load16 = [](__amdgpu_buffer_rsrc_t r, unsigned byte_off) {
    const u32x4 v = __builtin_amdgcn_raw_buffer_load_b128(r, int(byte_off), 0, 0);
    return make_float4(__uint_as_float(v.x), __uint_as_float(v.y), __uint_as_float(v.z),
                       __uint_as_float(v.w));
  };
  auto load_tile = [&](int kt, float4* ra, float4* rb, bool nx = false) {
    const int k0 = kt * kBBK;
    const int tm0 = CHAIN && nx ? m0n : m0, tn0 = CHAIN && nx ? n0n : n0;
    uint64_t a_stage = 0, b_stage = 0;  // WIDE: the stage's first k, from the block bases
    unsigned a_kbase = 0, b_kbase = 0;  // otherwise: its offsets from the operand bases
    if constexpr (!WIDE) {
      if constexpr (AM != 0) {
        a_kbase = unsigned(ktab[2 * kt]);
        b_kbase = unsigned(ktab[2 * kt + 1]);
      }
    } else if constexpr (AM != 0) {
      // uniform values read from LDS: made scalar so the resources stay in SGPRs
      const uint64_t ta = ktab[2 * kt], tb = ktab[2 * kt + 1];
      a_stage = uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(ta))))) |
                (uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(ta >> 32))))) << 32);
      b_stage = uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(tb))))) |
                (uint64_t(uint32_t(__builtin_amdgcn_readfirstlane(int(uint32_t(tb >> 32))))) << 32);
    } else {
      a_stage = uint64_t(k0);
      b_stage = uint64_t(k0) * uint64_t(N);
    }
    const __amdgpu_buffer_rsrc_t rsrc_a =
        WIDE ? __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(A + a_base + a_stage), 0,
                                                 a_rec, 0x00020000)
             : rsrc_a_whole;
    const __amdgpu_buffer_rsrc_t rsrc_b =
        WIDE ? __builtin_amdgcn_make_buffer_rsrc(const_cast<float2*>(B + b_base + b_stage), 0,
                                                 b_rec, 0x00020000)
             : rsrc_b_whole;
    // row and column origins of the offsets: the block's (WIDE) or 0
    const int mo = WIDE ? m0 : 0, no = WIDE ? n0 : 0, ko = WIDE ? k0 : 0;
#pragma unroll
    for (int h = 0; h < kAH; ++h) {
      const int f = tid + NTH * h;
      int gm, gk;
      unsigned off;
      unsigned arow;
      if constexpr (kNextInLds)
        arow = nx ? next_slots()[h * NTH + tid] : unsigned(a_off[h]);
      else
        arow = CHAIN && nx ? a_offn[h] : unsigned(a_off[h]);
      if constexpr (AM == 2) {
        gm = tm0 + 2 * (f >> 3);
        gk = k0 + (f & 7);
        off = arow + a_kbase + unsigned(a_coff);
      } else {
        gm = tm0 + (f >> 2);
        gk = k0 + 2 * (f & 3);
        off = AM == 0 ? unsigned(gm - mo) * unsigned(K) + unsigned(gk - ko)
                      : arow + a_kbase + unsigned(a_coff);
      }
      ra[h] = load16(rsrc_a, (gm < M && gk < K) ? off * unsigned(sizeof(float2)) : kOutside);
    }
#pragma unroll
    for (int h = 0; h < kBH; ++h) {
      const int g = tid + NTH * h;
      const int gk = k0 + (g & 7), gn = tn0 + 2 * (g >> 3);
      unsigned bcol;
      if constexpr (kNextInLds)
        bcol = nx ? next_slots()[(kAH + h) * NTH + tid] : unsigned(b_off[h]);
      else
        bcol = CHAIN && nx ? b_offn[h] : unsigned(b_off[h]);
      const unsigned off = AM == 0 ? unsigned(gk - ko) * unsigned(N) + unsigned(gn - no)
                                   : b_kbase + unsigned(b_roff) + bcol;
      const bool in = (kBT % NTH == 0 || g < kBT) && gk < K && gn < N;
      rb[h] = load16(rsrc_b, in ? off * unsigned(sizeof(float2)) : kOutside);
    }
  };
  typedef T T2 __attribute__((ext_vector_type(2)));
  typedef T T4 __attribute__((ext_vector_type(4)));
  auto store_tile = [&](int buf, const float4* ra, const float4* rb) {
    T* base = smem + buf * kBuf;
#pragma unroll
    for (int h = 0; h < kAH; ++h) {
      const int f = tid + NTH * h;
      float v[4] = {ra[h].x * sa, ra[h].y * sa, ra[h].z * sa, ra[h].w * sa};
      T t3[4][3];
#pragma unroll
      for (int e = 0; e < 4; ++e) split_terms<MODE>(v[e], t3[e]);
      if constexpr (AM == 2) {
        // (re, im) of rows row and row+1 at complex k = kk
        const int row = 2 * (f >> 3), kk = f & 7;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          T2 w0 = {t3[0][t], t3[1][t]};
          T2 w1 = {t3[2][t], t3[3][t]};
          *reinterpret_cast<T2*>(base + t * kAPlane + row * kRowL + 2 * kk) = w0;
          *reinterpret_cast<T2*>(base + t * kAPlane + (row + 1) * kRowL + 2 * kk) = w1;
        }
      } else {
        const int row = f >> 2, q = f & 3;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          T4 w = {t3[0][t], t3[1][t], t3[2][t], t3[3][t]};
          *reinterpret_cast<T4*>(base + t * kAPlane + row * kRowL + 4 * q) = w;
        }
      }
    }
    T* bb = base + NT * kAPlane;
#pragma unroll
    for (int h = 0; h < kBH; ++h) {
      const int g = tid + NTH * h;
      if (kBT % NTH != 0 && g >= kBT) continue;
      const int kk = g & 7, nq = g >> 3;
      // (Br, Bi) of n, then of n+1
      float v[4] = {rb[h].x * sb, rb[h].y * sb, rb[h].z * sb, rb[h].w * sb};
      T t3[4][3];
#pragma unroll
      for (int e = 0; e < 4; ++e) split_terms<MODE>(v[e], t3[e]);
#pragma unroll
      for (int t = 0; t < NT; ++t) {
#pragma unroll
        for (int c = 0; c < 2; ++c) {
          const int n = 2 * nq + c;
          const T br = t3[2 * c][t], bi = t3[2 * c + 1][t];
          T2 p0 = {br, static_cast<T>(-static_cast<float>(bi))};
          T2 p1 = {bi, br};
          *reinterpret_cast<T2*>(bb + t * kBPlane + (2 * n) * kRowL + 2 * kk) = p0;
          *reinterpret_cast<T2*>(bb + t * kBPlane + (2 * n + 1) * kRowL + 2 * kk) = p1;
        }
      }
    }
  };

  const int half = lane >> 5;
  const int col = lane & 31;
  const int part = col & 1;

  floatx16 acc[I][J];
#pragma unroll
  for (int i = 0; i < I; ++i)
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;

  auto compute = [&](int buf) {
    const T* base = smem + buf * kBuf;
    const T* bb = base + NT * kAPlane;
    V a[NT][I], b[NT][J];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
#pragma unroll
      for (int i = 0; i < I; ++i)
        a[t][i] = *reinterpret_cast<const V*>(base + t * kAPlane +
                                              (wm * 32 * I + i * 32 + col) * kRowL + 8 * half);
#pragma unroll
      for (int j = 0; j < J; ++j) {
        const int rc = wn * 32 * J + j * 32 + col;  // real column of Bhat in the block
        b[t][j] = *reinterpret_cast<const V*>(bb + t * kBPlane + rc * kRowL + 8 * half);
      }
    }
#pragma unroll
    for (int i = 0; i < I; ++i)
#pragma unroll
      for (int j = 0; j < J; ++j) {
        if constexpr (SplitTraits<MODE>::PRODUCTS == 6) {
          acc[i][j] = mfma16<MODE>(a[2][i], b[0][j], acc[i][j]);
          acc[i][j] = mfma16<MODE>(a[0][i], b[2][j], acc[i][j]);
          acc[i][j] = mfma16<MODE>(a[1][i], b[1][j], acc[i][j]);
        }
        acc[i][j] = mfma16<MODE>(a[1][i], b[0][j], acc[i][j]);
        acc[i][j] = mfma16<MODE>(a[0][i], b[1][j], acc[i][j]);
        acc[i][j] = mfma16<MODE>(a[0][i], b[0][j], acc[i][j]);
      }
  };
  // Interleave of one stage: global loads, fragment reads, then MFMAs in
  // pairs around groups of VALU (operand split) and one LDS store.
  auto schedule = [&]() {
    constexpr int kMfma = I * J * SplitTraits<MODE>::PRODUCTS;
    __builtin_amdgcn_sched_group_barrier(0x100, (I + J) * NT, 0);  // DS read
#pragma unroll
    for (int i = 0; i < kMfma / 2; ++i) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
      __builtin_amdgcn_sched_group_barrier(0x002, 6, 0);  // VALU
      __builtin_amdgcn_sched_group_barrier(0x200, 1, 0);  // DS write
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
    }
  };

  // Two register sets: stage s is loaded into set s & 1 at stage s - 2 and
  // split into LDS buffer s & 1 at stage s - 1.  (A third set, hiding two
  // stages of latency, measured slower: the kernel is not latency-bound.)
  float4 ra[2][kAH], rb[2][kBH];
  load_tile(0, ra[0], rb[0]);
  // the scales are reduced while the first loads are in flight (in the
  // second stage buffer, first written after the prologue's barrier)
  if constexpr (MODE == 2)
    operand_scales<NTH>(amax, reinterpret_cast<float*>(smem + kBuf), sa, sb);
  load_tile(min(1, nkt - 1), ra[1], rb[1]);
  store_tile(0, ra[0], rb[0]);
  __syncthreads();
  // One stage: the loads of stage s + 2 stay at its top (sched_barrier),
  // then the MFMAs of stage s interleaved with the split of stage s + 1.
  // The main loop runs pairs of stages unconditionally (a conditional stage
  // would let the compiler sink earlier loads into it); an odd last stage
  // follows it.
  auto stage = [&](auto si, int s) {
    constexpr int SI = decltype(si)::value;
    if constexpr (CHAIN) {
      // the tile's last two stages load the next tile's first two
      const bool nx = has_next && s + 2 >= nkt;
      load_tile(nx ? s + 2 - nkt : min(s + 2, nkt - 1), ra[SI], rb[SI], nx);
    } else {
      load_tile(min(s + 2, nkt - 1), ra[SI], rb[SI]);
    }
    __builtin_amdgcn_sched_barrier(0);
    compute(SI);
    store_tile(SI ^ 1, ra[SI ^ 1], rb[SI ^ 1]);
    schedule();
    __syncthreads();
  };
  const float unscale = 1.f / (sa * sb);
  for (;;) {  // the tiles of a chain (one tile otherwise)
  int s0 = 0;
  for (; s0 + 2 <= nkt; s0 += 2) {
    stage(std::integral_constant<int, 0>{}, s0);
    stage(std::integral_constant<int, 1>{}, s0 + 1);
  }
  if (s0 < nkt) compute(0);  // stage nkt - 1

  float* Cf = reinterpret_cast<float*>(C);
  float cm = 0.f;  // max |re|, |im| of the values this thread stores
#pragma unroll
  for (int i = 0; i < I; ++i)
#pragma unroll
    for (int j = 0; j < J; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int row = m0 + wm * 32 * I + i * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
        const int cn = n0 + wn * 16 * J + j * 16 + (col >> 1);
        if (row < M && cn < N) {
          const float v = acc[i][j][r] * unscale;
          // Streaming stores: C of these products is 128 MiB to 1 GiB, and
          // its lines left dirty in the last-level cache slow the kernels
          // that stream through it next (0.9% of the 7 x 7 depth-32 job).
          __builtin_nontemporal_store(v, &Cf[(uint64_t(row) * N + cn) * 2 + part]);
          cm = fmaxf(cm, fabsf(v));
        }
      }
  if (cmax_slot) {
    // one partial maximum per tile, for the split of the GEMM that reads C
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) cm = fmaxf(cm, __shfl_xor(cm, off));
    // (a chain's next tile already has its first stage in buffer 0: the
    // maxima then go behind the stage buffers and the k table)
    float* red = reinterpret_cast<float*>(
        CHAIN ? smem_raw + 2 * kBuf * sizeof(T) + (AM != 0 ? 16 * nkt : 0) : smem_raw);
    __syncthreads();  // the last stage's fragment reads are done
    if (lane == 0) red[wave] = cm;
    __syncthreads();
    if (tid == 0) {
      float b = red[0];
#pragma unroll
      for (int w = 1; w < NTH / 64; ++w) b = fmaxf(b, red[w]);
      cmax_slot[CHAIN ? Lc : 0] = b;
    }
  }
  if constexpr (!CHAIN) {
    break;
  } else {
    if (!has_next) break;
    Lc += Lstride;
    m0 = m0n;
    n0 = n0n;
    if constexpr (kNextInLds) {
      const unsigned* nl = next_slots();
#pragma unroll
      for (int h = 0; h < kAH; ++h) a_off[h] = nl[h * NTH + tid];
#pragma unroll
      for (int h = 0; h < kBH; ++h) b_off[h] = nl[(kAH + h) * NTH + tid];
    } else if constexpr (AM != 0) {
#pragma unroll
      for (int h = 0; h < kAH; ++h) a_off[h] = a_offn[h];
#pragma unroll
      for (int h = 0; h < kBH; ++h) b_off[h] = b_offn[h];
    }
    set_next();
#pragma unroll
    for (int i = 0; i < I; ++i)
#pragma unroll
      for (int j = 0; j < J; ++j)
#pragma unroll
        for (int r = 0; r < 16; ++r) acc[i][j][r] = 0.f;
  }
  }
}

template <int MODE, int WM, int WN, int J, int AM, int I = 2, bool WIDE = false>
__global__ __launch_bounds__(64 * WM * WN) void cgemm_split_big_kernel(
    const float2* __restrict__ A, const float2* __restrict__ B, float2* __restrict__ C, int M,
    int N, int K, AmaxRef amax, qfx_index_map a_rows,
    qfx_index_map a_cols, qfx_index_map b_rows, qfx_index_map b_cols,
    float* __restrict__ cmax) {
  // Column tiles vary fastest in dispatch order (the grid's x counts row
  // tiles): 2-4% faster than row tiles fastest on the 65536-131072 x 256-512
  // products, the same on 4096^3; the tile computation is unchanged.
  const unsigned L = blockIdx.x + gridDim.x * blockIdx.y;
  const unsigned tn = L % gridDim.y, tm = L / gridDim.y;
  split_big_tile<MODE, WM, WN, J, AM, I, WIDE>(
      A, B, C, M, N, K, amax, a_rows, a_cols, b_rows, b_cols, cmax ? cmax + L : nullptr,
      int(tm) * 32 * I * WM, int(tn) * 16 * J * WN);
}

// Chained tiles (split_big_tile with CHAIN): one block per CU, block b
// computing tiles b, b + gridDim.x, ... in the column-fastest order of
// cgemm_split_big_kernel; cmax (if any) gets one maximum per tile.
template <int MODE, int WM, int WN, int J, int AM, int I = 2>
__global__ __launch_bounds__(64 * WM * WN) void cgemm_split_chain_kernel(
    const float2* __restrict__ A, const float2* __restrict__ B, float2* __restrict__ C, int M,
    int N, int K, AmaxRef amax, qfx_index_map a_rows, qfx_index_map a_cols,
    qfx_index_map b_rows, qfx_index_map b_cols, float* __restrict__ cmax, unsigned ntiles,
    unsigned gy) {
  const unsigned L = blockIdx.x;
  split_big_tile<MODE, WM, WN, J, AM, I, false, true>(
      A, B, C, M, N, K, amax, a_rows, a_cols, b_rows, b_cols, cmax, int(L / gy) * 32 * I * WM,
      int(L % gy) * 16 * J * WN, L, gridDim.x, ntiles, gy);
}

// One problem of a grouped launch: row-major operands, 256 x 128 tiles.
struct GroupTile {
  const float2* A;
  const float2* B;
  float2* C;
  int M, N, K, tiles_m;
  uint64_t first_tile;  // index of the problem's first tile in the launch
  AmaxRef amax;
};

// Persistent grouped launch: the blocks walk the tiles of all problems
// (problems of any shape, tiles numbered problem by problem), each block
// taking every gridDim.x-th tile.
template <int MODE>
__global__ __launch_bounds__(512) void cgemm_split_grouped_kernel(const GroupTile* __restrict__ probs,
                                                                  int count, uint64_t total) {
  const qfx_index_map none{};
  for (uint64_t t = blockIdx.x; t < total; t += gridDim.x) {
    int lo = 0, hi = count - 1;  // the last problem whose first tile is <= t
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (probs[mid].first_tile <= t) lo = mid;
      else hi = mid - 1;
    }
    const GroupTile g = probs[lo];
    const uint64_t local = t - g.first_tile;
    const int tm = int(local % uint64_t(g.tiles_m)), tn = int(local / uint64_t(g.tiles_m));
    split_big_tile<MODE, 4, 2, 4, 0>(g.A, g.B, g.C, g.M, g.N, g.K, g.amax, none, none, none, none,
                                     nullptr, tm * 256, tn * 128);
    __syncthreads();  // the next tile's prologue rewrites the LDS
  }
}

template <int MODE, int WM, int WN, int J, int I = 2>
constexpr size_t split_big_smem() {
  return size_t(2) * SplitTraits<MODE>::NT * (32 * I * WM * kRowL + 2 * 16 * J * WN * kRowL) * 2;
}

constexpr size_t kMaxLds = 160 * 1024;

// LDS of one launch: the two stage buffers, plus 16 bytes per stage for the
// per-stage k offsets of the mapped forms.
template <int MODE, int WM, int WN, int J, int AM, int I = 2>
size_t split_big_lds(uint64_t k) {
  return split_big_smem<MODE, WM, WN, J, I>() + (AM != 0 ? 16 * ((k + kBBK - 1) / kBBK) : 0);
}

// Dynamic LDS above 64 KiB must be allowed per kernel once per process.
template <int MODE, int WM, int WN, int J, int AM, int I = 2, bool WIDE = false>
hipError_t split_big_prepare() {
  static hipError_t e = hipFuncSetAttribute(
      reinterpret_cast<const void*>(&cgemm_split_big_kernel<MODE, WM, WN, J, AM, I, WIDE>),
      hipFuncAttributeMaxDynamicSharedMemorySize,
      int(AM != 0 ? kMaxLds : split_big_smem<MODE, WM, WN, J, I>()));
  return e;
}

// Compute units of the current device (the grid of a chained launch).
unsigned device_cus() {
  static int cached[64] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cached[dev]) {
    int c = 0;
    if (hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      c = 0;
    cached[dev] = c > 0 ? c : 256;
  }
  return unsigned(cached[dev]);
}

// Products whose tiles run as chains: whole-operand addressing, an even
// number of k-stages, and at least two tiles for every compute unit.
bool split_chain_applies(bool wide, uint64_t tiles, uint64_t k) {
  const uint64_t nkt = (k + kBBK - 1) / kBBK;
  return !wide && nkt % 2 == 0 && tiles >= 2 * uint64_t(device_cus()) && tiles <= 0xffffffffULL;
}

template <int MODE, int WM, int WN, int J, int AM, bool WIDE, int I = 2>
hipError_t launch_split_big(hipStream_t stream, const float2* A, const float2* B, float2* C,
                            uint64_t m, uint64_t n, uint64_t k, AmaxRef amax,
                            const qfx_index_map* maps, float* cmax) {
  hipError_t e = split_big_prepare<MODE, WM, WN, J, AM, I, WIDE>();
  if (e != hipSuccess) return e;
  constexpr int kBM = 32 * I * WM, kBN = 16 * J * WN;
  dim3 g{unsigned((m + kBM - 1) / kBM), unsigned((n + kBN - 1) / kBN), 1};
  qfx_index_map none{};
  // (the three-plane BF16 split would spill in a chain: the next tile's
  // offsets do not fit beside its 256 VGPRs)
  if constexpr (!WIDE && MODE != 0) {
    const uint64_t tiles = uint64_t(g.x) * g.y;
    // the per-tile maxima go behind the stage buffers and the k table, and
    // for the 128 x 256 tile's digit-mapped forms the next tile's lane
    // offsets (4 bytes per lane and load) behind them
    constexpr int kAH = 32 * I * WM * 4 / (64 * WM * WN);
    constexpr int kBH = (4 * 16 * J * WN + 64 * WM * WN - 1) / (64 * WM * WN);
    const size_t lds = split_big_lds<MODE, WM, WN, J, AM, I>(k) + 64 +
                       (WM == 2 && WN == 4 && AM != 0 ? size_t(4) * (kAH + kBH) * 64 * WM * WN : 0);
    if (lds <= kMaxLds && split_chain_applies(WIDE, tiles, k)) {
      static hipError_t ce = hipFuncSetAttribute(
          reinterpret_cast<const void*>(&cgemm_split_chain_kernel<MODE, WM, WN, J, AM, I>),
          hipFuncAttributeMaxDynamicSharedMemorySize, int(kMaxLds));
      if (ce != hipSuccess) return ce;
      const unsigned grid = unsigned(std::min<uint64_t>(tiles, device_cus()));
      cgemm_split_chain_kernel<MODE, WM, WN, J, AM, I><<<grid, 64 * WM * WN, lds, stream>>>(
          A, B, C, int(m), int(n), int(k), amax, maps ? maps[0] : none, maps ? maps[1] : none,
          maps ? maps[2] : none, maps ? maps[3] : none, cmax, unsigned(tiles), g.y);
      return hipGetLastError();
    }
  }
  cgemm_split_big_kernel<MODE, WM, WN, J, AM, I, WIDE>
      <<<g, 64 * WM * WN, split_big_lds<MODE, WM, WN, J, AM, I>(k), stream>>>(
          A, B, C, int(m), int(n), int(k), amax, maps ? maps[0] : none, maps ? maps[1] : none,
          maps ? maps[2] : none, maps ? maps[3] : none, cmax);
  return hipGetLastError();
}

// Tile layouts of the large-tile kernel (8 waves as WM x WN, J column tiles
// per wave): 4x2 (256 x 128 complex), 2x4 (128 x 256), 8x1 with J = 2
// (512 x 32, for n < 128).
enum BigLayout { kLayoutNone = 0, kLayout4x2, kLayout2x4, kLayout8x1 };

// The large-tile kernel addresses each operand with 32-bit byte offsets
// into a buffer resource based at the block's first row of A (first column
// of B) and at the stage's first k: one stage of a block reads at most 512
// rows of A over 8 k, and 8 rows of B over at most 256 complex columns.
constexpr uint64_t kSpanLimit = (uint64_t(1) << 31) - 64;
bool big_operands_fit(uint64_t m, uint64_t n, uint64_t k) {
  return 512 * k * sizeof(float2) <= kSpanLimit && 8 * n * sizeof(float2) <= kSpanLimit &&
         m <= 0x7fffffffULL && n <= 0x7fffffffULL && k <= 0x7fffffffULL;
}

static uint64_t map_offset_host(const qfx_index_map& d, uint64_t idx) {
  uint64_t off = 0;
  for (int j = 0; j < d.ndigits; ++j) {
    off += (idx & ((uint64_t(1) << d.bits[j]) - 1)) << d.shift[j];
    idx >>= d.bits[j];
  }
  return off;
}

// Digit-mapped operands: what one stage of one block reads (every tile
// shape: at most 512 rows of A and 256 complex columns of B, 8 k) must fit
// the buffer offsets.
bool big_mapped_spans_fit(const qfx_index_map& a_rows, const qfx_index_map& a_cols,
                          const qfx_index_map& b_rows, const qfx_index_map& b_cols, uint64_t m,
                          uint64_t n, uint64_t k) {
  const uint64_t a_span = map_offset_host(a_rows, std::min<uint64_t>(512, m) - 1) +
                          map_offset_host(a_cols, std::min<uint64_t>(kBBK, k) - 1) + 1;
  const uint64_t b_span = map_offset_host(b_cols, std::min<uint64_t>(256, n) - 1) +
                          map_offset_host(b_rows, std::min<uint64_t>(kBBK, k) - 1) + 1;
  return a_span * sizeof(float2) <= kSpanLimit && b_span * sizeof(float2) <= kSpanLimit &&
         m <= 0x7fffffffULL && n <= 0x7fffffffULL && k <= 0x7fffffffULL;
}

// Whether each operand, whole, fits the 32-bit offsets of one buffer
// resource (the default large-tile addressing); otherwise the kernel bases
// its resources per block and stage (WIDE).
bool big_operands_whole(uint64_t m, uint64_t n, uint64_t k) {
  return m * k * sizeof(float2) <= kSpanLimit && k * n * sizeof(float2) <= kSpanLimit;
}
bool big_mapped_whole(const qfx_index_map& a_rows, const qfx_index_map& a_cols,
                      const qfx_index_map& b_rows, const qfx_index_map& b_cols, uint64_t m,
                      uint64_t n, uint64_t k) {
  const uint64_t a_span = map_offset_host(a_rows, m - 1) + map_offset_host(a_cols, k - 1) + 1;
  const uint64_t b_span = map_offset_host(b_rows, k - 1) + map_offset_host(b_cols, n - 1) + 1;
  return a_span * sizeof(float2) <= kSpanLimit && b_span * sizeof(float2) <= kSpanLimit;
}

BigLayout choose_big_layout(int precision, uint64_t m, uint64_t n) {
  auto blocks = [&](uint64_t bm, uint64_t bn) { return ((m + bm - 1) / bm) * ((n + bn - 1) / bn); };
  // 128 x 256 measured faster for tall operands (m >= 128 n), 256 x 128 otherwise
  if (n >= 512 && m >= 128 * n && precision != QFX_PREC_SPLIT_BF16X3 && blocks(128, 256) >= 128)
    return kLayout2x4;
  if (n >= 128 && blocks(256, 128) >= 128) return kLayout4x2;
  if (n >= 512 && precision != QFX_PREC_SPLIT_BF16X3 && blocks(128, 256) >= 128) return kLayout2x4;
  // three BF16 planes of the 512-row tile would exceed the 160 KiB of LDS
  if (n >= 32 && n < 128 && precision != QFX_PREC_SPLIT_BF16X3 && blocks(512, 32) >= 128)
    return kLayout8x1;
  return kLayoutNone;
}

template <int MODE, int AM, bool WIDE>
hipError_t launch_split_big_layout(BigLayout lay, hipStream_t stream, const float2* A,
                                   const float2* B, float2* C, uint64_t m, uint64_t n,
                                   uint64_t k, AmaxRef amax,
                                   const qfx_index_map* maps, float* cmax) {
  if constexpr (MODE == 0) {
    return launch_split_big<0, 4, 2, 4, AM, WIDE>(stream, A, B, C, m, n, k, amax, maps, cmax);
  } else {
    if (lay == kLayout2x4)
      return launch_split_big<MODE, 2, 4, 4, AM, WIDE>(stream, A, B, C, m, n, k, amax, maps, cmax);
    if (lay == kLayout8x1)
      return launch_split_big<MODE, 8, 1, 2, AM, WIDE>(stream, A, B, C, m, n, k, amax, maps, cmax);
    return launch_split_big<MODE, 4, 2, 4, AM, WIDE>(stream, A, B, C, m, n, k, amax, maps, cmax);
  }
}

// Whether the mapped form's per-stage table fits beside the stage buffers.
bool split_big_mapped_fits(int precision, BigLayout lay, uint64_t k) {
  if (precision == QFX_PREC_SPLIT_BF16X3) return split_big_lds<0, 4, 2, 4, 1>(k) <= kMaxLds;
  if (lay == kLayout2x4) return split_big_lds<2, 2, 4, 4, 1>(k) <= kMaxLds;
  if (lay == kLayout8x1) return split_big_lds<2, 8, 1, 2, 1>(k) <= kMaxLds;
  return split_big_lds<2, 4, 2, 4, 1>(k) <= kMaxLds;
}

template <int AM, bool WIDE>
hipError_t launch_split_big_mode(int precision, BigLayout lay, hipStream_t stream,
                                 const float2* A, const float2* B, float2* C, uint64_t m,
                                 uint64_t n, uint64_t k, AmaxRef amax,
                                 const qfx_index_map* maps, float* cmax) {
  if (precision == QFX_PREC_SPLIT_BF16X3)
    return launch_split_big_layout<0, AM, WIDE>(lay, stream, A, B, C, m, n, k, amax, maps, cmax);
  if (precision == QFX_PREC_SPLIT_BF16)
    return launch_split_big_layout<1, AM, WIDE>(lay, stream, A, B, C, m, n, k, amax, maps, cmax);
  return launch_split_big_layout<2, AM, WIDE>(lay, stream, A, B, C, m, n, k, amax, maps, cmax);
}

// wide: at least one operand spans more than the 32-bit offsets of one
// resource (per-stage resources then).
template <int AM>
hipError_t launch_split_big_prec(int precision, BigLayout lay, hipStream_t stream,
                                 const float2* A, const float2* B, float2* C, uint64_t m,
                                 uint64_t n, uint64_t k, AmaxRef amax,
                                 const qfx_index_map* maps, float* cmax, bool wide) {
  return wide ? launch_split_big_mode<AM, true>(precision, lay, stream, A, B, C, m, n, k, amax,
                                                maps, cmax)
              : launch_split_big_mode<AM, false>(precision, lay, stream, A, B, C, m, n, k, amax,
                                                 maps, cmax);
}

// Blocks (one partial maximum each) of a large-tile launch.
uint64_t big_layout_blocks(BigLayout lay, uint64_t m, uint64_t n) {
  const uint64_t bm = lay == kLayout2x4 ? 128 : (lay == kLayout8x1 ? 512 : 256);
  const uint64_t bn = lay == kLayout2x4 ? 256 : (lay == kLayout8x1 ? 32 : 128);
  return ((m + bm - 1) / bm) * ((n + bn - 1) / bn);
}

// Tiles that can run a product of this precision and operand form.
std::vector<BigLayout> layout_candidates(int precision, int am, uint64_t m, uint64_t n,
                                         uint64_t k) {
  std::vector<BigLayout> out;
  if (precision == QFX_PREC_SPLIT_BF16X3) {  // its three planes fit the 256 x 128 tile only
    out.push_back(kLayout4x2);
    return out;
  }
  for (BigLayout l : {kLayout4x2, kLayout2x4, kLayout8x1}) {
    if (l == kLayout8x1 && n > 256) continue;  // 32-column tiles for narrow products only
    if (am != 0 && !split_big_mapped_fits(precision, l, k)) continue;
    const uint64_t bn = l == kLayout2x4 ? 256 : (l == kLayout8x1 ? 32 : 128);
    if ((n + bn - 1) / bn > 65535) continue;
    out.push_back(l);
  }
  (void)m;
  return out;
}

// The tile of a large-tile product: the fixed rule's choice, or, with the
// autotuner on, the fastest candidate (each run twice on first use; every
// candidate writes the same C, so the caller's own launch follows).
template <class Launch>
BigLayout tuned_layout(qfx_context* ctx, int precision, int am, uint64_t m, uint64_t n,
                       uint64_t k, BigLayout def, Launch launch, hipError_t* err) {
  *err = hipSuccess;
  if (!ctx->autotune || def == kLayoutNone) return def;
  const std::array<uint64_t, 5> key{uint64_t(precision), uint64_t(am), m, n, k};
  auto it = ctx->tuned.find(key);
  if (it != ctx->tuned.end()) return BigLayout(it->second);
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(ctx->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
    return def;  // no timing (and no synchronization) inside a graph capture
  const std::vector<BigLayout> cands = layout_candidates(precision, am, m, n, k);
  if (cands.size() < 2) {
    ctx->tuned[key] = def;
    return def;
  }
  for (auto& e : ctx->tune_ev)
    if (!e && (*err = hipEventCreate(&e)) != hipSuccess) return def;
  BigLayout best = def;
  float best_ms = 1e30f;
  for (BigLayout l : cands)
    for (int rep = 0; rep < 2; ++rep) {
      if ((*err = hipEventRecord(ctx->tune_ev[0], ctx->stream)) != hipSuccess) return def;
      if ((*err = launch(l)) != hipSuccess) return def;
      if ((*err = hipEventRecord(ctx->tune_ev[1], ctx->stream)) != hipSuccess) return def;
      if ((*err = hipEventSynchronize(ctx->tune_ev[1])) != hipSuccess) return def;
      float ms = 0.f;
      if ((*err = hipEventElapsedTime(&ms, ctx->tune_ev[0], ctx->tune_ev[1])) != hipSuccess)
        return def;
      if (ms < best_ms) {
        best_ms = ms;
        best = l;
      }
    }
  ctx->tuned[key] = best;
  return best;
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

// y[c] = sum_r x[r] A[r][c] (A is m x k, row-major), split over rows.
// Stage 1: block s reduces rows [s*rows_per, ...) for every column into
// part[s][c]: 64 column lanes x 4 row groups, fixed order, LDS tree.
__global__ __launch_bounds__(256) void gemv_t_partial_kernel(const float2* __restrict__ A,
                                                             const float2* __restrict__ x,
                                                             float2* __restrict__ part,
                                                             uint64_t M, uint64_t K,
                                                             uint64_t rows_per_split) {
  __shared__ float2 red[4][64];
  const uint64_t s = blockIdx.x;
  const int cl = threadIdx.x & 63;
  const int rg = threadIdx.x >> 6;
  const uint64_t r0 = s * rows_per_split;
  const uint64_t r1 = r0 + rows_per_split < M ? r0 + rows_per_split : M;
  for (uint64_t c0 = 0; c0 < K; c0 += 64) {
    const uint64_t c = c0 + cl;
    float2 acc = make_float2(0.f, 0.f);
    if (c < K)
      for (uint64_t r = r0 + rg; r < r1; r += 4) acc = cfma(x[r], A[r * K + c], acc);
    red[rg][cl] = acc;
    __syncthreads();
    if (rg == 0 && c < K) {
      float2 t = red[0][cl];
#pragma unroll
      for (int g = 1; g < 4; ++g) {
        t.x += red[g][cl].x;
        t.y += red[g][cl].y;
      }
      part[s * K + c] = t;
    }
    __syncthreads();
  }
}

// Stage 2 for few columns and many partials: one block per column, each
// thread sums a fixed strided subset, then a fixed LDS tree.
__global__ __launch_bounds__(256) void sum_partials_narrow_kernel(const float2* __restrict__ part,
                                                                  float2* __restrict__ y,
                                                                  uint64_t nsplit, uint64_t K) {
  __shared__ float2 red[256];
  const uint64_t c = blockIdx.x;
  float2 acc = make_float2(0.f, 0.f);
  for (uint64_t s = threadIdx.x; s < nsplit; s += 256) {
    const float2 v = part[s * K + c];
    acc.x += v.x;
    acc.y += v.y;
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) {
      red[threadIdx.x].x += red[threadIdx.x + w].x;
      red[threadIdx.x].y += red[threadIdx.x + w].y;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) y[c] = red[0];
}

// Stage 2: y[c] = sum_s part[s][c]; 64 columns x 4 split groups per block,
// grid-stride over column tiles.
__global__ __launch_bounds__(256) void sum_partials_kernel(const float2* __restrict__ part,
                                                           float2* __restrict__ y,
                                                           uint64_t nsplit, uint64_t K) {
  __shared__ float2 red[4][64];
  const int cl = threadIdx.x & 63;
  const int sg = threadIdx.x >> 6;
  for (uint64_t c0 = uint64_t(blockIdx.x) * 64; c0 < K; c0 += uint64_t(gridDim.x) * 64) {
    const uint64_t c = c0 + cl;
    float2 acc = make_float2(0.f, 0.f);
    if (c < K)
      for (uint64_t s = sg; s < nsplit; s += 4) {
        float2 v = part[s * K + c];
        acc.x += v.x;
        acc.y += v.y;
      }
    red[sg][cl] = acc;
    __syncthreads();
    if (sg == 0 && c < K) {
      float2 t = red[0][cl];
#pragma unroll
      for (int g = 1; g < 4; ++g) {
        t.x += red[g][cl].x;
        t.y += red[g][cl].y;
      }
      y[c] = t;
    }
    __syncthreads();
  }
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

// Few outputs, long k: C[m x n] with m*n <= 4096.  Block (x = k chunk,
// y = tile of 64 outputs): 4 groups of 64 lanes split the chunk, fixed
// order, LDS tree; part[chunk][o] is then summed over chunks by
// sum_partials_kernel.
__global__ __launch_bounds__(256) void cgemm_splitk_partial_kernel(
    const float2* __restrict__ A, const float2* __restrict__ B, float2* __restrict__ part,
    uint64_t M, uint64_t N, uint64_t K, uint64_t chunk) {
  __shared__ float2 red[4][64];
  const int ol = threadIdx.x & 63;
  const int kg = threadIdx.x >> 6;
  const uint64_t o = uint64_t(blockIdx.y) * 64 + ol;
  const uint64_t k0 = uint64_t(blockIdx.x) * chunk;
  const uint64_t k1 = k0 + chunk < K ? k0 + chunk : K;
  float2 acc = make_float2(0.f, 0.f);
  if (o < M * N) {
    const uint64_t m = o / N, n = o - m * N;
    const float2* a = A + m * K;
    for (uint64_t k = k0 + kg; k < k1; k += 4) acc = cfma(a[k], B[k * N + n], acc);
  }
  red[kg][ol] = acc;
  __syncthreads();
  if (kg == 0 && o < M * N) {
    float2 t = red[0][ol];
#pragma unroll
    for (int g = 1; g < 4; ++g) {
      t.x += red[g][ol].x;
      t.y += red[g][ol].y;
    }
    part[uint64_t(blockIdx.x) * M * N + o] = t;
  }
}

// Split-K for at most 8 x 8 outputs and long k (the final merges of a
// contraction, e.g. 8 x 8 x 2^20).  Each thread keeps all 64 outputs in
// registers over its k values (A rows and the B row of one k are
// contiguous loads across lanes), then a butterfly reduce-scatter leaves
// output o = i*8 + j on lane o of each wave; the 4 waves are summed in LDS
// in fixed order and sum_partials_kernel adds the blocks' partials in
// fixed order, so results are deterministic.
constexpr int kTinyChunk = 2048;  // k values per block (a power of two >= 256)
__global__ __launch_bounds__(256) void cgemm_tiny_mn_partial_kernel(
    const float2* __restrict__ A, const float2* __restrict__ B, float2* __restrict__ part,
    int M, int N, uint64_t K) {
  __shared__ float2 red[4][64];
  float2 acc[64];
#pragma unroll
  for (int o = 0; o < 64; ++o) acc[o] = make_float2(0.f, 0.f);
  const uint64_t k0 = uint64_t(blockIdx.x) * kTinyChunk;
  const uint64_t k1 = k0 + kTinyChunk < K ? k0 + kTinyChunk : K;
  for (uint64_t k = k0 + threadIdx.x; k < k1; k += 256) {
    float2 a[8], b[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = i < M ? A[uint64_t(i) * K + k] : make_float2(0.f, 0.f);
#pragma unroll
    for (int j = 0; j < 8; ++j) b[j] = j < N ? B[k * N + j] : make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[i * 8 + j] = cfma(a[i], b[j], acc[i * 8 + j]);
  }
  // reduce-scatter: after the step with distance s, a lane keeps the half
  // of its remaining outputs selected by (lane & s).
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int s = 32, half = 32; s > 0; s >>= 1, half >>= 1) {
    const bool upper = (lane & s) != 0;
#pragma unroll
    for (int o = 0; o < half; ++o) {
      float2 send = upper ? acc[o] : acc[o + half];
      float2 keep = upper ? acc[o + half] : acc[o];
      keep.x += __shfl_xor(send.x, s);
      keep.y += __shfl_xor(send.y, s);
      acc[o] = keep;
    }
  }
  red[threadIdx.x >> 6][lane] = acc[0];
  __syncthreads();
  if (threadIdx.x < 64) {
    float2 t = red[0][lane];
#pragma unroll
    for (int w = 1; w < 4; ++w) {
      t.x += red[w][lane].x;
      t.y += red[w][lane].y;
    }
    const int i = lane >> 3, j = lane & 7;
    if (i < M && j < N) part[uint64_t(blockIdx.x) * M * N + i * N + j] = t;
  }
}

// Small common dimension (k <= 16): exact FP32 on the vector ALUs, one
// fmaf chain per output in increasing k (the same operation order as
// cgemm_scalar_kernel, hence bit-identical to it).  Block tile BM x BN
// (BM = 4096 / BN), each thread 4 rows x 4 columns; A and B tiles staged
// through LDS with coalesced loads.
constexpr int kSmallK = 16;
// MAPPED: operands through digit maps (offset tables in LDS); A is read
// along m when a_mfast (a_rows digit 0 at stride 1), else along k; B along
// n when b_nfast, else along k.
// Split-K use (blockIdx.z = chunk > 0 or part != nullptr): the block covers
// k in [16*z, 16*z + K) of a common dimension of length ldk and writes its
// partial tile to part[z][m][n] (summed afterwards in fixed order).
// Max of v over a 256-thread block, written by thread 0 (the partial maxima
// a split-FP16 consumer of the output reduces instead of a max-abs pass).
__device__ __forceinline__ void store_block_max(float v, float* slot) {
  __shared__ float red[4];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) *slot = fmaxf(fmaxf(red[0], red[1]), fmaxf(red[2], red[3]));
}

template <int BN, bool MAPPED>
__global__ __launch_bounds__(256) void cgemm_smallk_kernel(
    const float2* __restrict__ A, const float2* __restrict__ B, float2* __restrict__ C,
    uint64_t M, int N, int K, qfx_index_map a_rows, qfx_index_map a_cols, qfx_index_map b_rows,
    qfx_index_map b_cols, bool a_mfast, bool b_nfast, uint64_t ldk, float2* __restrict__ part,
    float* __restrict__ cmax) {
  const uint64_t kbeg = uint64_t(blockIdx.z) * kSmallK;
  constexpr int TX = BN / 4;        // column groups
  constexpr int BM = 4 * 256 / TX;  // rows per block
  __shared__ float2 As[BM][kSmallK + 1];
  __shared__ float2 Bs[kSmallK][BN];
  __shared__ uint64_t arow[MAPPED ? BM : 1], bcol[MAPPED ? BN : 1];
  __shared__ uint64_t acol[MAPPED ? kSmallK : 1], brow[MAPPED ? kSmallK : 1];
  const int tx = threadIdx.x % TX, ty = threadIdx.x / TX;
  const uint64_t m0 = uint64_t(blockIdx.x) * BM;
  const int n0 = blockIdx.y * BN;
  if constexpr (MAPPED) {
    for (int i = threadIdx.x; i < BM; i += 256) arow[i] = map_offset(a_rows, m0 + i);
    for (int i = threadIdx.x; i < BN; i += 256) bcol[i] = map_offset(b_cols, uint64_t(n0 + i));
    if (threadIdx.x < K) {
      acol[threadIdx.x] = map_offset(a_cols, kbeg + threadIdx.x);
      brow[threadIdx.x] = map_offset(b_rows, kbeg + threadIdx.x);
    }
    __syncthreads();
  }
  // A tile: rows m0.., all K columns
  for (int e = threadIdx.x; e < BM * K; e += 256) {
    int r, c;
    if (MAPPED && a_mfast) {
      c = e / BM;
      r = e - c * BM;
    } else {
      r = e / K;
      c = e - r * K;
    }
    float2 v = make_float2(0.f, 0.f);
    if (m0 + r < M) v = MAPPED ? A[arow[r] + acol[c]] : A[(m0 + r) * ldk + kbeg + c];
    As[r][c] = v;
  }
  for (int e = threadIdx.x; e < K * BN; e += 256) {
    int r, c;
    if (MAPPED && !b_nfast) {
      c = e / K;
      r = e - c * K;
    } else {
      r = e / BN;
      c = e - r * BN;
    }
    float2 v = make_float2(0.f, 0.f);
    if (n0 + c < N) v = MAPPED ? B[brow[r] + bcol[c]] : B[(kbeg + r) * N + n0 + c];
    Bs[r][c] = v;
  }
  __syncthreads();
  float2 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = make_float2(0.f, 0.f);
  for (int kk = 0; kk < K; ++kk) {
    float2 a[4], b[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) a[i] = As[ty * 4 + i][kk];
#pragma unroll
    for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx + TX * j];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = cfma(a[i], b[j], acc[i][j]);
  }
  float2* out = part ? part + uint64_t(blockIdx.z) * M * N : C;
  float cm = 0.f;  // max |re|, |im| of the values this thread stores
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const uint64_t r = m0 + ty * 4 + i;
    if (r >= M) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = n0 + tx + TX * j;
      if (c < N) {
        out[r * N + c] = acc[i][j];
        cm = fmaxf(cm, fmaxf(fabsf(acc[i][j].x), fabsf(acc[i][j].y)));
      }
    }
  }
  if (cmax) store_block_max(cm, cmax + blockIdx.x + uint64_t(gridDim.x) * blockIdx.y);
}


// Mapped form of cgemm_tiny_mn_partial_kernel: A(i, k) = A[a_rows(i) +
// a_cols(k)], B(k, j) = B[b_rows(k) + b_cols(j)]; same order of operations.
__global__ __launch_bounds__(256) void cgemm_tiny_mn_mapped_partial_kernel(
    const float2* __restrict__ A, const float2* __restrict__ B, float2* __restrict__ part,
    int M, int N, uint64_t K, qfx_index_map a_rows, qfx_index_map a_cols, qfx_index_map b_rows,
    qfx_index_map b_cols) {
  __shared__ float2 red[4][64];
  uint64_t ar[8], bc[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) ar[i] = i < M ? map_offset(a_rows, i) : 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) bc[j] = j < N ? map_offset(b_cols, j) : 0;
  float2 acc[64];
#pragma unroll
  for (int o = 0; o < 64; ++o) acc[o] = make_float2(0.f, 0.f);
  const uint64_t k0 = uint64_t(blockIdx.x) * kTinyChunk;
  const uint64_t k1 = k0 + kTinyChunk < K ? k0 + kTinyChunk : K;
  // A digit map moves each index bit to its own offset bit, so indices with
  // disjoint bits have additive offsets: k0 (a multiple of kTinyChunk =
  // 2^11), the thread (< 2^8) and the step 256 it (< 2^11) are such.
  const uint64_t ak0 = map_offset(a_cols, k0 + threadIdx.x), bk0 = map_offset(b_rows, k0 + threadIdx.x);
  for (uint64_t k = k0 + threadIdx.x; k < k1; k += 256) {
    const uint64_t st = k - k0 - threadIdx.x;
    const uint64_t ak = ak0 + map_offset(a_cols, st), bk = bk0 + map_offset(b_rows, st);
    float2 a[8], b[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = i < M ? A[ar[i] + ak] : make_float2(0.f, 0.f);
#pragma unroll
    for (int j = 0; j < 8; ++j) b[j] = j < N ? B[bk + bc[j]] : make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < 8; ++i)
#pragma unroll
      for (int j = 0; j < 8; ++j) acc[i * 8 + j] = cfma(a[i], b[j], acc[i * 8 + j]);
  }
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int s = 32, half = 32; s > 0; s >>= 1, half >>= 1) {
    const bool upper = (lane & s) != 0;
#pragma unroll
    for (int o = 0; o < half; ++o) {
      float2 send = upper ? acc[o] : acc[o + half];
      float2 keep = upper ? acc[o + half] : acc[o];
      keep.x += __shfl_xor(send.x, s);
      keep.y += __shfl_xor(send.y, s);
      acc[o] = keep;
    }
  }
  red[threadIdx.x >> 6][lane] = acc[0];
  __syncthreads();
  if (threadIdx.x < 64) {
    float2 t = red[0][lane];
#pragma unroll
    for (int w = 1; w < 4; ++w) {
      t.x += red[w][lane].x;
      t.y += red[w][lane].y;
    }
    const int i = lane >> 3, j = lane & 7;
    if (i < M && j < N) part[uint64_t(blockIdx.x) * M * N + i * N + j] = t;
  }
}

// k <= 16 and n <= 16 (tall, narrow products such as 2^20 x 16 x 16): one
// thread per row of C.  B sits in LDS; the thread reads its row of A once;
// the block's rows of C (one contiguous range of 256 n elements) go through
// LDS so that the stores are contiguous across the wave.  Each output is the
// same fmaf chain in increasing k as in cgemm_smallk_kernel (bit-identical).
// AROW: each row of A holds its K (even) values contiguously and 16-byte
// aligned; the block's rows are then read into LDS as 16-byte pieces, K / 2
// consecutive lanes per row, instead of one 8-byte load per lane and k
// strided by the row pitch.
// The 256 rows of C from row 256 * block on (block = blockIdx.x of a single
// product, or the block's index inside its product in a grouped launch).
template <bool MAPPED, bool AROW>
__device__ __forceinline__ void rows_tile(const float2* __restrict__ A,
                                          const float2* __restrict__ B, float2* __restrict__ C,
                                          uint64_t M, int N, int K, const qfx_index_map& a_rows,
                                          const qfx_index_map& a_cols,
                                          const qfx_index_map& b_rows,
                                          const qfx_index_map& b_cols, float* __restrict__ cmax,
                                          uint64_t block) {
  constexpr int kAPitch = 18;  // 144-byte rows: 16-byte reads of 8 or 16 rows hit distinct banks
  constexpr int kCPitch = 17;
  __shared__ float2 Bs[kSmallK][16];
  __shared__ uint64_t acol[kSmallK];
  // The block's rows of A on their way in (AROW), then its rows of C on
  // their way out; a barrier separates the two uses.
  __shared__ __align__(16) float2 X[256 * kAPitch];
  const uint64_t m0 = block * 256;
  {
    const int r = threadIdx.x >> 4, c = threadIdx.x & 15;
    float2 v = make_float2(0.f, 0.f);
    if (r < K && c < N)
      v = MAPPED ? B[map_offset(b_rows, r) + map_offset(b_cols, c)] : B[r * N + c];
    Bs[r][c] = v;
    if (threadIdx.x < kSmallK)
      acol[threadIdx.x] = MAPPED && int(threadIdx.x) < K ? map_offset(a_cols, threadIdx.x)
                                                          : threadIdx.x;
  }
  if constexpr (AROW) {
    const int per_row = K >> 1;  // 16-byte pieces per row
    for (int q = threadIdx.x; q < 256 * per_row; q += 256) {
      const int r = q / per_row, p = q - r * per_row;
      const uint64_t grow = m0 + r;
      float4 v = make_float4(0.f, 0.f, 0.f, 0.f);
      if (grow < M) {
        const uint64_t off = MAPPED ? map_offset(a_rows, grow) : grow * uint64_t(K);
        v = *reinterpret_cast<const float4*>(A + off + 2 * p);
      }
      *reinterpret_cast<float4*>(&X[r * kAPitch + 2 * p]) = v;
    }
  }
  __syncthreads();
  const uint64_t row = m0 + threadIdx.x;
  float2 acc[16];
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = make_float2(0.f, 0.f);
  if (row < M) {
    if constexpr (AROW) {
#pragma unroll
      for (int kk = 0; kk < kSmallK; kk += 2) {
        if (kk < K) {
          const float4 a2 = *reinterpret_cast<const float4*>(&X[threadIdx.x * kAPitch + kk]);
          const float2 a0 = make_float2(a2.x, a2.y), a1 = make_float2(a2.z, a2.w);
#pragma unroll
          for (int j = 0; j < 16; ++j) acc[j] = cfma(a0, Bs[kk][j], acc[j]);
#pragma unroll
          for (int j = 0; j < 16; ++j) acc[j] = cfma(a1, Bs[kk + 1][j], acc[j]);
        }
      }
    } else {
      const uint64_t abase = MAPPED ? map_offset(a_rows, row) : row * uint64_t(K);
#pragma unroll
      for (int kk = 0; kk < kSmallK; ++kk) {
        if (kk < K) {
          const float2 a = A[abase + acol[kk]];
#pragma unroll
          for (int j = 0; j < 16; ++j) acc[j] = cfma(a, Bs[kk][j], acc[j]);
        }
      }
    }
  }
  if constexpr (AROW) __syncthreads();  // every row of A is read before X takes C
  float cm = 0.f;
  if (row < M) {
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      X[threadIdx.x * kCPitch + j] = acc[j];
      if (j < N) cm = fmaxf(cm, fmaxf(fabsf(acc[j].x), fabsf(acc[j].y)));
    }
  }
  __syncthreads();
  const uint64_t rows = M - m0 < 256 ? M - m0 : 256;
  const int total = int(rows) * N;
  float2* out = C + m0 * uint64_t(N);
  for (int e = threadIdx.x; e < total; e += 256) out[e] = X[(e / N) * kCPitch + e % N];
  if (cmax) store_block_max(cm, cmax + block);
}

template <bool MAPPED, bool AROW>
__global__ __launch_bounds__(256) void cgemm_rows_kernel(
    const float2* __restrict__ A, const float2* __restrict__ B, float2* __restrict__ C,
    uint64_t M, int N, int K, qfx_index_map a_rows, qfx_index_map a_cols, qfx_index_map b_rows,
    qfx_index_map b_cols, float* __restrict__ cmax) {
  rows_tile<MAPPED, AROW>(A, B, C, M, N, K, a_rows, a_cols, b_rows, b_cols, cmax, blockIdx.x);
}

// The row kernel for many 256-row tiles with A read one value per lane and
// k (rows_tile without AROW): each block walks the tiles blockIdx.x,
// blockIdx.x + gridDim.x, ... and loads the next tile's row of A into
// registers while it computes the current one.  With one tile per block the
// blocks on a CU load together and then compute together (16 x 16 complex
// fmaf chains per row, 27 us of vector work per 2^20 rows), so the loads and
// the arithmetic do not overlap.  Same fmaf chains as rows_tile
// (bit-identical); per-tile maxima in cmax[tile].
constexpr unsigned kRowsStreamBlocks = 1024;  // 4 blocks on each of 256 CUs

template <bool MAPPED>
__global__ __launch_bounds__(256, 4) void cgemm_rows_stream_kernel(
    const float2* __restrict__ A, const float2* __restrict__ B, float2* __restrict__ C,
    uint64_t M, int N, int K, qfx_index_map a_rows, qfx_index_map a_cols, qfx_index_map b_rows,
    qfx_index_map b_cols, float* __restrict__ cmax, uint64_t ntiles) {
  constexpr int kCPitch = 17;
  __shared__ float2 Bs[kSmallK][16];
  __shared__ uint64_t acol[kSmallK];
  __shared__ float2 X[256 * kCPitch];
  {
    const int r = threadIdx.x >> 4, c = threadIdx.x & 15;
    float2 v = make_float2(0.f, 0.f);
    if (r < K && c < N)
      v = MAPPED ? B[map_offset(b_rows, r) + map_offset(b_cols, c)] : B[r * N + c];
    Bs[r][c] = v;
    if (threadIdx.x < kSmallK)
      acol[threadIdx.x] = MAPPED && int(threadIdx.x) < K ? map_offset(a_cols, threadIdx.x)
                                                          : threadIdx.x;
  }
  __syncthreads();
  auto load_row = [&](uint64_t tile, float2* a) {
    const uint64_t row = tile * 256 + threadIdx.x;
#pragma unroll
    for (int kk = 0; kk < kSmallK; ++kk) a[kk] = make_float2(0.f, 0.f);
    if (row < M) {
      const uint64_t abase = MAPPED ? map_offset(a_rows, row) : row * uint64_t(K);
#pragma unroll
      for (int kk = 0; kk < kSmallK; ++kk)
        if (kk < K) a[kk] = A[abase + acol[kk]];
    }
  };
  float2 a_cur[kSmallK], a_next[kSmallK];
  load_row(blockIdx.x, a_cur);
  for (uint64_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
    if (t + gridDim.x < ntiles) load_row(t + gridDim.x, a_next);
    const uint64_t m0 = t * 256, row = m0 + threadIdx.x;
    float2 acc[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) acc[j] = make_float2(0.f, 0.f);
    float cm = 0.f;
    if (row < M) {
#pragma unroll
      for (int kk = 0; kk < kSmallK; ++kk)
        if (kk < K) {
#pragma unroll
          for (int j = 0; j < 16; ++j) acc[j] = cfma(a_cur[kk], Bs[kk][j], acc[j]);
        }
#pragma unroll
      for (int j = 0; j < 16; ++j) {
        X[threadIdx.x * kCPitch + j] = acc[j];
        if (j < N) cm = fmaxf(cm, fmaxf(fabsf(acc[j].x), fabsf(acc[j].y)));
      }
    }
    __syncthreads();
    const uint64_t rows = M - m0 < 256 ? M - m0 : 256;
    const int total = int(rows) * N;
    float2* out = C + m0 * uint64_t(N);
    for (int e = threadIdx.x; e < total; e += 256) out[e] = X[(e / N) * kCPitch + e % N];
    if (cmax) store_block_max(cm, cmax + t);
    __syncthreads();  // X and the maximum's scratch are rewritten by the next tile
#pragma unroll
    for (int kk = 0; kk < kSmallK; ++kk) a_cur[kk] = a_next[kk];
  }
}

// One problem of a grouped launch of the row kernel (row-major operands).
struct RowsProb {
  const float2* A;
  const float2* B;
  float2* C;
  uint64_t M;
  int N, K;
  uint64_t first_block;  // index of the problem's first block in the launch
};

// Grouped row kernel: block b runs block b - first_block of the last
// problem whose first block is <= b.
__global__ __launch_bounds__(256) void cgemm_rows_grouped_kernel(const RowsProb* __restrict__ probs,
                                                                 int count) {
  const uint64_t b = blockIdx.x;
  int lo = 0, hi = count - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (probs[mid].first_block <= b) lo = mid;
    else hi = mid - 1;
  }
  const RowsProb p = probs[lo];
  const qfx_index_map none{};
  rows_tile<false, false>(p.A, p.B, p.C, p.M, p.N, p.K, none, none, none, none, nullptr,
                          b - p.first_block);
}

unsigned cap_grid(uint64_t g) { return unsigned(std::max<uint64_t>(1, std::min<uint64_t>(g, 65535))); }

}  // namespace

// Small-output GEMMs with a longer common dimension (m*n <= 2^16,
// 16 < k <= 4096, k a multiple of 16): split-K over chunks of 16 with the
// vector-ALU tile kernel, partial tiles summed in fixed chunk order.
// Enough blocks to fill the chip where one MFMA tile per output block
// would leave most CUs idle.
static bool smallk_splitk_fits(uint64_t m, uint64_t n, uint64_t k) {
  return m * n <= (uint64_t(1) << 16) && k > uint64_t(kSmallK) && k <= 4096 &&
         k % kSmallK == 0;
}

static int launch_smallk_splitk(qfx_context* ctx, const float2* A, const float2* B, float2* C,
                                uint64_t m, uint64_t n, uint64_t k, bool mapped,
                                const qfx_index_map* maps, bool a_mfast, bool b_nfast) {
  const uint64_t nchunk = k / kSmallK;
  void* ws = nullptr;
  int st = qfx_workspace(ctx, nchunk * m * n * sizeof(float2), &ws);
  if (st != QFX_OK) return st;
  float2* part = static_cast<float2*>(ws);
  const qfx_index_map none{};
  const qfx_index_map& ar = mapped ? maps[0] : none;
  const qfx_index_map& ac = mapped ? maps[1] : none;
  const qfx_index_map& br = mapped ? maps[2] : none;
  const qfx_index_map& bc = mapped ? maps[3] : none;
  if (n <= 16) {
    dim3 g{unsigned((m + 255) / 256), unsigned((n + 15) / 16), unsigned(nchunk)};
    if (mapped)
      cgemm_smallk_kernel<16, true><<<g, 256, 0, ctx->stream>>>(
          A, B, C, m, int(n), kSmallK, ar, ac, br, bc, a_mfast, b_nfast, k, part, nullptr);
    else
      cgemm_smallk_kernel<16, false><<<g, 256, 0, ctx->stream>>>(
          A, B, C, m, int(n), kSmallK, ar, ac, br, bc, false, true, k, part, nullptr);
  } else {
    dim3 g{unsigned((m + 63) / 64), unsigned((n + 63) / 64), unsigned(nchunk)};
    if (mapped)
      cgemm_smallk_kernel<64, true><<<g, 256, 0, ctx->stream>>>(
          A, B, C, m, int(n), kSmallK, ar, ac, br, bc, a_mfast, b_nfast, k, part, nullptr);
    else
      cgemm_smallk_kernel<64, false><<<g, 256, 0, ctx->stream>>>(
          A, B, C, m, int(n), kSmallK, ar, ac, br, bc, false, true, k, part, nullptr);
  }
  sum_partials_kernel<<<cap_grid((m * n + 63) / 64), 256, 0, ctx->stream>>>(part, C, nchunk,
                                                                           m * n);
  QFX_HIP(hipGetLastError());
  return QFX_OK;
}

static int launch_gemv_t(qfx_context* ctx, uint64_t m, uint64_t k, const float2* A,
                         const float2* x, float2* y) {
  // About 1024 row blocks of at least 16 rows each.
  uint64_t nsplit = std::max<uint64_t>(1, std::min<uint64_t>(1024, (m + 15) / 16));
  uint64_t rows_per = (m + nsplit - 1) / nsplit;
  nsplit = (m + rows_per - 1) / rows_per;
  void* ws = nullptr;
  int st = qfx_workspace(ctx, nsplit * k * sizeof(float2), &ws);
  if (st != QFX_OK) return st;
  float2* part = static_cast<float2*>(ws);
  gemv_t_partial_kernel<<<unsigned(nsplit), 256, 0, ctx->stream>>>(A, x, part, m, k, rows_per);
  sum_partials_kernel<<<cap_grid((k + 63) / 64), 256, 0, ctx->stream>>>(part, y, nsplit, k);
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
  QFX_EXTENT(ctx, A, 8.0 * double(m) * double(k), "qfx_cgemv_c64: A");
  QFX_EXTENT(ctx, x, 8.0 * double(trans ? m : k), "qfx_cgemv_c64: x");
  QFX_EXTENT(ctx, y, 8.0 * double(ny), "qfx_cgemv_c64: y");
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
  QFX_EXTENT(ctx, x, 8.0 * double(n), "qfx_cdotu_c64: x");
  QFX_EXTENT(ctx, y, 8.0 * double(n), "qfx_cdotu_c64: y");
  QFX_EXTENT(ctx, out, 8.0, "qfx_cdotu_c64: out");
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

// Operand scales of the split-FP16 form: partial maxima of |A| and |B| in
// one launch (at most 256 blocks per operand, at least 16 Ki elements
// each), skipping an operand whose maxima its producer already wrote; the
// GEMM kernel reduces them (no host round trip, no atomics).
static int launch_amax(qfx_context* ctx, const float2* A, uint64_t na, const float2* B,
                       uint64_t nb, AmaxRef* ref, const float* a_known = nullptr, int a_n = 0,
                       const float* b_known = nullptr, int b_n = 0) {
  const int ba = a_known ? 0 : int(std::min<uint64_t>(256, std::max<uint64_t>(1, (na + 16383) / 16384)));
  const int bb = b_known ? 0 : int(std::min<uint64_t>(256, std::max<uint64_t>(1, (nb + 16383) / 16384)));
  float* part = nullptr;
  if (ba + bb > 0) {
    void* ws = nullptr;
    int st = qfx_workspace(ctx, size_t(ba + bb) * sizeof(float), &ws);
    if (st != QFX_OK) return st;
    part = static_cast<float*>(ws);
    amax_partial_kernel<<<unsigned(ba + bb), 256, 0, ctx->stream>>>(A, na, B, nb, ba, part);
  }
  *ref = AmaxRef{a_known ? a_known : part, a_known ? a_n : ba, b_known ? b_known : part + ba,
                 b_known ? b_n : bb};
  return QFX_OK;
}

extern "C" int qfx_amax_partials_c64(qfx_context* ctx, const qfx_c64* X, uint64_t n,
                                     float* out, int cap, int* count) {
  QFX_CHECK_CTX(ctx);
  if (count) *count = 0;
  if (n == 0) return QFX_OK;
  if (!X || !out) return qfx_set_error(QFX_ERR_INVALID, "qfx_amax_partials_c64: null pointer");
  const int b = int(std::min<uint64_t>(256, std::max<uint64_t>(1, (n + 16383) / 16384)));
  if (b > cap)
    return qfx_set_error(QFX_ERR_INVALID, "qfx_amax_partials_c64: " + std::to_string(b) +
                                              " partial maxima do not fit in " +
                                              std::to_string(cap));
  QfxLaunchScope scope(ctx, QFX_K_COPY, 8.0 * n, 0);
  amax_partial_kernel<<<unsigned(b), 256, 0, ctx->stream>>>(
      reinterpret_cast<const float2*>(X), n, nullptr, 0, b, out);
  QFX_HIP(hipGetLastError());
  if (count) *count = b;
  return QFX_OK;
}

extern "C" int qfx_cgemm_c64(qfx_context* ctx, int precision, uint64_t m, uint64_t n,
                             uint64_t k, const qfx_c64* A, const qfx_c64* B, qfx_c64* C) {
  QFX_CHECK_CTX(ctx);
  if (precision < QFX_PREC_FP32 || precision > QFX_PREC_FP8)
    return qfx_set_error(QFX_ERR_INVALID, "qfx_cgemm_c64: unknown precision");
  if (m == 0 || n == 0) return QFX_OK;
  if (!C) return qfx_set_error(QFX_ERR_INVALID, "qfx_cgemm_c64: null C");
  if (k == 0) return qfx_zero(ctx, C, m * n * sizeof(qfx_c64));
  if (!A || !B) return qfx_set_error(QFX_ERR_INVALID, "qfx_cgemm_c64: null A or B");
  QFX_EXTENT(ctx, A, 8.0 * double(m) * double(k), "qfx_cgemm_c64: A");
  QFX_EXTENT(ctx, B, 8.0 * double(k) * double(n), "qfx_cgemm_c64: B");
  QFX_EXTENT(ctx, C, 8.0 * double(m) * double(n), "qfx_cgemm_c64: C");
  const float2* A2 = reinterpret_cast<const float2*>(A);
  const float2* B2 = reinterpret_cast<const float2*>(B);
  float2* C2 = reinterpret_cast<float2*>(C);
  const double flops = 8.0 * double(m) * double(n) * double(k);
  const double bytes = 8.0 * (double(m) * k + double(k) * n + double(m) * n);

  if (k <= uint64_t(kSmallK) && n <= 65535ULL * 64) {
    // Small common dimension: exact FP32 on the vector ALUs for every
    // precision (bit-identical to the defined-order scalar kernel).
    QfxLaunchScope scope(ctx, QFX_K_CGEMM_SMALL, bytes, flops);
    if (n <= 16 && (m + 255) / 256 <= 0x7fffffffULL) {
      const unsigned blocks = unsigned((m + 255) / 256);
      if (k % 2 == 0 && reinterpret_cast<uintptr_t>(A) % 16 == 0)
        cgemm_rows_kernel<false, true><<<blocks, 256, 0, ctx->stream>>>(
            A2, B2, C2, m, int(n), int(k), {}, {}, {}, {}, nullptr);
      else if (blocks >= 2 * kRowsStreamBlocks)
        cgemm_rows_stream_kernel<false><<<kRowsStreamBlocks, 256, 0, ctx->stream>>>(
            A2, B2, C2, m, int(n), int(k), {}, {}, {}, {}, nullptr, blocks);
      else
        cgemm_rows_kernel<false, false><<<blocks, 256, 0, ctx->stream>>>(
            A2, B2, C2, m, int(n), int(k), {}, {}, {}, {}, nullptr);
    } else if (n <= 16) {
      dim3 g{unsigned((m + 255) / 256), unsigned((n + 15) / 16), 1};
      cgemm_smallk_kernel<16, false><<<g, 256, 0, ctx->stream>>>(A2, B2, C2, m, int(n), int(k),
                                                                 {}, {}, {}, {}, false, true, k,
                                                                 nullptr, nullptr);
    } else {
      dim3 g{unsigned((m + 63) / 64), unsigned((n + 63) / 64), 1};
      cgemm_smallk_kernel<64, false><<<g, 256, 0, ctx->stream>>>(A2, B2, C2, m, int(n), int(k),
                                                                 {}, {}, {}, {}, false, true, k,
                                                                 nullptr, nullptr);
    }
    QFX_HIP(hipGetLastError());
    return QFX_OK;
  }
  if (precision != QFX_PREC_FP32_EXACT && m > 8 && n > 1 && smallk_splitk_fits(m, n, k)) {
    QfxLaunchScope scope(ctx, QFX_K_CGEMM_SMALL, bytes, flops);
    return launch_smallk_splitk(ctx, A2, B2, C2, m, n, k, false, nullptr, false, true);
  }
  if (precision != QFX_PREC_FP32_EXACT && m <= 8 && n <= 8 && k >= 1024) {
    QfxLaunchScope scope(ctx, QFX_K_GEMV, bytes, flops);
    const uint64_t nchunk = (k + kTinyChunk - 1) / kTinyChunk;
    void* ws = nullptr;
    int st = qfx_workspace(ctx, nchunk * m * n * sizeof(float2), &ws);
    if (st != QFX_OK) return st;
    float2* part = static_cast<float2*>(ws);
    cgemm_tiny_mn_partial_kernel<<<unsigned(nchunk), 256, 0, ctx->stream>>>(A2, B2, part, int(m),
                                                                            int(n), k);
    sum_partials_narrow_kernel<<<unsigned(m * n), 256, 0, ctx->stream>>>(part, C2, nchunk, m * n);
    QFX_HIP(hipGetLastError());
    return QFX_OK;
  }
  if (precision == QFX_PREC_FP8 && qfx_fp8_applies(m, n, k)) {
    QfxLaunchScope scope(ctx, QFX_K_CGEMM, bytes, flops);
    return qfx_cgemm_fp8(ctx, A2, B2, C2, m, n, k, nullptr);
  }
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
    if (precision != QFX_PREC_FP32_EXACT && m * n <= 4096 && k >= 1024) {
      QfxLaunchScope scope(ctx, QFX_K_GEMV, bytes, flops);
      const uint64_t outs = m * n;
      uint64_t nchunk = std::max<uint64_t>(1, std::min<uint64_t>(2048 / ((outs + 63) / 64), k / 256));
      const uint64_t chunk = (k + nchunk - 1) / nchunk;
      nchunk = (k + chunk - 1) / chunk;
      void* ws = nullptr;
      int st = qfx_workspace(ctx, nchunk * outs * sizeof(float2), &ws);
      if (st != QFX_OK) return st;
      float2* part = static_cast<float2*>(ws);
      dim3 g{unsigned(nchunk), unsigned((outs + 63) / 64), 1};
      cgemm_splitk_partial_kernel<<<g, 256, 0, ctx->stream>>>(A2, B2, part, m, n, k, chunk);
      sum_partials_kernel<<<cap_grid((outs + 63) / 64), 256, 0, ctx->stream>>>(part, C2, nchunk,
                                                                            outs);
      QFX_HIP(hipGetLastError());
      return QFX_OK;
    }
    QfxLaunchScope scope(ctx, QFX_K_CGEMM_SMALL, bytes, flops);
    cgemm_scalar_kernel<<<cap_grid((m * n + 255) / 256), 256, 0, ctx->stream>>>(A2, B2, C2, m, n,
                                                                               k);
    QFX_HIP(hipGetLastError());
    return QFX_OK;
  }
  QfxLaunchScope scope(ctx, QFX_K_CGEMM, bytes, flops);
  dim3 grid(unsigned((m + kBM - 1) / kBM), unsigned((n + kBN - 1) / kBN));
  if (grid.y > 65535) return qfx_set_error(QFX_ERR_UNSUPPORTED, "qfx_cgemm_c64: n too large");
  // Large tiles (one block per CU) when there are enough of them to fill
  // the chip; float4 operand loads need even k and n and 16-byte alignment.
  const BigLayout layout =
      (precision != QFX_PREC_FP32 && precision != QFX_PREC_FP8 && (k % 2 == 0) && (n % 2 == 0) &&
       big_operands_fit(m, n, k) &&
       (reinterpret_cast<uintptr_t>(A) % 16 == 0) && (reinterpret_cast<uintptr_t>(B) % 16 == 0))
          ? choose_big_layout(precision, m, n)
          : kLayoutNone;
  AmaxRef amax{nullptr, 0, nullptr, 0};
  if (precision == QFX_PREC_SPLIT_FP16) {
    int st = launch_amax(ctx, A2, m * k, B2, k * n, &amax);
    if (st != QFX_OK) return st;
  }
  if (layout != kLayoutNone) {
    if ((n + 31) / 32 > 65535) return qfx_set_error(QFX_ERR_UNSUPPORTED, "qfx_cgemm_c64: n too large");
    hipError_t te = hipSuccess;
    const bool wide = !big_operands_whole(m, n, k);
    const BigLayout use = tuned_layout(
        ctx, precision, 0, m, n, k, layout,
        [&](BigLayout l) {
          return launch_split_big_prec<0>(precision, l, ctx->stream, A2, B2, C2, m, n, k, amax,
                                          nullptr, nullptr, wide);
        },
        &te);
    QFX_HIP(te);
    QFX_HIP(launch_split_big_prec<0>(precision, use, ctx->stream, A2, B2, C2, m, n, k, amax,
                                     nullptr, nullptr, wide));
    return QFX_OK;
  }
  if (precision == QFX_PREC_SPLIT_FP16) {
    cgemm_fp16x2_mfma_kernel<<<grid, 256, 0, ctx->stream>>>(A2, B2, C2, int(m), int(n), int(k),
                                                            amax);
    QFX_HIP(hipGetLastError());
    return QFX_OK;
  }
  // FP8 below its tile sizes runs the FP32 MFMA kernel.
  if (precision == QFX_PREC_FP32 || precision == QFX_PREC_FP8)
    cgemm_f32_mfma_kernel<<<grid, 256, 0, ctx->stream>>>(A2, B2, C2, int(m), int(n), int(k));
  else if (precision == QFX_PREC_SPLIT_BF16)
    cgemm_bf16_mfma_kernel<3><<<grid, 256, 0, ctx->stream>>>(A2, B2, C2, int(m), int(n), int(k));
  else
    cgemm_bf16_mfma_kernel<6><<<grid, 256, 0, ctx->stream>>>(A2, B2, C2, int(m), int(n), int(k));
  QFX_HIP(hipGetLastError());
  return QFX_OK;
}

// The same product with no device allocation: every kernel's scratch
// (split-K partial sums, operand maxima, FP8 codes) comes from the fixed
// scratch array of the context's device, the large tiles follow the fixed
// rule (no autotuner timing), and a product that needs more scratch than the
// array holds is refused before anything is launched for it.
extern "C" int qfx_cgemm_noalloc_c64(qfx_context* ctx, int precision, uint64_t m, uint64_t n,
                                     uint64_t k, const qfx_c64* A, const qfx_c64* B,
                                     qfx_c64* C) {
  QFX_CHECK_CTX(ctx);
  if (!ctx->fixed_scratch)
    return qfx_set_error(QFX_ERR_INVALID, "qfx_cgemm_noalloc_c64: no fixed scratch on this device");
  struct Restore {
    qfx_context* c;
    bool autotune;
    ~Restore() {
      c->use_fixed_scratch = false;
      c->autotune = autotune;
    }
  } restore{ctx, ctx->autotune};
  ctx->use_fixed_scratch = true;
  ctx->autotune = false;
  return qfx_cgemm_c64(ctx, precision, m, n, k, A, B, C);
}

// ------------------------------------------------------------ mapped operands

static bool map_size(const qfx_index_map* d, uint64_t* size) {
  if (!d || d->ndigits < 0 || d->ndigits > 16) return false;
  int total = 0;
  for (int j = 0; j < d->ndigits; ++j) {
    if (d->bits[j] == 0 || d->shift[j] > 47) return false;
    total += d->bits[j];
  }
  if (total > 47) return false;
  *size = uint64_t(1) << total;
  return true;
}

// Offset of the map's last index (every digit at its largest value).
static uint64_t map_last(const qfx_index_map& d) {
  uint64_t off = 0;
  for (int j = 0; j < d.ndigits; ++j) off += ((uint64_t(1) << d.bits[j]) - 1) << d.shift[j];
  return off;
}

// The map is the identity: its digits in order from stride 1, index i at
// offset i.
static bool identity_map(const qfx_index_map& d) {
  int at = 0;
  for (int j = 0; j < d.ndigits; ++j) {
    if (d.shift[j] != at) return false;
    at += d.bits[j];
  }
  return true;
}

// Some digit of the map sits at stride 2^shift.
static bool has_digit_at(const qfx_index_map& d, int shift) {
  for (int j = 0; j < d.ndigits; ++j)
    if (d.shift[j] == shift) return true;
  return false;
}

// Digit 0 at stride 1 with at least `bits` bits: runs of 2^bits contiguous
// elements along this matrix dimension.
static bool contiguous(const qfx_index_map& d, int bits) {
  return d.ndigits > 0 && d.shift[0] == 0 && d.bits[0] >= bits;
}

extern "C" int qfx_cgemm_mapped_c64(qfx_context* ctx, int precision, const qfx_c64* A,
                                    const qfx_index_map* a_rows, const qfx_index_map* a_cols,
                                    const qfx_c64* B, const qfx_index_map* b_rows,
                                    const qfx_index_map* b_cols, qfx_c64* C) {
  return qfx_cgemm_mapped_ex_c64(ctx, precision, A, a_rows, a_cols, B, b_rows, b_cols, C,
                                 nullptr);
}

extern "C" int qfx_cgemm_mapped_ex_c64(qfx_context* ctx, int precision, const qfx_c64* A,
                                       const qfx_index_map* a_rows,
                                       const qfx_index_map* a_cols, const qfx_c64* B,
                                       const qfx_index_map* b_rows,
                                       const qfx_index_map* b_cols, qfx_c64* C,
                                       qfx_scale_io* io) {
  QFX_CHECK_CTX(ctx);
  if (io) io->c_max_n = 0;
  if (precision < QFX_PREC_FP32 || precision > QFX_PREC_FP8)
    return qfx_set_error(QFX_ERR_INVALID, "qfx_cgemm_mapped_c64: unknown precision");
  uint64_t m, k, k2, n;
  if (!map_size(a_rows, &m) || !map_size(a_cols, &k) || !map_size(b_rows, &k2) ||
      !map_size(b_cols, &n))
    return qfx_set_error(QFX_ERR_INVALID, "qfx_cgemm_mapped_c64: invalid index map");
  if (k != k2)
    return qfx_set_error(QFX_ERR_INVALID, "qfx_cgemm_mapped_c64: a_cols and b_rows differ in size");
  if (!A || !B || !C) return qfx_set_error(QFX_ERR_INVALID, "qfx_cgemm_mapped_c64: null pointer");
  QFX_EXTENT(ctx, A, 8.0 * double(map_last(*a_rows) + map_last(*a_cols) + 1),
             "qfx_cgemm_mapped_c64: A");
  QFX_EXTENT(ctx, B, 8.0 * double(map_last(*b_rows) + map_last(*b_cols) + 1),
             "qfx_cgemm_mapped_c64: B");
  QFX_EXTENT(ctx, C, 8.0 * double(m) * double(n), "qfx_cgemm_mapped_c64: C");
  const float2* A2 = reinterpret_cast<const float2*>(A);
  const float2* B2 = reinterpret_cast<const float2*>(B);
  float2* C2 = reinterpret_cast<float2*>(C);
  const double flops = 8.0 * double(m) * double(n) * double(k);
  const double bytes = 8.0 * (double(m) * k + double(k) * n + double(m) * n);

  if (k <= uint64_t(kSmallK) && n <= 65535ULL * 64) {
    QfxLaunchScope scope(ctx, QFX_K_CGEMM_SMALL, bytes, flops);
    const bool a_mfast = contiguous(*a_rows, 1) && !contiguous(*a_cols, 3);
    const bool b_nfast = contiguous(*b_cols, 1) || !contiguous(*b_rows, 1);
    // per-block maxima of C for a split-FP16 product that reads it next
    auto cmax_for = [&](uint64_t blocks) -> float* {
      if (precision != QFX_PREC_SPLIT_FP16 || !io || !io->c_max ||
          blocks > uint64_t(std::max(io->c_max_cap, 0)))
        return nullptr;
      io->c_max_n = int(blocks);
      return io->c_max;
    };
    if (n <= 16 && (m + 255) / 256 <= 0x7fffffffULL) {
      const uint64_t blocks = (m + 255) / 256;
      // rows of A contiguous along k (A's k digits in order from stride 1)
      // and starting at even offsets (16-byte aligned)
      if (k % 2 == 0 && identity_map(*a_cols) && !has_digit_at(*a_rows, 0) &&
          reinterpret_cast<uintptr_t>(A) % 16 == 0)
        cgemm_rows_kernel<true, true><<<unsigned(blocks), 256, 0, ctx->stream>>>(
            A2, B2, C2, m, int(n), int(k), *a_rows, *a_cols, *b_rows, *b_cols, cmax_for(blocks));
      else if (blocks >= 2 * kRowsStreamBlocks)
        cgemm_rows_stream_kernel<true><<<kRowsStreamBlocks, 256, 0, ctx->stream>>>(
            A2, B2, C2, m, int(n), int(k), *a_rows, *a_cols, *b_rows, *b_cols, cmax_for(blocks),
            blocks);
      else
        cgemm_rows_kernel<true, false><<<unsigned(blocks), 256, 0, ctx->stream>>>(
            A2, B2, C2, m, int(n), int(k), *a_rows, *a_cols, *b_rows, *b_cols, cmax_for(blocks));
    } else if (n <= 16) {
      dim3 g{unsigned((m + 255) / 256), unsigned((n + 15) / 16), 1};
      cgemm_smallk_kernel<16, true><<<g, 256, 0, ctx->stream>>>(
          A2, B2, C2, m, int(n), int(k), *a_rows, *a_cols, *b_rows, *b_cols, a_mfast, b_nfast, k,
          nullptr, cmax_for(uint64_t(g.x) * g.y));
    } else {
      dim3 g{unsigned((m + 63) / 64), unsigned((n + 63) / 64), 1};
      cgemm_smallk_kernel<64, true><<<g, 256, 0, ctx->stream>>>(
          A2, B2, C2, m, int(n), int(k), *a_rows, *a_cols, *b_rows, *b_cols, a_mfast, b_nfast, k,
          nullptr, cmax_for(uint64_t(g.x) * g.y));
    }
    QFX_HIP(hipGetLastError());
    return QFX_OK;
  }
  if (precision != QFX_PREC_FP32_EXACT && m > 8 && n > 1 && smallk_splitk_fits(m, n, k)) {
    QfxLaunchScope scope(ctx, QFX_K_CGEMM_SMALL, bytes, flops);
    const qfx_index_map maps[4] = {*a_rows, *a_cols, *b_rows, *b_cols};
    const bool a_mfast = contiguous(*a_rows, 1) && !contiguous(*a_cols, 3);
    const bool b_nfast = contiguous(*b_cols, 1) || !contiguous(*b_rows, 1);
    return launch_smallk_splitk(ctx, A2, B2, C2, m, n, k, true, maps, a_mfast, b_nfast);
  }
  if (precision != QFX_PREC_FP32_EXACT && m <= 8 && n <= 8 && k >= 1024) {
    // Consecutive k go to consecutive lanes: B must be near-contiguous along
    // k (one row of B within 128 bytes), else the loads are scattered; the
    // caller then reorders B.
    if (b_rows->ndigits == 0 || b_rows->shift[0] > 4)
      return qfx_set_error(QFX_ERR_UNSUPPORTED, "qfx_cgemm_mapped_c64: B is not read along k");
    QfxLaunchScope scope(ctx, QFX_K_GEMV, bytes, flops);
    const uint64_t nchunk = (k + kTinyChunk - 1) / kTinyChunk;
    void* ws = nullptr;
    int st = qfx_workspace(ctx, nchunk * m * n * sizeof(float2), &ws);
    if (st != QFX_OK) return st;
    float2* part = static_cast<float2*>(ws);
    cgemm_tiny_mn_mapped_partial_kernel<<<unsigned(nchunk), 256, 0, ctx->stream>>>(
        A2, B2, part, int(m), int(n), k, *a_rows, *a_cols, *b_rows, *b_cols);
    sum_partials_narrow_kernel<<<unsigned(m * n), 256, 0, ctx->stream>>>(part, C2, nchunk, m * n);
    QFX_HIP(hipGetLastError());
    return QFX_OK;
  }
  if (precision == QFX_PREC_FP8) {
    // The quantize kernels read the operands through the maps: every form.
    if (!qfx_fp8_applies(m, n, k))
      return qfx_set_error(QFX_ERR_UNSUPPORTED, "qfx_cgemm_mapped_c64: no kernel reads this form");
    QfxLaunchScope scope(ctx, QFX_K_CGEMM, bytes, flops);
    const qfx_index_map maps[4] = {*a_rows, *a_cols, *b_rows, *b_cols};
    return qfx_cgemm_fp8(ctx, A2, B2, C2, m, n, k, maps);
  }
  const bool split = precision == QFX_PREC_SPLIT_FP16 || precision == QFX_PREC_SPLIT_BF16 ||
                     precision == QFX_PREC_SPLIT_BF16X3;
  const BigLayout layout =
      (split && big_mapped_spans_fit(*a_rows, *a_cols, *b_rows, *b_cols, m, n, k) && k >= 8 &&
       reinterpret_cast<uintptr_t>(A) % 16 == 0 && reinterpret_cast<uintptr_t>(B) % 16 == 0)
          ? choose_big_layout(precision, m, n)
          : kLayoutNone;
  const int am = contiguous(*a_cols, 3) ? 1 : (contiguous(*a_rows, 3) ? 2 : 0);
  if (layout == kLayoutNone || am == 0 || !contiguous(*b_cols, 3) || (n + 31) / 32 > 65535 ||
      !split_big_mapped_fits(precision, layout, k))
    return qfx_set_error(QFX_ERR_UNSUPPORTED, "qfx_cgemm_mapped_c64: no kernel reads this form");
  QfxLaunchScope scope(ctx, QFX_K_CGEMM, bytes, flops);
  AmaxRef amax{nullptr, 0, nullptr, 0};
  float* cmax = nullptr;
  if (precision == QFX_PREC_SPLIT_FP16) {
    const bool a_known = io && io->a_max && io->a_max_n > 0;
    const bool b_known = io && io->b_max && io->b_max_n > 0;
    int st = launch_amax(ctx, A2, m * k, B2, k * n, &amax, a_known ? io->a_max : nullptr,
                         a_known ? io->a_max_n : 0, b_known ? io->b_max : nullptr,
                         b_known ? io->b_max_n : 0);
    if (st != QFX_OK) return st;
  }
  const qfx_index_map maps[4] = {*a_rows, *a_cols, *b_rows, *b_cols};
  // per-block maxima of C for the next product, when the caller's array holds them
  auto cmax_for = [&](BigLayout l) -> float* {
    return (precision == QFX_PREC_SPLIT_FP16 && io && io->c_max &&
            big_layout_blocks(l, m, n) <= uint64_t(std::max(io->c_max_cap, 0)))
               ? io->c_max
               : nullptr;
  };
  const bool wide = !big_mapped_whole(*a_rows, *a_cols, *b_rows, *b_cols, m, n, k);
  auto launch = [&](BigLayout l) {
    return am == 1 ? launch_split_big_prec<1>(precision, l, ctx->stream, A2, B2, C2, m, n, k,
                                              amax, maps, cmax_for(l), wide)
                   : launch_split_big_prec<2>(precision, l, ctx->stream, A2, B2, C2, m, n, k,
                                              amax, maps, cmax_for(l), wide);
  };
  hipError_t te = hipSuccess;
  const BigLayout use = tuned_layout(ctx, precision, am, m, n, k, layout, launch, &te);
  QFX_HIP(te);
  cmax = cmax_for(use);
  if (cmax) io->c_max_n = int(big_layout_blocks(use, m, n));
  QFX_HIP(launch(use));
  return QFX_OK;
}

// ------------------------------------------------------------ grouped products

// The row-kernel problems of a grouped call in one launch.  The problem
// list goes to the device through the staged uploads (flushed right before
// the launch) into the workspace.
static int launch_rows_grouped(qfx_context* ctx, const qfx_gemm_desc* descs,
                               const std::vector<int>& rows) {
  std::vector<RowsProb> probs(rows.size());
  uint64_t blocks = 0;
  double flops = 0, bytes = 0;
  for (size_t r = 0; r < rows.size(); ++r) {
    const qfx_gemm_desc& d = descs[rows[r]];
    probs[r] = RowsProb{reinterpret_cast<const float2*>(d.A), reinterpret_cast<const float2*>(d.B),
                        reinterpret_cast<float2*>(d.C), d.m, int(d.n), int(d.k), blocks};
    blocks += (d.m + 255) / 256;
    flops += 8.0 * double(d.m) * double(d.n) * double(d.k);
    bytes += 8.0 * (double(d.m) * d.k + double(d.k) * d.n + double(d.m) * d.n);
  }
  if (blocks > 0x7fffffffULL)
    return qfx_set_error(QFX_ERR_UNSUPPORTED, "qfx_cgemm_grouped_c64: too many row blocks");
  void* ws = nullptr;
  int st = qfx_workspace(ctx, probs.size() * sizeof(RowsProb), &ws);
  if (st != QFX_OK) return st;
  st = qfx_copy_h2d_async(ctx, ws, probs.data(), probs.size() * sizeof(RowsProb));
  if (st != QFX_OK) return st;
  st = qfx_flush_uploads(ctx);
  if (st != QFX_OK) return st;
  QfxLaunchScope scope(ctx, QFX_K_CGEMM_SMALL, bytes, flops);
  cgemm_rows_grouped_kernel<<<unsigned(blocks), 256, 0, ctx->stream>>>(
      static_cast<const RowsProb*>(ws), int(probs.size()));
  QFX_HIP(hipGetLastError());
  return QFX_OK;
}

extern "C" int qfx_cgemm_grouped_c64(qfx_context* ctx, int precision, int count,
                                     const qfx_gemm_desc* descs) {
  QFX_CHECK_CTX(ctx);
  QFX_NOT_CAPTURING(ctx, "qfx_cgemm_grouped_c64");
  if (count < 0 || (count > 0 && !descs))
    return qfx_set_error(QFX_ERR_INVALID, "qfx_cgemm_grouped_c64: invalid problem list");
  if (precision < QFX_PREC_FP32 || precision > QFX_PREC_FP8)
    return qfx_set_error(QFX_ERR_INVALID, "qfx_cgemm_grouped_c64: unknown precision");
  const bool split = precision == QFX_PREC_SPLIT_FP16 || precision == QFX_PREC_SPLIT_BF16 ||
                     precision == QFX_PREC_SPLIT_BF16X3;
  // Problems for the large tiles go into one launch, and so do those of the
  // row kernel (k, n <= 16, any precision: exact FP32, the kernel
  // qfx_cgemm_c64 runs them on); the others (other precisions, odd sizes,
  // unaligned or too large operands) run one by one exactly as
  // qfx_cgemm_c64 would run them.
  std::vector<int> grouped, rows;
  for (int i = 0; i < count; ++i) {
    const qfx_gemm_desc& d = descs[i];
    if (d.m > 0 && d.n > 0 && d.k > 0 && d.k <= uint64_t(kSmallK) && d.n <= 16 && d.A && d.B &&
        d.C && (d.m + 255) / 256 <= 0x7fffffffULL) {
      rows.push_back(i);
      continue;
    }
    const bool eligible = split && d.m > 0 && d.n > 0 && d.k > uint64_t(kSmallK) &&
                          d.k % 2 == 0 && d.n % 2 == 0 && d.A && d.B && d.C &&
                          reinterpret_cast<uintptr_t>(d.A) % 16 == 0 &&
                          reinterpret_cast<uintptr_t>(d.B) % 16 == 0 &&
                          big_operands_fit(d.m, d.n, d.k) && big_operands_whole(d.m, d.n, d.k) &&
                          d.m <= 0x7fffffffULL &&
                          d.n <= 0x7fffffffULL;
    if (eligible) {
      grouped.push_back(i);
      continue;
    }
    const int st = qfx_cgemm_c64(ctx, precision, d.m, d.n, d.k, d.A, d.B, d.C);
    if (st != QFX_OK) return st;
  }
  if (!rows.empty()) {
    const int st = launch_rows_grouped(ctx, descs, rows);
    if (st != QFX_OK) return st;
  }
  if (grouped.empty()) return QFX_OK;
  double flops = 0, bytes = 0;
  size_t nparts = 0;
  std::vector<std::array<int, 2>> blocks(grouped.size(), {0, 0});
  for (size_t g = 0; g < grouped.size(); ++g) {
    const qfx_gemm_desc& d = descs[grouped[g]];
    flops += 8.0 * double(d.m) * double(d.n) * double(d.k);
    bytes += 8.0 * (double(d.m) * d.k + double(d.k) * d.n + double(d.m) * d.n);
    if (precision == QFX_PREC_SPLIT_FP16) {
      blocks[g][0] = int(std::min<uint64_t>(256, std::max<uint64_t>(1, (d.m * d.k + 16383) / 16384)));
      blocks[g][1] = int(std::min<uint64_t>(256, std::max<uint64_t>(1, (d.k * d.n + 16383) / 16384)));
      nparts += size_t(blocks[g][0] + blocks[g][1]);
    }
  }
  QfxLaunchScope scope(ctx, QFX_K_CGEMM, bytes, flops);
  const size_t part_bytes = (nparts * sizeof(float) + 255) / 256 * 256;
  const size_t desc_bytes = grouped.size() * sizeof(GroupTile);
  void* ws = nullptr;
  int st = qfx_workspace(ctx, part_bytes + desc_bytes, &ws);
  if (st != QFX_OK) return st;
  float* part = static_cast<float*>(ws);
  GroupTile* dev = reinterpret_cast<GroupTile*>(static_cast<char*>(ws) + part_bytes);
  // host staging of the problem list: the previous grouped call's copy
  // must have been read before the staging buffer is rewritten
  if (!ctx->group_ev) QFX_HIP(hipEventCreateWithFlags(&ctx->group_ev, hipEventDisableTiming));
  else QFX_HIP(hipEventSynchronize(ctx->group_ev));
  ctx->group_host.resize(desc_bytes);
  GroupTile* host = reinterpret_cast<GroupTile*>(ctx->group_host.data());
  uint64_t tile = 0;
  size_t off = 0;
  for (size_t g = 0; g < grouped.size(); ++g) {
    const qfx_gemm_desc& d = descs[grouped[g]];
    const float2* A2 = reinterpret_cast<const float2*>(d.A);
    const float2* B2 = reinterpret_cast<const float2*>(d.B);
    GroupTile t{};
    t.A = A2;
    t.B = B2;
    t.C = reinterpret_cast<float2*>(d.C);
    t.M = int(d.m);
    t.N = int(d.n);
    t.K = int(d.k);
    t.tiles_m = int((d.m + 255) / 256);
    t.first_tile = tile;
    tile += uint64_t(t.tiles_m) * ((d.n + 127) / 128);
    if (precision == QFX_PREC_SPLIT_FP16) {
      const int ba = blocks[g][0], bb = blocks[g][1];
      amax_partial_kernel<<<unsigned(ba + bb), 256, 0, ctx->stream>>>(A2, d.m * d.k, B2,
                                                                      d.k * d.n, ba, part + off);
      t.amax = AmaxRef{part + off, ba, part + off + ba, bb};
      off += size_t(ba + bb);
    }
    host[g] = t;
  }
  QFX_HIP(hipGetLastError());
  QFX_HIP(hipMemcpyAsync(dev, host, desc_bytes, hipMemcpyHostToDevice, ctx->stream));
  QFX_HIP(hipEventRecord(ctx->group_ev, ctx->stream));
  if (ctx->cus <= 0) {
    int cus = 0;
    QFX_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
    ctx->cus = std::max(cus, 1);
  }
  const unsigned grid = unsigned(std::min<uint64_t>(tile, uint64_t(ctx->cus)));
  auto go = [&](auto mode) -> hipError_t {
    constexpr int MODE = decltype(mode)::value;
    static hipError_t e = hipFuncSetAttribute(
        reinterpret_cast<const void*>(&cgemm_split_grouped_kernel<MODE>),
        hipFuncAttributeMaxDynamicSharedMemorySize, int(split_big_smem<MODE, 4, 2, 4>()));
    if (e != hipSuccess) return e;
    cgemm_split_grouped_kernel<MODE><<<grid, 512, split_big_smem<MODE, 4, 2, 4>(), ctx->stream>>>(
        dev, int(grouped.size()), tile);
    return hipGetLastError();
  };
  if (precision == QFX_PREC_SPLIT_BF16X3) QFX_HIP(go(std::integral_constant<int, 0>{}));
  else if (precision == QFX_PREC_SPLIT_BF16) QFX_HIP(go(std::integral_constant<int, 1>{}));
  else QFX_HIP(go(std::integral_constant<int, 2>{}));
  return QFX_OK;
}

// ------------------------------------------------------------ autotuner

extern "C" int qfx_autotune_enable(qfx_context* ctx, int on) {
  QFX_CHECK_CTX(ctx);
  ctx->autotune = on != 0;
  return QFX_OK;
}

extern "C" int qfx_autotune_entries(qfx_context* ctx, uint64_t* out, int cap, int* count) {
  QFX_CHECK_CTX(ctx);
  int i = 0;
  for (const auto& kv : ctx->tuned) {
    if (out && i < cap) {
      for (int j = 0; j < 5; ++j) out[6 * i + j] = kv.first[j];
      out[6 * i + 5] = uint64_t(kv.second);
    }
    ++i;
  }
  if (count) *count = i;
  return QFX_OK;
}

extern "C" int qfx_autotune_set(qfx_context* ctx, int precision, int form, uint64_t m,
                                uint64_t n, uint64_t k, int tile) {
  QFX_CHECK_CTX(ctx);
  if (precision < QFX_PREC_FP32 || precision > QFX_PREC_FP8 || form < 0 || form > 2 ||
      tile < 0 || tile > 3)
    return qfx_set_error(QFX_ERR_INVALID, "qfx_autotune_set: invalid precision, form or tile");
  const std::array<uint64_t, 5> key{uint64_t(precision), uint64_t(form), m, n, k};
  if (tile == 0) {
    ctx->tuned.erase(key);
    return QFX_OK;
  }
  const auto cands = layout_candidates(precision, form, m, n, k);
  if (std::find(cands.begin(), cands.end(), BigLayout(tile)) == cands.end())
    return qfx_set_error(QFX_ERR_INVALID, "qfx_autotune_set: the tile cannot run this product");
  ctx->tuned[key] = tile;
  return QFX_OK;
}
