// Complex GEMM in FP8 (OCP E4M3) with one power-of-two scale per 128 x 128
// tile of each real operand: the low-fidelity precision QFX_PREC_FP8.
//
// C = A B (complex64, C row-major m x n) is computed as the real product
//   C_real [m x 2n] = A_real [m x 2k] * Bhat [2k x 2n]
// where A_real is A's interleaved storage read as real numbers and Bhat
// holds each complex b = br + i bi as the 2 x 2 block [[br, bi], [-bi, br]]
// (rows 2kk, 2kk+1; columns 2j, 2j+1), so C_real is C's interleaved storage.
//
// Two quantize kernels write A_real (m x K, K contiguous) and Bhat^T
// (2n x K, K contiguous) as E4M3 codes, K = 2k padded to 128: each 128 x 128
// tile is multiplied by 2^-s, s = e - 8 for the tile's max |x| = f 2^e,
// f in [0.5, 1), so every scaled value is below 256 (E4M3's largest finite
// value is 448), and rounded to nearest even (v_cvt_pk_fp8_f32).  The tile
// exponents are kept as E8M0 codes (s + 127).
//
// The GEMM kernel multiplies the codes on v_mfma_scale_f32_32x32x64_f8f6f4,
// whose block scales apply 2^sA * 2^sB to each product, so the FP32
// accumulator receives the dequantized products directly.  The reference
// has no FP8 path: its arithmetic is cblas_cgemm in FP32
// (src/tensor.cpp:960-980).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "qfx_internal.h"

namespace {

constexpr int kQT = 128;        // scale tile edge (real elements)
constexpr int kF8BM = 256;      // GEMM block tile: rows of C_real
constexpr int kF8BN = 256;      //                  columns of C_real
constexpr int kF8BK = 128;      //                  K per stage (= one scale tile)
constexpr int kF8StageBytes = (kF8BM + kF8BN) * kF8BK;  // A and B codes of one stage

typedef int v8i __attribute__((ext_vector_type(8)));
typedef float v16f __attribute__((ext_vector_type(16)));

__device__ __forceinline__ int tile_shift(float amax) {
  if (!(amax > 0.f)) return 0;
  int e;
  frexpf(amax, &e);
  return min(max(e - 8, -127), 127);
}

// Codes of two floats (already scaled) in the low 16 bits.
__device__ __forceinline__ uint32_t e4m3_pair(float lo, float hi) {
  return uint32_t(__builtin_amdgcn_cvt_pk_fp8_f32(lo, hi, 0, false)) & 0xffffu;
}

__device__ __forceinline__ float block_max(float v, float* red) {
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  if ((threadIdx.x & 63) == 0) red[w] = v;
  __syncthreads();
  v = red[0];
  for (int i = 1; i < nw; ++i) v = fmaxf(v, red[i]);
  return v;
}

// Source offset (in complex elements) of matrix element (r, c): row-major
// with `ld` columns, or through the two digit maps.
template <bool MAPPED>
__device__ __forceinline__ uint64_t elem_offset(uint64_t r, uint64_t c, uint64_t ld,
                                                const qfx_index_map& rows,
                                                const qfx_index_map& cols) {
  if constexpr (MAPPED)
    return map_offset(rows, r) + map_offset(cols, c);
  else
    return r * ld + c;
}

// A (m x k complex) -> Aq (Mp x Kp codes), sa[(row tile) * KT + k tile].
// Block (x, y) = (128-row tile, 128-wide K tile = 64 complex columns);
// 256 threads, thread t owns complex column t % 64 of rows t / 64 + 4 i.
template <bool MAPPED>
__global__ __launch_bounds__(256) void fp8_quant_a_kernel(const float2* __restrict__ A,
                                                          uint64_t m, uint64_t k,
                                                          qfx_index_map rows,
                                                          qfx_index_map cols,
                                                          uint8_t* __restrict__ Aq,
                                                          uint64_t Kp, int* __restrict__ sa) {
  __shared__ float red[4];
  const uint64_t r0 = uint64_t(blockIdx.x) * kQT;
  const uint64_t c = uint64_t(blockIdx.y) * (kQT / 2) + (threadIdx.x & 63);
  const int rr = threadIdx.x >> 6;
  float2 v[32];
  float amax = 0.f;
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    const uint64_t r = r0 + rr + 4 * i;
    v[i] = (r < m && c < k) ? A[elem_offset<MAPPED>(r, c, k, rows, cols)] : float2{0.f, 0.f};
    amax = fmaxf(amax, fmaxf(fabsf(v[i].x), fabsf(v[i].y)));
  }
  amax = block_max(amax, red);
  const int s = tile_shift(amax);
  if (threadIdx.x == 0) sa[uint64_t(blockIdx.x) * gridDim.y + blockIdx.y] = s + 127;
#pragma unroll
  for (int i = 0; i < 32; ++i) {
    const uint64_t r = r0 + rr + 4 * i;
    const uint32_t q = e4m3_pair(ldexpf(v[i].x, -s), ldexpf(v[i].y, -s));
    *reinterpret_cast<uint16_t*>(Aq + r * Kp + 2 * c) = uint16_t(q);
  }
}

// B (k x n complex) -> Bt = Bhat^T (Np x Kp codes), sb[(row tile of Bt) * KT + k tile].
// Block (x, y) = (64 complex columns of B = 128 rows of Bt, 64 complex rows
// of B = 128 K); the tile is transposed through LDS so that both the reads
// of B and the 16-byte writes of Bt are contiguous.
template <bool MAPPED>
__global__ __launch_bounds__(256) void fp8_quant_b_kernel(const float2* __restrict__ B,
                                                          uint64_t k, uint64_t n,
                                                          qfx_index_map rows,
                                                          qfx_index_map cols,
                                                          uint8_t* __restrict__ Bt,
                                                          uint64_t Kp, int* __restrict__ sb) {
  __shared__ float red[4];
  __shared__ __align__(16) uint8_t tile[kQT][kQT + 16];
  const uint64_t j = uint64_t(blockIdx.x) * (kQT / 2) + (threadIdx.x & 63);
  const uint64_t kk0 = uint64_t(blockIdx.y) * (kQT / 2);
  const int jl = threadIdx.x & 63, kr = threadIdx.x >> 6;
  float2 v[16];
  float amax = 0.f;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const uint64_t kk = kk0 + kr + 4 * i;
    v[i] = (kk < k && j < n) ? B[elem_offset<MAPPED>(kk, j, n, rows, cols)] : float2{0.f, 0.f};
    amax = fmaxf(amax, fmaxf(fabsf(v[i].x), fabsf(v[i].y)));
  }
  amax = block_max(amax, red);
  const int s = tile_shift(amax);
  if (threadIdx.x == 0) sb[uint64_t(blockIdx.x) * gridDim.y + blockIdx.y] = s + 127;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int kl = kr + 4 * i;
    const float br = ldexpf(v[i].x, -s), bi = ldexpf(v[i].y, -s);
    // Bt[2j][2kk..2kk+1] = (br, -bi), Bt[2j+1][2kk..2kk+1] = (bi, br)
    *reinterpret_cast<uint16_t*>(&tile[2 * jl][2 * kl]) = uint16_t(e4m3_pair(br, -bi));
    *reinterpret_cast<uint16_t*>(&tile[2 * jl + 1][2 * kl]) = uint16_t(e4m3_pair(bi, br));
  }
  __syncthreads();
  const uint64_t row0 = uint64_t(blockIdx.x) * kQT, col0 = uint64_t(blockIdx.y) * kQT;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int q = threadIdx.x + 256 * i;  // 1024 chunks of 16 bytes
    const int r = q >> 3, c = (q & 7) * 16;
    *reinterpret_cast<uint4*>(Bt + (row0 + r) * Kp + col0 + c) =
        *reinterpret_cast<const uint4*>(&tile[r][c]);
  }
}

// LDS address of 16-byte chunk c of row r in a 128-byte-row tile: chunks are
// XOR-swizzled by the row so that 8 consecutive rows read at the same chunk
// cover all banks.
__device__ __forceinline__ int swz(int r, int c) { return r * kF8BK + ((c ^ (r & 7)) << 4); }

// MFMA operand of one lane: 32 consecutive K bytes (chunks c0, c0 + 1) of row r.
__device__ __forceinline__ v8i lds_frag(const unsigned char* base, int r, int c0) {
  const uint4 lo = *reinterpret_cast<const uint4*>(base + swz(r, c0));
  const uint4 hi = *reinterpret_cast<const uint4*>(base + swz(r, c0 + 1));
  return v8i{int(lo.x), int(lo.y), int(lo.z), int(lo.w), int(hi.x), int(hi.y), int(hi.z),
             int(hi.w)};
}

// C_real (M x N, row stride N) = Aq (Mp x Kp) * Bt^T (Np x Kp), all codes
// zero-padded, so no operand bound checks are needed; C stores are guarded.
// WM x WN waves; a wave owns (256 / WM) x (256 / WN) of the block tile as
// 32 x 32 MFMA tiles; its rows and columns lie in one scale tile each.
template <int WM, int WN>
__global__ __launch_bounds__(64 * WM * WN) void cgemm_fp8_kernel(
    const uint8_t* __restrict__ Aq, const uint8_t* __restrict__ Bt, const int* __restrict__ sa,
    const int* __restrict__ sb, float* __restrict__ C, uint64_t M, uint64_t N, uint64_t Kp) {
  constexpr int kThreads = 64 * WM * WN;
  constexpr int TM = kF8BM / WM, TN = kF8BN / WN;  // wave tile
  constexpr int IM = TM / 32, JN = TN / 32;         // MFMA tiles per wave
  constexpr int kLoads = kF8BM * kF8BK / 16 / kThreads;  // 16-byte chunks per thread per operand
  constexpr int kRowStep = kThreads / 8;
  static_assert(TM <= kQT && TN <= kQT && kLoads >= 1, "wave tile must fit one scale tile");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave / WN, wn = wave % WN;
  const uint64_t m0 = uint64_t(blockIdx.x) * kF8BM, n0 = uint64_t(blockIdx.y) * kF8BN;
  const int KT = int(Kp / kF8BK);
  // Global -> LDS copy: rows tid / 8 + kRowStep i, chunk tid % 8 (the
  // swizzle depends on row % 8 only, and kRowStep is a multiple of 8).
  const uint8_t* ga = Aq + (m0 + (tid >> 3)) * Kp + (tid & 7) * 16;
  const uint8_t* gb = Bt + (n0 + (tid >> 3)) * Kp + (tid & 7) * 16;
  const uint64_t gstep = uint64_t(kRowStep) * Kp;
  const int la = swz(tid >> 3, tid & 7);
  const int* sa_row = sa + ((m0 + uint64_t(wm) * TM) / kQT) * uint64_t(KT);
  const int* sb_row = sb + ((n0 + uint64_t(wn) * TN) / kQT) * uint64_t(KT);

  v16f acc[IM][JN];
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.f;

  const int fr = lane & 31, fh = lane >> 5;  // fragment row, K half
  uint4 ra[kLoads], rb[kLoads];
#pragma unroll
  for (int i = 0; i < kLoads; ++i) {
    ra[i] = *reinterpret_cast<const uint4*>(ga + i * gstep);
    rb[i] = *reinterpret_cast<const uint4*>(gb + i * gstep);
  }
#pragma unroll
  for (int i = 0; i < kLoads; ++i) {
    *reinterpret_cast<uint4*>(smem_raw + la + kRowStep * kF8BK * i) = ra[i];
    *reinterpret_cast<uint4*>(smem_raw + kF8BM * kF8BK + la + kRowStep * kF8BK * i) = rb[i];
  }
  __syncthreads();
  // Unconditional loads and stores keep ra / rb in registers: the last
  // stage reloads itself into the idle buffer, which nothing reads again.
  for (int kt = 0; kt < KT; ++kt) {
    const uint64_t knext = uint64_t(min(kt + 1, KT - 1)) * kF8BK;
#pragma unroll
    for (int i = 0; i < kLoads; ++i) {
      ra[i] = *reinterpret_cast<const uint4*>(ga + i * gstep + knext);
      rb[i] = *reinterpret_cast<const uint4*>(gb + i * gstep + knext);
    }
    const unsigned char* sA = smem_raw + (kt & 1) * kF8StageBytes;
    const unsigned char* sB = sA + kF8BM * kF8BK;
    const int scale_a = sa_row[kt], scale_b = sb_row[kt];
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      const int c0 = 4 * ks + 2 * fh;  // K bytes 64 ks + 32 fh .. +31
      v8i a[IM];
#pragma unroll
      for (int i = 0; i < IM; ++i) a[i] = lds_frag(sA, wm * TM + 32 * i + fr, c0);
#pragma unroll
      for (int j = 0; j < JN; ++j) {
        const v8i b = lds_frag(sB, wn * TN + 32 * j + fr, c0);
#pragma unroll
        for (int i = 0; i < IM; ++i)
          acc[i][j] = __builtin_amdgcn_mfma_scale_f32_32x32x64_f8f6f4(
              a[i], b, acc[i][j], 0, 0, 0, scale_a, 0, scale_b);
      }
    }
    unsigned char* dA = smem_raw + ((kt + 1) & 1) * kF8StageBytes;
#pragma unroll
    for (int i = 0; i < kLoads; ++i) {
      *reinterpret_cast<uint4*>(dA + la + kRowStep * kF8BK * i) = ra[i];
      *reinterpret_cast<uint4*>(dA + kF8BM * kF8BK + la + kRowStep * kF8BK * i) = rb[i];
    }
    __syncthreads();
  }

  // Accumulator layout of a 32 x 32 tile: lane -> column lane % 32, register
  // e -> row 8 (e / 4) + 4 (lane / 32) + e % 4.
#pragma unroll
  for (int i = 0; i < IM; ++i)
#pragma unroll
    for (int j = 0; j < JN; ++j) {
      const uint64_t col = n0 + wn * TN + 32 * j + fr;
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const uint64_t row = m0 + wm * TM + 32 * i + 8 * (e >> 2) + 4 * fh + (e & 3);
        if (row < M && col < N) C[row * N + col] = acc[i][j][e];
      }
    }
}

constexpr int kF8WM = 2, kF8WN = 2;  // 4 waves of 128 x 128

hipError_t fp8_prepare() {
  static hipError_t e = hipFuncSetAttribute(
      reinterpret_cast<const void*>(&cgemm_fp8_kernel<kF8WM, kF8WN>),
                                            hipFuncAttributeMaxDynamicSharedMemorySize,
                                            2 * kF8StageBytes);
  return e;
}

uint64_t round_up(uint64_t x, uint64_t r) { return (x + r - 1) / r * r; }

}  // namespace

bool qfx_fp8_applies(uint64_t m, uint64_t n, uint64_t k) {
  // the quantize grid has Mp / 128 blocks along x (below 2^31)
  return m >= 128 && n >= 64 && k >= 64 && m <= (uint64_t(1) << 36) &&
         round_up(2 * n, kF8BN) / kF8BN <= 65535 && round_up(2 * k, kQT) / kQT <= 65535;
}

int qfx_cgemm_fp8(qfx_context* ctx, const float2* A, const float2* B, float2* C, uint64_t m,
                  uint64_t n, uint64_t k, const qfx_index_map* maps) {
  if (!qfx_fp8_applies(m, n, k))
    return qfx_set_error(QFX_ERR_UNSUPPORTED, "qfx_cgemm_c64: shape too small for FP8 tiles");
  const uint64_t Mp = round_up(m, kF8BM), Np = round_up(2 * n, kF8BN), Kp = round_up(2 * k, kQT);
  const uint64_t KT = Kp / kQT;
  // launch limits of the three grids (qfx_fp8_applies keeps every shape it
  // accepts inside them; checked again here before anything is launched)
  if (Mp / kQT > 0x7fffffffULL || KT > 65535 || Np / kQT > 0x7fffffffULL || Np / kF8BN > 65535)
    return qfx_set_error(QFX_ERR_INVALID, "qfx_cgemm_c64: FP8 grid beyond the launch limits");
  const uint64_t a_bytes = Mp * Kp, b_bytes = Np * Kp;
  const uint64_t sa_n = (Mp / kQT) * KT, sb_n = (Np / kQT) * KT;
  void* ws = nullptr;
  int st = qfx_workspace(ctx, a_bytes + b_bytes + 4 * (sa_n + sb_n), &ws);
  if (st != QFX_OK) return st;
  uint8_t* Aq = static_cast<uint8_t*>(ws);
  uint8_t* Bt = Aq + a_bytes;
  int* sa = reinterpret_cast<int*>(Bt + b_bytes);
  int* sb = sa + sa_n;
  const qfx_index_map none{};
  const dim3 ga(unsigned(Mp / kQT), unsigned(KT)), gb(unsigned(Np / kQT), unsigned(KT));
  if (maps) {
    fp8_quant_a_kernel<true><<<ga, 256, 0, ctx->stream>>>(A, m, k, maps[0], maps[1], Aq, Kp, sa);
    fp8_quant_b_kernel<true><<<gb, 256, 0, ctx->stream>>>(B, k, n, maps[2], maps[3], Bt, Kp, sb);
  } else {
    fp8_quant_a_kernel<false><<<ga, 256, 0, ctx->stream>>>(A, m, k, none, none, Aq, Kp, sa);
    fp8_quant_b_kernel<false><<<gb, 256, 0, ctx->stream>>>(B, k, n, none, none, Bt, Kp, sb);
  }
  QFX_HIP(hipGetLastError());
  QFX_HIP(fp8_prepare());
  const dim3 g(unsigned(Mp / kF8BM), unsigned(Np / kF8BN));
  cgemm_fp8_kernel<kF8WM, kF8WN><<<g, 64 * kF8WM * kF8WN, 2 * kF8StageBytes, ctx->stream>>>(
      Aq, Bt, sa, sb, reinterpret_cast<float*>(C), m, 2 * n, Kp);
  QFX_HIP(hipGetLastError());
  return QFX_OK;
}
