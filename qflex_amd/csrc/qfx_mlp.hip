// The cost model's network: an MLP 64 -> 256 -> 256 -> 1 (ReLU) in BF16
// with FP32 accumulation, trained with MSE and Adam, all on the GPU.
// (No counterpart in the reference: it scores candidate contraction
// orderings; amplitudes always come from exact contraction.)
//
// Parameters live in one FP32 buffer (layout in include/qfx.h); every GEMM
// rounds its operands to BF16 (round to nearest even) and accumulates in
// FP32 on v_mfma_f32_32x32x16_bf16.  Activations are kept as BF16, the
// values the next GEMM consumes, so the backward pass differentiates
// exactly the function the forward pass computed.
#include <algorithm>

#include "qfx_internal.h"

namespace {

typedef float floatx16 __attribute__((ext_vector_type(16)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kIn = QFX_MLP_IN, kHid = QFX_MLP_HIDDEN;
constexpr int kW1 = 0, kB1 = kW1 + kIn * kHid, kW2 = kB1 + kHid, kB2 = kW2 + kHid * kHid,
              kW3 = kB2 + kHid, kB3 = kW3 + kHid, kParams = kB3 + 1;
static_assert(kParams == QFX_MLP_PARAMS, "parameter layout");

// C[M x N] (FP32, row-major, ldc = N) = op(A) . op(B) with BF16 operands
// read through strides: A(i, k) = A[i*as0 + k*as1], B(k, j) = B[k*bs0 +
// j*bs1].  Epilogue: + bias[j] (if bias), ReLU (if relu), and a BF16 copy
// (if Cb).  Block 64 x 64 outputs, 4 waves of 32 x 32, k in steps of 16
// staged through LDS as [row][k] for A and [col][k] for B.
__global__ __launch_bounds__(256) void mlp_gemm_kernel(const __bf16* __restrict__ A, int64_t as0,
                                                       int64_t as1, const __bf16* __restrict__ B,
                                                       int64_t bs0, int64_t bs1,
                                                       float* __restrict__ C,
                                                       __bf16* __restrict__ Cb,
                                                       const float* __restrict__ bias, int relu,
                                                       int M, int N, int K) {
  constexpr int kRow = 24;  // 16 k + 8 pad (48-byte rows)
  __shared__ __align__(16) __bf16 As[64 * kRow];
  __shared__ __align__(16) __bf16 Bs[64 * kRow];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const int m0 = blockIdx.x * 64, n0 = blockIdx.y * 64;
  const int half = lane >> 5, col = lane & 31;
  floatx16 acc;
#pragma unroll
  for (int r = 0; r < 16; ++r) acc[r] = 0.f;
  const __bf16 zero = static_cast<__bf16>(0.f);
  for (int k0 = 0; k0 < K; k0 += 16) {
    // 64 x 16 elements of each operand, 4 per thread
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const int e = tid + 256 * h;
      const int r = e >> 4, kk = e & 15;
      const int gi = m0 + r, gj = n0 + r, gk = k0 + kk;
      As[r * kRow + kk] = (gi < M && gk < K) ? A[gi * as0 + gk * as1] : zero;
      Bs[r * kRow + kk] = (gj < N && gk < K) ? B[gk * bs0 + gj * bs1] : zero;
    }
    __syncthreads();
    const bf16x8 a = *reinterpret_cast<const bf16x8*>(As + (wm * 32 + col) * kRow + 8 * half);
    const bf16x8 b = *reinterpret_cast<const bf16x8*>(Bs + (wn * 32 + col) * kRow + 8 * half);
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    __syncthreads();
  }
  // register r of lane l: row 8*(r/4) + 4*(l/32) + r%4, column l%32
#pragma unroll
  for (int r = 0; r < 16; ++r) {
    const int i = m0 + wm * 32 + 8 * (r >> 2) + 4 * half + (r & 3);
    const int j = n0 + wn * 32 + col;
    if (i >= M || j >= N) continue;
    float v = acc[r];
    if (bias) v += bias[j];
    if (relu) v = fmaxf(v, 0.f);
    C[int64_t(i) * N + j] = v;
    if (Cb) Cb[int64_t(i) * N + j] = static_cast<__bf16>(v);
  }
}

__global__ void to_bf16_kernel(const float* __restrict__ x, __bf16* __restrict__ y, int64_t n) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    y[i] = static_cast<__bf16>(x[i]);
}

// dy = 2 (y - t) / B; partial sums of (y - t)^2 / B per block (fixed order).
__global__ __launch_bounds__(256) void mse_grad_kernel(const float* __restrict__ y,
                                                       const float* __restrict__ t,
                                                       float* __restrict__ dy,
                                                       float* __restrict__ loss_part, int B) {
  __shared__ float red[256];
  const int i = blockIdx.x * 256 + threadIdx.x;
  float l = 0.f;
  if (i < B) {
    const float d = y[i] - t[i];
    dy[i] = 2.f * d / float(B);
    l = d * d / float(B);
  }
  red[threadIdx.x] = l;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0];
}

// dZ[b][j] = (H[b][j] > 0) * g[b][j], with g = dy[b] * w3[j] (layer 3) or
// g = dH[b][j] (layer 2 -> 1); BF16 copy for the next GEMM.
__global__ void relu_grad_kernel(const __bf16* __restrict__ H, const float* __restrict__ dH,
                                 const float* __restrict__ dy, const float* __restrict__ w3,
                                 float* __restrict__ dZ, __bf16* __restrict__ dZb, int B, int N) {
  const int64_t n = int64_t(B) * N;
  for (int64_t e = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; e < n;
       e += int64_t(gridDim.x) * blockDim.x) {
    const int b = int(e / N), j = int(e - int64_t(b) * N);
    const float g = dH ? dH[e] : dy[b] * static_cast<float>(static_cast<__bf16>(w3[j]));
    const float v = static_cast<float>(H[e]) > 0.f ? g : 0.f;
    dZ[e] = v;
    dZb[e] = static_cast<__bf16>(v);
  }
}

// out[j] = sum_b X[b][j] (column sums over the batch, fixed order).
__global__ __launch_bounds__(256) void col_sum_kernel(const float* __restrict__ X,
                                                      float* __restrict__ out, int B, int N) {
  __shared__ float red[256];
  const int j = blockIdx.x;
  float s = 0.f;
  for (int b = threadIdx.x; b < B; b += 256) s += X[int64_t(b) * N + j];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (int(threadIdx.x) < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = red[0];
}

__global__ void adam_kernel(float* __restrict__ p, const float* __restrict__ g,
                            float* __restrict__ m, float* __restrict__ v, int64_t n, float lr,
                            float b1, float b2, float eps, float c1, float c2) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

unsigned grid_for(int64_t n) { return unsigned(std::min<int64_t>((n + 255) / 256, 4096)); }

// Workspace layout (bytes, 256-aligned pieces) for a batch of B.
struct MlpWork {
  __bf16 *w_b, *x_b, *h1_b, *h2_b, *dz2_b, *dz1_b;
  float *h1, *h2, *y, *dy, *dz2, *dh1, *dz1, *loss_part;
};

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

int mlp_work(qfx_context* ctx, int B, MlpWork* w) {
  const size_t bh = size_t(B) * kHid;
  size_t sizes[] = {kParams * 2, size_t(B) * kIn * 2, bh * 2, bh * 2, bh * 2, bh * 2,
                    bh * 4,      bh * 4,              size_t(B) * 4, size_t(B) * 4, bh * 4,
                    bh * 4,      bh * 4,              size_t((B + 255) / 256) * 4};
  size_t total = 0;
  for (size_t s : sizes) total += align256(s);
  void* base = nullptr;
  int st = qfx_workspace(ctx, total, &base);
  if (st != QFX_OK) return st;
  char* p = static_cast<char*>(base);
  void** slots[] = {reinterpret_cast<void**>(&w->w_b),  reinterpret_cast<void**>(&w->x_b),
                    reinterpret_cast<void**>(&w->h1_b), reinterpret_cast<void**>(&w->h2_b),
                    reinterpret_cast<void**>(&w->dz2_b), reinterpret_cast<void**>(&w->dz1_b),
                    reinterpret_cast<void**>(&w->h1),   reinterpret_cast<void**>(&w->h2),
                    reinterpret_cast<void**>(&w->y),    reinterpret_cast<void**>(&w->dy),
                    reinterpret_cast<void**>(&w->dz2),  reinterpret_cast<void**>(&w->dh1),
                    reinterpret_cast<void**>(&w->dz1),  reinterpret_cast<void**>(&w->loss_part)};
  for (int i = 0; i < 14; ++i) {
    *slots[i] = p;
    p += align256(sizes[i]);
  }
  return QFX_OK;
}

void gemm(hipStream_t s, const __bf16* A, int64_t as0, int64_t as1, const __bf16* B, int64_t bs0,
          int64_t bs1, float* C, __bf16* Cb, const float* bias, int relu, int M, int N, int K) {
  dim3 g{unsigned((M + 63) / 64), unsigned((N + 63) / 64), 1};
  mlp_gemm_kernel<<<g, 256, 0, s>>>(A, as0, as1, B, bs0, bs1, C, Cb, bias, relu, M, N, K);
}

// Forward pass into the workspace; y = w.y.
int forward(qfx_context* ctx, int B, const float* X, const float* params, const MlpWork& w) {
  hipStream_t s = ctx->stream;
  to_bf16_kernel<<<grid_for(kParams), 256, 0, s>>>(params, w.w_b, kParams);
  to_bf16_kernel<<<grid_for(int64_t(B) * kIn), 256, 0, s>>>(X, w.x_b, int64_t(B) * kIn);
  // H1 = relu(X W1 + b1), W1[i][j] at W1 + i*kHid + j
  gemm(s, w.x_b, kIn, 1, w.w_b + kW1, kHid, 1, w.h1, w.h1_b, params + kB1, 1, B, kHid, kIn);
  gemm(s, w.h1_b, kHid, 1, w.w_b + kW2, kHid, 1, w.h2, w.h2_b, params + kB2, 1, B, kHid, kHid);
  // y = H2 w3 + b3 (w3 as a 256 x 1 matrix)
  gemm(s, w.h2_b, kHid, 1, w.w_b + kW3, 1, 1, w.y, nullptr, params + kB3, 0, B, 1, kHid);
  QFX_HIP(hipGetLastError());
  return QFX_OK;
}

}  // namespace

extern "C" int qfx_mlp_forward(qfx_context* ctx, uint64_t batch, const float* X,
                               const float* params, float* y) {
  QFX_CHECK_CTX(ctx);
  if (batch == 0) return QFX_OK;
  if (!X || !params || !y) return qfx_set_error(QFX_ERR_INVALID, "qfx_mlp_forward: null pointer");
  if (batch > (uint64_t(1) << 20))
    return qfx_set_error(QFX_ERR_INVALID, "qfx_mlp_forward: batch above 2^20");
  const int B = int(batch);
  MlpWork w;
  int st = mlp_work(ctx, B, &w);
  if (st != QFX_OK) return st;
  st = forward(ctx, B, X, params, w);
  if (st != QFX_OK) return st;
  QFX_HIP(hipMemcpyAsync(y, w.y, size_t(B) * sizeof(float), hipMemcpyDeviceToDevice,
                         ctx->stream));
  return QFX_OK;
}

extern "C" int qfx_mlp_backward(qfx_context* ctx, uint64_t batch, const float* X,
                                const float* target, const float* params, float* grads,
                                float* loss) {
  QFX_CHECK_CTX(ctx);
  if (batch == 0) return qfx_set_error(QFX_ERR_INVALID, "qfx_mlp_backward: empty batch");
  if (!X || !target || !params || !grads || !loss)
    return qfx_set_error(QFX_ERR_INVALID, "qfx_mlp_backward: null pointer");
  if (batch > (uint64_t(1) << 20))
    return qfx_set_error(QFX_ERR_INVALID, "qfx_mlp_backward: batch above 2^20");
  const int B = int(batch);
  MlpWork w;
  int st = mlp_work(ctx, B, &w);
  if (st != QFX_OK) return st;
  st = forward(ctx, B, X, params, w);
  if (st != QFX_OK) return st;
  hipStream_t s = ctx->stream;
  const int nblk = (B + 255) / 256;
  mse_grad_kernel<<<nblk, 256, 0, s>>>(w.y, target, w.dy, w.loss_part, B);
  // loss = sum of the block partials (fixed order): a 1 x nblk . nblk x 1
  // column sum
  col_sum_kernel<<<1, 256, 0, s>>>(w.loss_part, loss, nblk, 1);
  // layer 3: dw3[j] = sum_b H2[b][j] dy[b]; db3 = sum_b dy[b]
  {
    // dy as BF16 column for the GEMM: reuse dz1_b's first B entries
    to_bf16_kernel<<<grid_for(B), 256, 0, s>>>(w.dy, w.dz1_b, B);
    // (H2^T)(j, b) = H2[b][j]: A strides (1, kHid); dy(b, 0) = dy[b]
    gemm(s, w.h2_b, 1, kHid, w.dz1_b, 1, 1, grads + kW3, nullptr, nullptr, 0, kHid, 1, B);
    col_sum_kernel<<<1, 256, 0, s>>>(w.dy, grads + kB3, B, 1);
  }
  // dZ2 = relu'(H2) * dy w3^T
  relu_grad_kernel<<<grid_for(int64_t(B) * kHid), 256, 0, s>>>(w.h2_b, nullptr, w.dy,
                                                                params + kW3, w.dz2, w.dz2_b, B,
                                                                kHid);
  // dW2 = H1^T dZ2 (256 x 256), db2 = column sums of dZ2
  gemm(s, w.h1_b, 1, kHid, w.dz2_b, kHid, 1, grads + kW2, nullptr, nullptr, 0, kHid, kHid, B);
  col_sum_kernel<<<kHid, 256, 0, s>>>(w.dz2, grads + kB2, B, kHid);
  // dH1 = dZ2 W2^T: (W2^T)(k, j) = W2[j][k]
  gemm(s, w.dz2_b, kHid, 1, w.w_b + kW2, 1, kHid, w.dh1, nullptr, nullptr, 0, B, kHid, kHid);
  relu_grad_kernel<<<grid_for(int64_t(B) * kHid), 256, 0, s>>>(w.h1_b, w.dh1, nullptr, nullptr,
                                                                w.dz1, w.dz1_b, B, kHid);
  // dW1 = X^T dZ1 (64 x 256), db1 = column sums of dZ1
  gemm(s, w.x_b, 1, kIn, w.dz1_b, kHid, 1, grads + kW1, nullptr, nullptr, 0, kIn, kHid, B);
  col_sum_kernel<<<kHid, 256, 0, s>>>(w.dz1, grads + kB1, B, kHid);
  QFX_HIP(hipGetLastError());
  return QFX_OK;
}

extern "C" int qfx_mlp_adam(qfx_context* ctx, float* params, const float* grads, float* m,
                            float* v, int step, float lr, float beta1, float beta2, float eps) {
  QFX_CHECK_CTX(ctx);
  if (!params || !grads || !m || !v)
    return qfx_set_error(QFX_ERR_INVALID, "qfx_mlp_adam: null pointer");
  if (step < 1) return qfx_set_error(QFX_ERR_INVALID, "qfx_mlp_adam: step counts from 1");
  const float c1 = 1.f - powf(beta1, float(step)), c2 = 1.f - powf(beta2, float(step));
  adam_kernel<<<grid_for(kParams), 256, 0, ctx->stream>>>(params, grads, m, v, kParams, lr, beta1,
                                                          beta2, eps, c1, c2);
  QFX_HIP(hipGetLastError());
  return QFX_OK;
}
