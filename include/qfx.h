/*
 * qfx.h -- C-ABI of the MI355X (gfx950) contraction engine for qFlex's grid
 * tensor-network path.
 *
 * Plain C: opaque context, plain pointers and sizes, int status codes, no
 * exceptions and no C++/torch types across the boundary.  Device pointers
 * handed to these calls come from qfx_alloc (the context's device arena).
 *
 * The reference (qFlex) performs every floating-point primitive of the path
 * on the host through CBLAS and scalar C++ loops.  Each entry point below
 * names the reference call site it replaces.
 *
 * All device work of a context is issued in order on the context's single
 * HIP stream; calls return after enqueueing unless stated otherwise.
 */
#ifndef QFX_H_
#define QFX_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Interleaved complex64, identical in layout to std::complex<float>
 * (reference s_type, src/tensor.h:30). */
typedef struct {
  float re, im;
} qfx_c64;

/* Interleaved complex128, identical to std::complex<double> (the amplitude
 * type of reference src/contraction_utils.h:218-220). */
typedef struct {
  double re, im;
} qfx_c128;

typedef struct qfx_context qfx_context;

enum qfx_status {
  QFX_OK = 0,
  QFX_ERR_INVALID = 1,     /* bad argument (null pointer, bad shape, ...) */
  QFX_ERR_OUT_OF_MEMORY = 2,
  QFX_ERR_HIP = 3,         /* HIP runtime error; see qfx_last_error() */
  QFX_ERR_UNSUPPORTED = 4
};

/* Arithmetic forms of the complex GEMM. */
enum qfx_precision {
  QFX_PREC_FP32 = 0,       /* FP32 MFMA, FP32 accumulate */
  QFX_PREC_SPLIT_BF16 = 1, /* A,B = hi+mid BF16; hi*hi + hi*mid + mid*hi, FP32 acc */
  QFX_PREC_FP32_EXACT = 2, /* scalar fmaf in K order (defined reduction order) */
  QFX_PREC_SPLIT_BF16X3 = 3, /* A,B = hi+mid+lo BF16; 6 cross products, FP32 acc */
  QFX_PREC_SPLIT_FP16 = 4,   /* A,B scaled by powers of two to max|x| in [2^14, 2^15),
                                split into FP16 hi + lo; hi*hi + hi*lo + lo*hi, FP32 acc */
  QFX_PREC_FP8 = 5           /* low fidelity: A and the real 2x2 expansion of B as OCP
                                E4M3, one power-of-two scale per 128x128 tile (max scaled
                                value < 256), FP32 acc.  Small products run as in the other
                                precisions (exact FP32 for k <= 16 or m*n <= 2^16); the
                                rest below m 128, n 64, k 64 run as QFX_PREC_FP32 */
};

/* Kernel classes reported by qfx_profile_read. */
enum qfx_kernel_class {
  QFX_K_PERMUTE = 0,
  QFX_K_CGEMM = 1,        /* MFMA complex GEMM kernels */
  QFX_K_GEMV = 2,
  QFX_K_DOT = 3,
  QFX_K_ACCUMULATE = 4,
  QFX_K_COPY = 5,
  QFX_K_CGEMM_SMALL = 6,  /* scalar complex GEMM (small shapes, FP32_EXACT) */
  QFX_K_NUM_CLASSES = 7
};

/* Message of the last failed call on this host thread ("" if none). */
const char* qfx_last_error(void);

/* ---- context, stream, arena ------------------------------------------- */

/* Creates a context on HIP device `device` with its own stream.  The device
 * arena grows in chunks of at least `arena_chunk_bytes` (0 = 1 GiB).
 * Replaces the host allocations of Tensor (src/tensor.cpp:87-155) and the
 * scratch planning storage of ContractionData::Initialize
 * (src/contraction_utils.cpp:179-218). */
int qfx_context_create(int device, size_t arena_chunk_bytes, qfx_context** out);
int qfx_context_destroy(qfx_context* ctx);

/* Device memory from the arena (256-byte aligned).  Freed blocks are reused
 * by later allocations of the same size class; memory returns to HIP only at
 * qfx_context_destroy. */
int qfx_alloc(qfx_context* ctx, size_t bytes, void** dptr);
int qfx_free(qfx_context* ctx, void* dptr);
/* Bytes currently held from HIP by the arena, and bytes handed out. */
int qfx_arena_stats(qfx_context* ctx, size_t* reserved, size_t* in_use);

/* Copies; h2d/d2h are synchronous with respect to the host buffer. */
int qfx_copy_h2d(qfx_context* ctx, void* dst, const void* src, size_t bytes);
int qfx_copy_d2h(qfx_context* ctx, void* dst, const void* src, size_t bytes);
/* Host-to-device copy that returns at once: the bytes are copied into a
 * pinned staging area of the context (so `src` may be reused right away).
 * The staged copies are batched and moved by one launch, which every later
 * call that puts work on the stream, reads device memory or waits for the
 * stream issues first, so they land in call order before any use of `dst`.
 * Copies above 4 MiB take the synchronous path.  For the many small gate
 * tensors of a network build (Tensor's data constructor,
 * src/tensor.cpp:170-188). */
int qfx_copy_h2d_async(qfx_context* ctx, void* dst, const void* src, size_t bytes);
/* Replaces std::copy of Tensor::_copy / project (src/tensor.cpp:107-139,
 * :307-315). */
int qfx_copy_d2d(qfx_context* ctx, void* dst, const void* src, size_t bytes);
/* Projection of a tensor viewed as [pre][dim][post] onto index values:
 * dst[a][j][r] = src[a][picks[j]][r], j < npicks (one launch per 64 values).
 * Replaces the copy loop of Tensor::project (src/tensor.cpp:280-316) and its
 * use on cut indices in ContractGrid (src/contraction_utils.cpp:296-310). */
int qfx_project_c64(qfx_context* ctx, const qfx_c64* src, qfx_c64* dst, uint64_t pre,
                    uint64_t dim, uint64_t post, uint64_t npicks, const uint64_t* picks);
/* Zero-fills with a kernel (graph-capture safe). */
int qfx_zero(qfx_context* ctx, void* dst, size_t bytes);
int qfx_synchronize(qfx_context* ctx);

/* ---- the hot-path primitives ------------------------------------------ */

/* Index permutation of a complex64 tensor (pure data movement, exact).
 * src has row-major dims[0..rank-1] (first index slowest); dst receives the
 * tensor whose axis k is src axis perm[k]:
 *   dst[y_0..y_{r-1}] = src[x] with x_{perm[k]} = y_k.
 * src and dst must not overlap.  Replaces Tensor::reorder's data movement
 * (_fast_reorder/_naive_reorder, src/tensor.cpp:377-864). */
int qfx_reorder_c64(qfx_context* ctx, const qfx_c64* src, qfx_c64* dst, int rank,
                    const uint64_t* dims, const int* perm);

/* C[m x n] = A[m x k] . B[k x n], complex64, row-major, alpha=1, beta=0.
 * Replaces cblas_cgemm(RowMajor, NoTrans, NoTrans, m, n, k, 1, A, k, B, n,
 * 0, C, n) in _multiply_MM (src/tensor.cpp:960-980).  k == 0 zero-fills C. */
int qfx_cgemm_c64(qfx_context* ctx, int precision, uint64_t m, uint64_t n, uint64_t k,
                  const qfx_c64* A, const qfx_c64* B, qfx_c64* C);

/* qfx_cgemm_c64 that allocates no device memory: no counterpart in the
 * reference.  For callers that cannot allocate (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  The kernels' scratch
 * (split-K partial sums, operand maxima, FP8 codes) lives in a fixed device
 * array of 16 MiB per GPU (QFX_NOALLOC_SCRATCH_BYTES; each GPU the library
 * runs on holds it until the process ends) instead of the context's growable
 * workspace; a product that needs more scratch than that returns
 * QFX_ERR_UNSUPPORTED before anything is launched for it.  The large tiles
 * follow the fixed rule (the autotuner times nothing here); results are
 * bit-identical to qfx_cgemm_c64's.  Every such call on one GPU shares that
 * GPU's array, which keeps the last call's contents: calls that overlap on
 * one GPU (contexts on different streams) race. */
#define QFX_NOALLOC_SCRATCH_BYTES (16u << 20)
int qfx_cgemm_noalloc_c64(qfx_context* ctx, int precision, uint64_t m, uint64_t n, uint64_t k,
                          const qfx_c64* A, const qfx_c64* B, qfx_c64* C);

/* Independent products C_i = A_i . B_i (row-major, as qfx_cgemm_c64) of any
 * shapes.  In the split precisions the products with k > 16, even k and n
 * and 16-byte aligned operands run as ONE persistent launch of the
 * 256 x 128 large-tile kernel: the launch's blocks walk the tiles of all
 * such products (numbered product by product), so many small or uneven
 * products fill the GPU together; each result is bit-identical to that
 * product computed alone on the 256 x 128 tile (split FP16: one max-abs
 * launch per product precedes the grouped launch).  Products with k <= 16
 * and n <= 16 (any precision) also share one launch, of the exact-FP32 row
 * kernel qfx_cgemm_c64 runs them on (same bits); the problem list reaches
 * the device through the staged uploads.  All other products run one by
 * one through qfx_cgemm_c64.  The C_i must not overlap each other or
 * any A_j, B_j.  Not allowed while recording a graph.  No counterpart in the
 * reference (one cblas_cgemm per pairwise contraction, src/tensor.cpp:973). */
typedef struct qfx_gemm_desc {
  uint64_t m, n, k;
  const qfx_c64* A;
  const qfx_c64* B;
  qfx_c64* C;
} qfx_gemm_desc;
int qfx_cgemm_grouped_c64(qfx_context* ctx, int precision, int count,
                          const qfx_gemm_desc* descs);

/* Index map of one matrix dimension of a tensor whose dimensions are all
 * powers of two.  The matrix index i is split into digits, least
 * significant first: digit j has bits[j] bits and is the tensor axis whose
 * element stride is 2^shift[j], so that
 *   offset(i) = sum_j ((i >> (bits[0] + ... + bits[j-1])) & (2^bits[j] - 1)) << shift[j].
 * ndigits == 0 is the map of a dimension of size 1. */
typedef struct qfx_index_map {
  int ndigits; /* 0..16 */
  unsigned char bits[16];
  unsigned char shift[16];
} qfx_index_map;

/* C[m x n] (row-major) = sum_k A[a_rows(m) + a_cols(k)] . B[b_rows(k) + b_cols(n)]:
 * the contraction of multiply(A, B, C) read straight from the stored index
 * orders of A and B, replacing the reorders of A to (left, common) and of
 * B to (common, right) plus _multiply_MM (src/tensor.cpp:1100-1128 and
 * :960-980).  m, n, k are the products of the maps' digit sizes.
 * Supported: the split precisions when m, n, k are large enough for the
 * MFMA tiles, any precision when k <= 16, and the non-exact precisions for
 * at most 8 x 8 outputs with k >= 1024; the operand read along its
 * contiguous axis needs digit 0 of that map at stride 1 with at least 3
 * bits.  Otherwise returns QFX_ERR_UNSUPPORTED without touching C (the
 * caller then reorders and calls qfx_cgemm_c64). */
int qfx_cgemm_mapped_c64(qfx_context* ctx, int precision, const qfx_c64* A,
                         const qfx_index_map* a_rows, const qfx_index_map* a_cols,
                         const qfx_c64* B, const qfx_index_map* b_rows,
                         const qfx_index_map* b_cols, qfx_c64* C);

/* Operand scales of QFX_PREC_SPLIT_FP16 across a chain of contractions (no
 * counterpart in the reference, whose FP32 GEMM needs none).  Each large-tile
 * GEMM can write one partial maximum of max(|re|,|im|) per block of its
 * output C into c_max (capacity c_max_cap floats; c_max_n receives the count,
 * 0 when this call wrote none).  A later GEMM that reads that tensor whole as
 * A or B passes those partials as a_max/a_max_n (b_max/b_max_n) instead of
 * having them recomputed by a pass over the operand.  NULL members are
 * ignored. */
typedef struct qfx_scale_io {
  const float* a_max;
  int a_max_n;
  const float* b_max;
  int b_max_n;
  float* c_max;
  int c_max_cap;
  int c_max_n; /* out */
} qfx_scale_io;

/* qfx_cgemm_mapped_c64 with the scale chain above (io may be NULL). */
/* Partial maxima of |re|, |im| of X[0..n) (at most 256 values, written to
 * device memory `out`, *count of them) in the form the producing GEMM's
 * qfx_scale_io.c_max has: a caller that reuses an operand (a grid tensor in
 * every slice) computes them once and passes them as a_max / b_max. */
int qfx_amax_partials_c64(qfx_context* ctx, const qfx_c64* X, uint64_t n, float* out, int cap,
                          int* count);
int qfx_cgemm_mapped_ex_c64(qfx_context* ctx, int precision, const qfx_c64* A,
                            const qfx_index_map* a_rows, const qfx_index_map* a_cols,
                            const qfx_c64* B, const qfx_index_map* b_rows,
                            const qfx_index_map* b_cols, qfx_c64* C, qfx_scale_io* io);

/* trans == 0: y[m] = A[m x k] . x[k]  (cblas_cgemv NoTrans in _multiply_Mv,
 *             src/tensor.cpp:982-1002)
 * trans == 1: y[k] = sum_r x[r] A[r][.], A is [m x k], x has m entries
 *             (cblas_cgemv Trans as called by _multiply_vM,
 *             src/tensor.cpp:1004-1025 and :1144-1151). */
int qfx_cgemv_c64(qfx_context* ctx, int trans, uint64_t m, uint64_t k,
                  const qfx_c64* A, const qfx_c64* x, qfx_c64* y);

/* *out = sum_i x[i] y[i], unconjugated, written to device memory.
 * Replaces cblas_cdotu_sub in _multiply_vv (src/tensor.cpp:1027-1042). */
int qfx_cdotu_c64(qfx_context* ctx, uint64_t n, const qfx_c64* x, const qfx_c64* y,
                  qfx_c64* out);

/* amps[offset + i] += (complex double) src[i], i < n, on the device.
 * Replaces the FP64 accumulation (*amplitudes_)[output_index] += scalar of
 * ContractionData::ContractGrid (src/contraction_utils.cpp:352), extended to
 * a vector of terminal-cut outputs. */
int qfx_accumulate_c128(qfx_context* ctx, const qfx_c64* src, uint64_t n, qfx_c128* amps,
                        uint64_t offset);

/* amps[offset + i] += (complex double) src[b][i] for b = 0 .. nbatch-1 (src is
 * [nbatch][n]), one FP64 addition per b in increasing b: the same sums, in
 * the same order, as nbatch calls of qfx_accumulate_c128.  For a contraction
 * whose bond-cut values are processed as a batch (the cut index kept as a
 * batch index of dimension nbatch); it replaces the per-slice accumulation
 * (*amplitudes_)[output_index] += scalar of ContractionData::ContractGrid
 * (src/contraction_utils.cpp:352) for nbatch slices at once. */
int qfx_accumulate_batched_c128(qfx_context* ctx, const qfx_c64* src, uint64_t n,
                                uint64_t nbatch, qfx_c128* amps, uint64_t offset);

/* ---- recorded sequences (HIP graphs) ------------------------------------- */
/* Between qfx_graph_begin and qfx_graph_end the context's calls are recorded
 * (HIP stream capture) instead of run; qfx_graph_launch replays the whole
 * sequence with one launch, on the same buffers.  Meant for a sequence that
 * repeats with the same pointers, such as the contraction after the cuts of
 * one slice.  The device workspace is reserved at `workspace_bytes` before
 * recording and may not grow while recording; calls that synchronize or
 * allocate device memory from HIP (copies to or from the host,
 * qfx_synchronize, qfx_profile_*, an arena that must grow) fail while
 * recording, the tile autotuner times nothing, and profiling is off.  Zero
 * fills are kernels (no memset nodes).  No counterpart in the reference. */
typedef struct qfx_graph qfx_graph;
int qfx_graph_begin(qfx_context* ctx, size_t workspace_bytes);
int qfx_graph_end(qfx_context* ctx, qfx_graph** out);
int qfx_graph_launch(qfx_context* ctx, qfx_graph* graph);
int qfx_graph_destroy(qfx_graph* graph);

/* ---- tile autotuner ------------------------------------------------------ */
/* The split-precision GEMMs run on one of three large tiles (complex rows x
 * columns): QFX_TILE_256x128, QFX_TILE_128x256 and QFX_TILE_512x32.  By
 * default a fixed rule picks one from m and n.  With the autotuner on, the
 * first product of each (precision, operand form, m, n, k) runs once on
 * every tile that fits (timed with HIP events; this synchronizes the
 * stream), the fastest is kept in the context and used from then on.  Every
 * tile gives bit-identical results (same per-element accumulation order),
 * so tuning changes only speed.  No counterpart in the reference (its
 * cblas_cgemm has no tile choice). */
enum qfx_tile { QFX_TILE_256x128 = 1, QFX_TILE_128x256 = 2, QFX_TILE_512x32 = 3 };
/* Operand forms of the tuning key. */
enum qfx_form { QFX_FORM_ROW_MAJOR = 0, QFX_FORM_A_ALONG_K = 1, QFX_FORM_A_ALONG_M = 2 };
int qfx_autotune_enable(qfx_context* ctx, int on);
/* Entries as 6 values each: precision, form, m, n, k, tile.  Writes at most
 * `cap` entries; *count = number of entries held. */
int qfx_autotune_entries(qfx_context* ctx, uint64_t* out, int cap, int* count);
/* Sets (or, with tile 0, removes) one entry, e.g. from a saved table. */
int qfx_autotune_set(qfx_context* ctx, int precision, int form, uint64_t m, uint64_t n,
                     uint64_t k, int tile);

/* ---- the cost model's network ------------------------------------------ */
/* No counterpart in the reference: the MLP that scores candidate contraction
 * orderings (features from the engine's symbolic plan, labels from its own
 * timings).  64 -> 256 -> 256 -> 1 with ReLU, BF16 operands (round to nearest
 * even) and FP32 accumulation.  params: one FP32 buffer of QFX_MLP_PARAMS
 * values: W1[64][256], b1[256], W2[256][256], b2[256], w3[256], b3[1]
 * (W[i][j] maps input i to output j).  X is [batch][64] FP32, row-major. */
#define QFX_MLP_IN 64
#define QFX_MLP_HIDDEN 256
#define QFX_MLP_PARAMS 82689

/* y[b] = network(X[b]), b < batch (batch <= 2^20). */
int qfx_mlp_forward(qfx_context* ctx, uint64_t batch, const float* X, const float* params,
                    float* y);
/* Forward + backward of the mean squared error against target[batch]:
 * grads (QFX_MLP_PARAMS FP32, same layout as params) and *loss (one FP32 in
 * device memory) = mean (y - target)^2. */
int qfx_mlp_backward(qfx_context* ctx, uint64_t batch, const float* X, const float* target,
                     const float* params, float* grads, float* loss);
/* One Adam update of params (FP32 master copy) with moments m, v; step >= 1. */
int qfx_mlp_adam(qfx_context* ctx, float* params, const float* grads, float* m, float* v, int step,
                 float lr, float beta1, float beta2, float eps);

/* ---- timing ------------------------------------------------------------ */

/* When enabled, every kernel the context launches is bracketed by a pair of
 * HIP events on the context stream, and qfx_profile_read returns per-class
 * launch counts, summed event time and summed algorithmic bytes and flops
 * (permute: 16 B per element; cgemm: 8*m*n*k flop and 8*(mk+kn+mn) B). */
int qfx_profile_enable(qfx_context* ctx, int on);
int qfx_profile_reset(qfx_context* ctx);
/* Synchronizes the stream, then fills the four arrays (QFX_K_NUM_CLASSES
 * entries each; any may be NULL). */
int qfx_profile_read(qfx_context* ctx, uint64_t* launches, double* ms, double* bytes,
                     double* flops);

/* Wall time of stream work: event at start, event at stop (synchronizes). */
int qfx_timer_start(qfx_context* ctx);
int qfx_timer_stop(qfx_context* ctx, float* ms);

#ifdef __cplusplus
}
#endif

#endif /* QFX_H_ */
