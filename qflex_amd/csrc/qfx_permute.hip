// Index permutation ("reorder") of complex64 tensors.
//
// Tensor::reorder (reference src/tensor.cpp:866-911) moves data with up to
// five left/right block passes over the tensor (_fast_reorder,
// src/tensor.cpp:494-710).  Here a permutation is one pass:
//
//  * axes of size 1 are dropped and axes that stay adjacent are fused;
//  * if every fused dimension is a power of two the permutation is a
//    permutation of the bits of the linear index.  A tile is the set of
//    elements that differ only in a chosen set T of index bits:
//    T = the lowest input bits (so a tile's loads are contiguous runs)
//    plus the lowest output bits not already in T (so its stores are too).
//    Each workgroup stages tiles through LDS: load in input order, store in
//    output order.  The bits outside T select the tile.
//  * otherwise a generic kernel computes each element's source offset.
#include <algorithm>
#include <numeric>

#include "qfx_internal.h"

namespace {

constexpr int kMaxBits = 48;
constexpr int kMaxRank = 24;

struct BitPermParams {
  int tbits;   // bits inside a tile
  int nfree;   // bits selecting the tile
  uint64_t ntiles;
  uint8_t tin[16];   // tile bit k (input order) -> input bit position
  uint8_t trank[16]; // tile bit k (input order) -> LDS position bit
  uint8_t tout[16];  // tile bit k (output order) -> output bit position
  uint8_t fin[kMaxBits];
  uint8_t fout[kMaxBits];
};

__device__ __forceinline__ uint32_t lds_slot(uint32_t s) { return s + (s >> 5); }

// Offsets of tile element e: the input offset, output offset and LDS
// position are OR-sums over e's set bits.  Elements e = t + e_j with t
// (this thread's bits) and e_j (the unrolled iteration's bits, the same in
// every lane) disjoint split each offset into a per-lane term computed once
// and a per-iteration term that lives in scalar registers.  lds_slot is
// additive over such a split: the low five bits of the two terms are disjoint,
// so their sum carries nothing past bit 4.
struct TileOffsets {
  uint64_t in, out;
  uint32_t slot;
};

template <int TBITS>
__device__ __forceinline__ TileOffsets tile_offsets(const BitPermParams& p, uint32_t e, int k0) {
  TileOffsets r{0, 0, 0};
  for (int k = k0; k < TBITS; ++k)
    if ((e >> k) & 1u) {
      r.in |= uint64_t(1) << p.tin[k];
      r.out |= uint64_t(1) << p.tout[k];
      r.slot |= 1u << p.trank[k];
    }
  return r;
}

__device__ __forceinline__ void tile_base(const BitPermParams& p, uint64_t t, uint64_t& bin,
                                          uint64_t& bout) {
  bin = 0;
  bout = 0;
  for (int k = 0; k < p.nfree; ++k)
    if ((t >> k) & 1u) {
      bin |= uint64_t(1) << p.fin[k];
      bout |= uint64_t(1) << p.fout[k];
    }
}

template <int TBITS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void permute_bits_kernel(const float2* __restrict__ src,
                                                             float2* __restrict__ dst,
                                                             BitPermParams p) {
  constexpr int TILE = 1 << TBITS;
  constexpr int PER = TILE / BLOCK;
  static_assert(PER >= 1, "tile smaller than block");
  __shared__ float2 lds[TILE + TILE / 32];

  // Tile-local element e = tid + j*BLOCK.
  const uint32_t tid = threadIdx.x;
  const TileOffsets lane = tile_offsets<TBITS>(p, tid, 0);
  const uint32_t lane_slot = lds_slot(lane.slot);

  for (uint64_t t = blockIdx.x; t < p.ntiles; t += gridDim.x) {
    uint64_t bin, bout;
    tile_base(p, t, bin, bout);
    float2 v[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j)
      v[j] = src[bin + (tile_offsets<TBITS>(p, uint32_t(j) * BLOCK, 0).in | lane.in)];
#pragma unroll
    for (int j = 0; j < PER; ++j)
      lds[lds_slot(tile_offsets<TBITS>(p, uint32_t(j) * BLOCK, 0).slot) + lane_slot] = v[j];
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const uint32_t ej = uint32_t(j) * BLOCK;
      dst[bout + (tile_offsets<TBITS>(p, ej, 0).out | lane.out)] = lds[lds_slot(ej) + lds_slot(tid)];
    }
    __syncthreads();
  }
}

// Pair variant: each thread moves element pairs, 16 bytes per global access.
// Input bit 0 is tile bit 0, so tile elements e, e+1 (input order) are
// adjacent in memory; output bit 0 is the lowest output bit of the tile, so
// output-order positions o, o+1 are adjacent too.  Needs TBITS >= 2.
template <int TBITS, int BLOCK>
__global__ __launch_bounds__(BLOCK) void permute_pairs_kernel(const float2* __restrict__ src,
                                                              float2* __restrict__ dst,
                                                              BitPermParams p) {
  constexpr int TILE = 1 << TBITS;
  constexpr int PER = TILE / (2 * BLOCK);  // pairs per thread
  static_assert(PER >= 1, "tile smaller than two elements per thread");
  __shared__ float2 lds[TILE + TILE / 32];

  // Pair tid + j*BLOCK covers tile elements e = 2*(tid + j*BLOCK) and e + 1.
  const uint32_t tid = threadIdx.x;
  const TileOffsets lane = tile_offsets<TBITS>(p, 2u * tid, 1);
  const uint32_t slot0 = lds_slot(lane.slot), slot1 = lds_slot(lane.slot | (1u << p.trank[0]));
  const uint32_t rslot = lds_slot(2u * tid);

  for (uint64_t t = blockIdx.x; t < p.ntiles; t += gridDim.x) {
    uint64_t bin, bout;
    tile_base(p, t, bin, bout);
    float4 v[PER];
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const uint64_t off = tile_offsets<TBITS>(p, 2u * uint32_t(j) * BLOCK, 1).in | lane.in;
      v[j] = *reinterpret_cast<const float4*>(src + bin + off);
    }
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const uint32_t sj = lds_slot(tile_offsets<TBITS>(p, 2u * uint32_t(j) * BLOCK, 1).slot);
      lds[sj + slot0] = make_float2(v[j].x, v[j].y);
      lds[sj + slot1] = make_float2(v[j].z, v[j].w);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const uint32_t ej = 2u * uint32_t(j) * BLOCK;
      const uint32_t r = lds_slot(ej) + rslot;
      const float2 x = lds[r];
      const float2 y = lds[r + 1];
      const uint64_t off = tile_offsets<TBITS>(p, ej, 1).out | lane.out;
      *reinterpret_cast<float4*>(dst + bout + off) = make_float4(x.x, x.y, y.x, y.y);
    }
    __syncthreads();
  }
}

struct GenericParams {
  int rank;
  uint64_t total;
  uint64_t out_dims[kMaxRank];
  uint64_t in_stride_of_out[kMaxRank];  // input stride of output axis k
};

__global__ void permute_generic_kernel(const float2* __restrict__ src, float2* __restrict__ dst,
                                       GenericParams p) {
  uint64_t i = uint64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint64_t stride = uint64_t(gridDim.x) * blockDim.x;
  for (; i < p.total; i += stride) {
    uint64_t rem = i, off = 0;
    for (int k = p.rank - 1; k >= 0; --k) {
      uint64_t d = p.out_dims[k];
      uint64_t x = rem % d;
      rem /= d;
      off += x * p.in_stride_of_out[k];
    }
    dst[i] = src[off];
  }
}

bool is_pow2(uint64_t x) { return x && !(x & (x - 1)); }
int log2u(uint64_t x) { return 63 - __builtin_clzll(x); }

template <int TBITS, int BLOCK>
void launch_bits(const BitPermParams& p, const float2* src, float2* dst, hipStream_t s) {
  uint64_t g = std::min<uint64_t>(p.ntiles, 256 * 16);
  permute_bits_kernel<TBITS, BLOCK><<<unsigned(g), BLOCK, 0, s>>>(src, dst, p);
}

template <int TBITS, int BLOCK>
void launch_pairs(const BitPermParams& p, const float2* src, float2* dst, hipStream_t s) {
  uint64_t g = std::min<uint64_t>(p.ntiles, 256 * 8);
  permute_pairs_kernel<TBITS, BLOCK><<<unsigned(g), BLOCK, 0, s>>>(src, dst, p);
}

}  // namespace

extern "C" int qfx_reorder_c64(qfx_context* ctx, const qfx_c64* src, qfx_c64* dst, int rank,
                               const uint64_t* dims, const int* perm) {
  QFX_CHECK_CTX(ctx);
  if (rank < 0 || rank > 64) return qfx_set_error(QFX_ERR_INVALID, "qfx_reorder_c64: bad rank");
  if (rank > 0 && (!dims || !perm))
    return qfx_set_error(QFX_ERR_INVALID, "qfx_reorder_c64: null dims/perm");
  std::vector<int> seen(rank, 0);
  uint64_t total = 1;
  for (int k = 0; k < rank; ++k) {
    if (perm[k] < 0 || perm[k] >= rank || seen[perm[k]]++)
      return qfx_set_error(QFX_ERR_INVALID, "qfx_reorder_c64: perm is not a permutation");
    if (dims[k] == 0) return QFX_OK;
    total *= dims[k];
  }
  if (!src || !dst) return qfx_set_error(QFX_ERR_INVALID, "qfx_reorder_c64: null data pointer");
  if (src == dst) return qfx_set_error(QFX_ERR_INVALID, "qfx_reorder_c64: src and dst alias");
  QFX_EXTENT(ctx, src, 8.0 * double(total), "qfx_reorder_c64: src");
  QFX_EXTENT(ctx, dst, 8.0 * double(total), "qfx_reorder_c64: dst");

  // Drop unit axes, then fuse axes that remain adjacent in the output.
  std::vector<int> keep_map(rank, -1);
  std::vector<uint64_t> d1;
  for (int a = 0; a < rank; ++a)
    if (dims[a] != 1) {
      keep_map[a] = int(d1.size());
      d1.push_back(dims[a]);
    }
  std::vector<int> p1;
  for (int k = 0; k < rank; ++k)
    if (keep_map[perm[k]] >= 0) p1.push_back(keep_map[perm[k]]);
  // runs of consecutive input axes in output order
  std::vector<std::pair<int, int>> runs;  // (first input axis, length)
  for (size_t k = 0; k < p1.size(); ++k) {
    if (!runs.empty() && runs.back().first + runs.back().second == p1[k])
      runs.back().second++;
    else
      runs.push_back({p1[k], 1});
  }
  const int r = int(runs.size());
  if (r <= 1) return qfx_copy_d2d(ctx, dst, src, total * sizeof(qfx_c64));
  // fused input axes sorted by first input axis
  std::vector<int> order(r);
  std::iota(order.begin(), order.end(), 0);
  std::sort(order.begin(), order.end(),
            [&](int x, int y) { return runs[x].first < runs[y].first; });
  std::vector<int> run_to_in(r);
  for (int i = 0; i < r; ++i) run_to_in[order[i]] = i;
  std::vector<uint64_t> fdims(r);  // fused input dims (input order)
  for (int i = 0; i < r; ++i) {
    uint64_t d = 1;
    for (int a = runs[order[i]].first; a < runs[order[i]].first + runs[order[i]].second; ++a)
      d *= d1[a];
    fdims[i] = d;
  }
  std::vector<int> fperm(r);  // output axis k -> fused input axis
  for (int k = 0; k < r; ++k) fperm[k] = run_to_in[k];

  QfxLaunchScope scope(ctx, QFX_K_PERMUTE, 16.0 * double(total), 0);
  const float2* s2 = reinterpret_cast<const float2*>(src);
  float2* d2 = reinterpret_cast<float2*>(dst);

  bool pow2 = std::all_of(fdims.begin(), fdims.end(), is_pow2);
  if (pow2) {
    // bit positions: input axis a occupies bits [s_a, s_a + b_a)
    std::vector<int> bits(r), in_start(r), out_start(r);
    int nbits = 0;
    for (int a = r - 1; a >= 0; --a) {
      bits[a] = log2u(fdims[a]);
      in_start[a] = nbits;
      nbits += bits[a];
    }
    if (nbits > kMaxBits) return qfx_set_error(QFX_ERR_UNSUPPORTED, "tensor too large");
    int ob_acc = 0;
    for (int k = r - 1; k >= 0; --k) {
      out_start[fperm[k]] = ob_acc;
      ob_acc += bits[fperm[k]];
    }
    std::vector<int> ob(nbits), ib_of_ob(nbits);
    for (int a = 0; a < r; ++a)
      for (int j = 0; j < bits[a]; ++j) ob[in_start[a] + j] = out_start[a] + j;
    for (int i = 0; i < nbits; ++i) ib_of_ob[ob[i]] = i;

    // Tiles of 2^12 elements (6 lowest input bits + 6 lowest output bits)
    // moved as 16-byte pairs for tensors of >= 2^16 elements; 2^10-element
    // tiles of single elements below that.
    const bool pairs = nbits >= 16 && (reinterpret_cast<uintptr_t>(src) % 16 == 0) &&
                       (reinterpret_cast<uintptr_t>(dst) % 16 == 0);
    const int tile_target = pairs ? 12 : 10;
    int tbits = std::min(nbits, tile_target);
    int abits = std::min(tbits, pairs ? 6 : 5);
    std::vector<char> in_t(nbits, 0);
    for (int i = 0; i < abits; ++i) in_t[i] = 1;
    int have = abits;
    for (int o = 0; o < nbits && have < tbits; ++o) {
      int ib = ib_of_ob[o];
      if (!in_t[ib]) {
        in_t[ib] = 1;
        ++have;
      }
    }
    // tbits may still exceed have only if nbits < tbits (not possible)
    BitPermParams p{};
    p.tbits = tbits;
    std::vector<int> tin, touts;
    for (int i = 0; i < nbits; ++i)
      if (in_t[i]) tin.push_back(i);
    for (int i = 0; i < nbits; ++i)
      if (in_t[i]) touts.push_back(ob[i]);
    std::sort(touts.begin(), touts.end());
    for (int k = 0; k < tbits; ++k) {
      p.tin[k] = uint8_t(tin[k]);
      p.trank[k] = uint8_t(std::lower_bound(touts.begin(), touts.end(), ob[tin[k]]) - touts.begin());
      p.tout[k] = uint8_t(touts[k]);
    }
    int nf = 0;
    for (int i = 0; i < nbits; ++i)
      if (!in_t[i]) {
        p.fin[nf] = uint8_t(i);
        p.fout[nf] = uint8_t(ob[i]);
        ++nf;
      }
    p.nfree = nf;
    p.ntiles = uint64_t(1) << nf;
    if (pairs) {
      launch_pairs<12, 256>(p, s2, d2, ctx->stream);
    } else switch (tbits) {
      case 10: launch_bits<10, 256>(p, s2, d2, ctx->stream); break;
      case 9: launch_bits<9, 256>(p, s2, d2, ctx->stream); break;
      case 8: launch_bits<8, 256>(p, s2, d2, ctx->stream); break;
      case 7: launch_bits<7, 128>(p, s2, d2, ctx->stream); break;
      case 6: launch_bits<6, 64>(p, s2, d2, ctx->stream); break;
      case 5: launch_bits<5, 32>(p, s2, d2, ctx->stream); break;
      case 4: launch_bits<4, 16>(p, s2, d2, ctx->stream); break;
      case 3: launch_bits<3, 8>(p, s2, d2, ctx->stream); break;
      case 2: launch_bits<2, 4>(p, s2, d2, ctx->stream); break;
      default: launch_bits<1, 2>(p, s2, d2, ctx->stream); break;
    }
  } else {
    if (r > kMaxRank) return qfx_set_error(QFX_ERR_UNSUPPORTED, "rank too large");
    GenericParams p{};
    p.rank = r;
    p.total = total;
    std::vector<uint64_t> in_stride(r);
    uint64_t st = 1;
    for (int a = r - 1; a >= 0; --a) {
      in_stride[a] = st;
      st *= fdims[a];
    }
    for (int k = 0; k < r; ++k) {
      p.out_dims[k] = fdims[fperm[k]];
      p.in_stride_of_out[k] = in_stride[fperm[k]];
    }
    uint64_t g = std::min<uint64_t>((total + 255) / 256, 65536);
    permute_generic_kernel<<<unsigned(g), 256, 0, ctx->stream>>>(s2, d2, p);
  }
  QFX_HIP(hipGetLastError());
  return QFX_OK;
}
