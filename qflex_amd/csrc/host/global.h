// Process-wide settings of the host (reference src/global.h:9-15).
#ifndef QFLEX_AMD_GLOBAL_H_
#define QFLEX_AMD_GLOBAL_H_

#include <cstddef>
#include <vector>

namespace qflex::global {

// Verbosity of timing / allocation messages on stderr.
inline int verbose = 0;

// Largest device footprint a contraction may plan (default 1 GiB, as the
// reference).  Checked by ContractionData::Initialize.
inline std::size_t memory_limit = std::size_t(1) << 30;

// Arithmetic form of every complex GEMM issued by the host; one of the
// qfx_precision values of include/qfx.h (0 = FP32 MFMA).
inline int precision = 0;

// When set, contractions whose operands are not yet in (left, common) /
// (common, right) order are read in their stored order by the mapped cGEMM
// (no reorder) wherever a kernel supports the form.
inline bool mapped_operands = true;

// Bond cuts whose values are contracted a batch at a time: entry i is the
// batch size of the i-th bond cut of the ordering (the last entry also
// applies to later cuts).  Empty, or 1, is the reference's slice loop (one
// value at a time).  A batch keeps the cut index, restricted to the batch's
// values, as a batch index: it is never summed inside a GEMM, and the
// amplitudes of the batch's slices are summed in FP64 as one per slice.
inline std::vector<std::size_t> slice_batch;

// Record the passes of the innermost bond cut's loop: the first pass runs
// eagerly, the second is recorded as one graph, later passes replay it
// (the projections of the cut tensors are copied into fixed buffers).
inline bool record_slices = false;

// Set while the library's per-launch profiling is on (recorded passes would
// escape it, so nothing is recorded then).
inline bool profiling = false;

}  // namespace qflex::global

#endif
