// Grid contraction: operation IR, ordering parser, validation and the
// device-resident contraction interpreter.
//
// Interface of the reference's src/contraction_utils.h:21-220.  The
// interpreter keeps the reference's semantics (expand / cut / merge, cut
// recursion, FP64 amplitude accumulation, output_index arithmetic) but runs
// on the GPU:
//   * tensors never leave the device arena; grid tensors and projections
//     are zero-copy views, and operands are reordered out of place;
//   * a normal cut iterates its values and sums into the same amplitudes;
//   * a terminal cut (one tensor: an output qubit) keeps its output index
//     open instead of re-running the rest of the ordering once per value,
//     so one pass yields the whole vector of outputs of that region; the
//     vector is scattered to the amplitudes in the reference's
//     output_index order;
//   * amplitudes are summed in FP64 on the device across slices.
#ifndef QFLEX_AMD_CONTRACTION_UTILS_H_
#define QFLEX_AMD_CONTRACTION_UTILS_H_

#include <chrono>
#include <deque>
#include <map>
#include <complex>
#include <list>
#include <optional>
#include <string>
#include <unordered_map>
#include <vector>

#include "global.h"
#include "qfx.h"
#include "input.h"
#include "tensor.h"
#include "utils.h"

namespace qflex {

std::string index_name(const std::vector<std::size_t>& p1, const std::vector<std::size_t>& p2);
std::string index_name(const std::vector<std::vector<std::size_t>>& tensors);
std::vector<std::size_t> get_qubit_coords(std::size_t q, std::size_t J);
bool find_grid_coord_in_list(const std::optional<std::vector<std::vector<std::size_t>>>& coord_list,
                             const std::size_t i, const std::size_t j);

struct ExpandPatch {
  ExpandPatch() {}
  ExpandPatch(std::string id, std::vector<std::size_t> tensor) : id(id), tensor(tensor) {}
  std::string id;
  std::vector<std::size_t> tensor;
};

struct CutIndex {
  CutIndex() {}
  CutIndex(std::vector<std::vector<std::size_t>> tensors, std::vector<std::size_t> values = {})
      : tensors(tensors), values(values) {}
  // Two tensors: a bond cut; one tensor: a terminal cut (output value).
  std::vector<std::vector<std::size_t>> tensors;
  // Values to assign; empty means all values of the index.
  std::vector<std::size_t> values;
};

struct MergePatches {
  MergePatches() {}
  MergePatches(std::string source_id, std::string target_id)
      : source_id(source_id), target_id(target_id) {}
  std::string source_id;
  std::string target_id;
};

struct ContractionOperation {
 public:
  ContractionOperation(ExpandPatch expand) : op_type(EXPAND), expand(expand) {}
  ContractionOperation(CutIndex cut) : op_type(CUT), cut(cut) {}
  ContractionOperation(MergePatches merge) : op_type(MERGE), merge(merge) {}

  enum OpType { EXPAND, CUT, MERGE };
  OpType op_type;
  const ExpandPatch expand;
  const CutIndex cut;
  const MergePatches merge;
};

void ordering_data_to_contraction_ordering(const QflexInput& input,
                                           std::list<ContractionOperation>* ordering);

// Work counters of the last ContractGrid (algorithmic, per the textbook
// definitions: cgemm 8*m*n*k flop; a reorder moves 16 bytes per element).
struct ContractionStats {
  std::size_t slices = 0;        // executions of the innermost suffix
  std::size_t multiplies = 0;
  double flops = 0;
  double reorder_elements = 0;
  std::size_t reorders = 0;
  std::size_t mapped_multiplies = 0;  // contractions read without reorder
  std::size_t replayed_passes = 0;    // cut passes run as a recorded graph
  std::size_t planned_bytes = 0;
  double seconds = 0;  // wall time of the contraction (device work finished)
};

// One pairwise contraction of an ordering, from the symbolic pass that
// also plans device memory: C[m x n] = A[m x k] . B[k x n], executed
// `multiplicity` times (the product of the value counts of the bond cuts
// before it); out_rank / out_size describe C.
struct ContractionStep {
  int op = 0;  // ContractionOperation::EXPAND or MERGE
  std::size_t m = 1, n = 1, k = 1;
  // independent products of this shape per execution: the product of the
  // dimensions of the batched cut indices both operands carry (1 without
  // batched cuts)
  std::size_t batch = 1;
  std::size_t multiplicity = 1;
  std::size_t out_rank = 0, out_size = 1;
};

// The symbolic pass over an ordering (index sets only, no data): patch
// sizes, cut buffer sizes and the pairwise-contraction steps.
struct SymbolicPlan {
  std::unordered_map<std::string, std::size_t> patch_max_size;
  std::unordered_map<std::string, std::size_t> side_buffers;
  std::size_t max_grid = 0;
  std::vector<ContractionStep> steps;
};
// `projected`: indices of the grid tensors that are fixed to one value
// before the contraction (the open output indices of a network flattened
// once for several final states); the pass leaves them out.
SymbolicPlan plan_contraction(const std::list<ContractionOperation>& ordering,
                              const std::vector<std::vector<Tensor>>& tensor_grid,
                              const std::vector<std::string>& projected = {});

// Batch size of every bond cut of `ordering` under global::slice_batch, by
// cut index name (at most the cut's value count; 1 = one value at a time).
std::unordered_map<std::string, std::size_t> bond_cut_batches(
    const std::list<ContractionOperation>& ordering,
    const std::vector<std::vector<Tensor>>& tensor_grid);

class ContractionData {
 public:
  static ContractionData Initialize(const std::list<ContractionOperation>& ordering,
                                    std::vector<std::vector<Tensor>>* tensor_grid,
                                    std::vector<std::complex<double>>* amplitudes,
                                    const std::vector<std::string>& projected = {});

  static constexpr char kGeneralSpace[] = "_general_internal_";
  static constexpr char kGeneralSpaceB[] = "_general_internal_b_";
  static constexpr char kResultSpace[] = "_result_internal_";
  static constexpr char kAmplitudeSpace[] = "_amplitudes_internal_";

  void ContractGrid(std::list<ContractionOperation> ordering, std::size_t output_index,
                    std::unordered_map<std::string, bool> active_patches);
  // One index of one grid tensor fixed to a value (a zero-copy view: the
  // index must be the tensor's first).
  struct SiteProjection {
    std::size_t i, j;
    std::string index;
    std::size_t value;
  };
  // The whole ordering once per entry of `states` (each a list of grid
  // tensor projections, e.g. the output bits of one final state), state s
  // writing its amplitudes at output index s; the amplitudes are zeroed
  // once and downloaded once.
  void ContractGridStates(const std::list<ContractionOperation>& ordering,
                          const std::unordered_map<std::string, bool>& active_patches,
                          const std::vector<std::vector<SiteProjection>>& states);

  Tensor& get_scratch(std::string id) { return scratch_[scratch_map_.at(id)]; }

  static std::string cut_copy_name(std::vector<std::vector<std::size_t>> index, std::size_t side) {
    return utils::concat("cut-", index_name(index), ":size-", side);
  }

  static std::string result_space(std::size_t size) { return utils::concat(kResultSpace, size); }

  std::vector<std::string> scratch_list() {
    std::vector<std::string> names;
    for (const auto& p : scratch_map_) names.push_back(p.first);
    return names;
  }

  const ContractionStats& stats() const { return stats_; }

  // A tensor seen by the interpreter: device data plus index metadata.
  struct View {
    const s_type* ptr = nullptr;
    std::vector<std::string> indices;
    std::vector<std::size_t> dims;
    // partial maxima of |re|,|im| written by the GEMM that produced this
    // tensor (split-FP16 scales), or none
    const float* amax = nullptr;
    int amax_n = 0;
    std::size_t size() const;
  };
  struct OpenOutput {
    std::string index;
    std::size_t count;
  };

 private:
  std::vector<Tensor> scratch_;
  // per result/patch buffer: device array for the partial maxima of the
  // tensor a GEMM writes there (pointer, capacity in floats)
  std::vector<Tensor> amax_store_;
  std::unordered_map<const s_type*, std::pair<float*, int>> amax_slots_;
  // split-FP16 partial maxima of grid-tensor operands (constant during one
  // ContractGrid), computed at their first use: (pointer, size) -> (array,
  // count); the address ranges of the grid tensors' buffers
  std::map<std::pair<const s_type*, std::size_t>, std::pair<const float*, int>> grid_amax_;
  std::deque<Tensor> grid_amax_store_;
  std::vector<std::pair<const s_type*, const s_type*>> grid_ranges_;
  void grid_operand_maxima(const View& v, const float** out, int* count);
  std::unordered_map<std::string, std::size_t> scratch_map_;
  std::chrono::steady_clock::time_point start_;
  std::unordered_map<std::string, std::size_t> patch_max_size_;
  std::vector<std::vector<Tensor>>* tensor_grid_ = nullptr;
  std::vector<std::complex<double>>* amplitudes_ = nullptr;

  std::vector<std::vector<View>> grid_views_;
  std::unordered_map<std::string, View> patches_;
  int depth_ = 0;
  ContractionStats stats_;
  void begin_contraction();
  void end_contraction();
  // batch size of each bond cut (by index name), and the cut indices whose
  // current values are a batch (kept open, never summed by a GEMM)
  std::unordered_map<std::string, std::size_t> cut_batches_;
  std::vector<std::string> batch_;
  // device arrays for the split-FP16 partial maxima of whole batched
  // operands (A, B)
  std::vector<Tensor> batch_amax_;
  bool is_batch(const std::string& index) const;
  bool recording_ = false;  // a pass of a cut loop is being recorded
  View contract_batched(const View& A, const View& B, s_type* out,
                        const std::vector<std::string>& batch_common);

  void run(std::list<ContractionOperation>& ordering, std::size_t output_index,
           std::unordered_map<std::string, bool>& active_patches,
           std::vector<OpenOutput> open_outputs);
  bool try_mapped(const s_type* a, const std::vector<std::string>& a_indices,
                  const std::vector<std::size_t>& a_dims, const s_type* b,
                  const std::vector<std::string>& b_indices, const std::vector<std::size_t>& b_dims,
                  const std::vector<std::string>& left, const std::vector<std::string>& common,
                  const std::vector<std::string>& right, s_type* out, qfx_scale_io* io);
  View contract_pair(const View& A, const View& B, s_type* out);
  View project_view(const View& v, const std::string& index, std::size_t value,
                    const std::vector<std::size_t>* gather_values, s_type* gather_buffer,
                    bool copy = false);
  // One bond cut's loop over its values (or batches of values): `groups`
  // lists the values of each pass; with global::record_slices the passes
  // after the first replay one recorded sequence.
  void cut_loop(const std::string& index, const std::vector<std::size_t>& pa,
                const std::vector<std::size_t>& pb, const View& saved_a, const View& saved_b,
                s_type* buf_a, s_type* buf_b, const std::vector<std::vector<std::size_t>>& groups,
                bool batched, const std::list<ContractionOperation>& ordering,
                std::size_t output_index,
                const std::unordered_map<std::string, bool>& active_patches,
                const std::vector<OpenOutput>& open_outputs);
  void emit_output(const View& output, std::size_t output_index,
                   const std::vector<OpenOutput>& open_outputs);
};

void ValidateOrdering(const std::list<ContractionOperation>& ordering);

void ContractGrid(const std::list<ContractionOperation>& ordering,
                  std::vector<std::vector<Tensor>>* tensor_grid,
                  std::vector<std::complex<double>>* amplitudes);

// Statistics of the most recent free-function ContractGrid call.
const ContractionStats& last_contraction_stats();

}  // namespace qflex

#endif
