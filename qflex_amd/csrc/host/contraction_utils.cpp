// Device-resident grid contraction interpreter (see contraction_utils.h).
// Parser, validation and planning follow reference
// src/contraction_utils.cpp:16-221 and :358-610; the interpreter keeps the
// semantics of :223-354.
#include "contraction_utils.h"

#include <algorithm>
#include <chrono>
#include <iostream>
#include <cctype>
#include <cstdlib>
#include <map>
#include <set>
#include <sstream>
#include <unordered_set>

#include "device.h"
#include "errors.h"

namespace qflex {

namespace {

ContractionStats g_last_stats;

std::size_t product(const std::vector<std::size_t>& d) {
  std::size_t p = 1;
  for (auto x : d) p *= x;
  return p;
}

}  // namespace

const ContractionStats& last_contraction_stats() { return g_last_stats; }

std::size_t ContractionData::View::size() const { return product(dims); }

// ------------------------------------------------------------ naming

std::string index_name(const std::vector<std::size_t>& p1, const std::vector<std::size_t>& p2) {
  if (p1.size() == 2 && p2.size() == 2)
    return utils::concat("(", p1[0], ",", p1[1], "),(", p2[0], ",", p2[1], ")");
  if (p1.size() == 3 && p2.size() == 3)
    return utils::concat("(", p1[0], ",", p1[1], ",", p1[2], "),(", p2[0], ",", p2[1], ",", p2[2],
                         ")");
  if (p1.size() == 2 && p2.empty()) return utils::concat("(", p1[0], ",", p1[1], "),(o)");
  std::stringstream ss;
  ss << "Failed to construct tensor name with the following vectors: p1 = [";
  for (const auto& p : p1) ss << p << ',';
  ss << "] and p2 = [";
  for (const auto& p : p2) ss << p << ',';
  ss << "].";
  throw ERROR_MSG(ss.str());
}

std::string index_name(const std::vector<std::vector<std::size_t>>& tensors) {
  if (tensors.size() == 2) {
    try {
      return index_name(tensors.at(0), tensors.at(1));
    } catch (const std::string& err_msg) {
      throw ERROR_MSG("Failed to call index_name(). Error:\n\t[", err_msg, "]");
    }
  }
  if (tensors.size() == 1) {
    try {
      return index_name(tensors.at(0), {});
    } catch (const std::string& err_msg) {
      throw ERROR_MSG("Failed to call index_name(). Error:\n\t[", err_msg, "]");
    }
  }
  throw ERROR_MSG("Failed to construct tensor name with input tensors size: ", tensors.size(),
                  ".");
}

std::vector<std::size_t> get_qubit_coords(std::size_t q, std::size_t J) {
  std::size_t i = q / J;
  return {i, q - i * J};
}

bool find_grid_coord_in_list(const std::optional<std::vector<std::vector<std::size_t>>>& coord_list,
                             const std::size_t i, const std::size_t j) {
  return coord_list.has_value() &&
         std::find(coord_list.value().begin(), coord_list.value().end(),
                   std::vector<std::size_t>({i, j})) != coord_list.value().end();
}

// ------------------------------------------------------------ parsing

namespace {

// The words of an ordering line: the operation, then its arguments.
std::vector<std::string> words_of(const std::string& line) {
  std::vector<std::string> w;
  std::size_t i = 0;
  while (i < line.size()) {
    while (i < line.size() && std::isspace(static_cast<unsigned char>(line[i]))) ++i;
    const std::size_t b = i;
    while (i < line.size() && !std::isspace(static_cast<unsigned char>(line[i]))) ++i;
    if (i > b) w.push_back(line.substr(b, i - b));
  }
  return w;
}

// A qubit number as stream extraction reads it: a missing word gives
// `missing`, a word without a leading integer gives 0, otherwise its leading
// integer ("7", "7x" -> 7).
long qubit_word(const std::vector<std::string>& w, std::size_t at, long missing) {
  if (at >= w.size()) return missing;
  const char* s = w[at].c_str();
  char* end = nullptr;
  const long v = std::strtol(s, &end, 10);
  return end == s ? 0 : v;
}

// "(v1,v2,...)" with digits and commas only, e.g. "()" or "(0,1,3)"; empty
// pieces read as 0.  False when the word has another form.
bool cut_values(const std::string& word, std::vector<std::size_t>* values) {
  if (word.size() < 2 || word.front() != '(' || word.back() != ')') return false;
  for (std::size_t i = 1; i + 1 < word.size(); ++i)
    if (!std::isdigit(static_cast<unsigned char>(word[i])) && word[i] != ',') return false;
  if (word.size() == 2) return true;
  std::size_t v = 0;
  for (std::size_t i = 1; i < word.size(); ++i) {
    if (word[i] == ',' || word[i] == ')') {
      values->push_back(v);
      v = 0;
    } else {
      v = 10 * v + std::size_t(word[i] - '0');
    }
  }
  return true;
}

}  // namespace

void ordering_data_to_contraction_ordering(const QflexInput& input,
                                           std::list<ContractionOperation>* ordering) {
  if (ordering == nullptr) throw ERROR_MSG("Ordering must be non-null.");
  const std::size_t sites = input.grid.I * input.grid.J;
  for (const auto& line : input.ordering.instructions) {
    if (line.empty() || line[0] == '#') continue;
    auto fail = [&line](const std::string& msg) {
      return ERROR_MSG("Parsing failed on line: '", line, "' with error: ", msg);
    };
    // Grid coordinates of qubit q, which must be on the grid and active;
    // `which` names the qubit in the error text ("", "1 ", "2 ").
    auto site = [&](long q, const char* which) {
      if (q < 0 || std::size_t(q) >= sites)
        throw fail(utils::concat("Index ", q, " must be within grid boundaries."));
      const std::vector<std::size_t> at = get_qubit_coords(std::size_t(q), input.grid.J);
      if (find_grid_coord_in_list(input.grid.qubits_off, at[0], at[1]))
        throw fail(utils::concat("Index ", which, "must specify an active qubit."));
      return at;
    };
    const std::vector<std::string> w = words_of(line);
    const std::string op = w.empty() ? std::string() : w[0];
    const std::string arg1 = w.size() > 1 ? w[1] : std::string();
    if (op == "expand") {
      ordering->emplace_back(ExpandPatch(arg1, site(qubit_word(w, 2, -1), "")));
    } else if (op == "merge") {
      ordering->emplace_back(MergePatches(arg1, w.size() > 2 ? w[2] : std::string()));
    } else if (op == "cut") {
      std::vector<std::size_t> values;
      if (!cut_values(arg1, &values))
        throw fail("Cut values must be comma-separated ints, e.g. (0,1,3).");
      const long q1 = qubit_word(w, 2, -1);
      auto p1 = site(q1, "1 ");
      if (w.size() <= 3) {  // a terminal cut: the output index of one qubit
        ordering->emplace_back(CutIndex({p1}, values));
        continue;
      }
      const long q2 = qubit_word(w, 3, -1);
      auto p2 = site(q2, "2 ");
      // the bond is named from its lower-numbered end
      if (q2 < q1) std::swap(p1, p2);
      ordering->emplace_back(CutIndex({p1, p2}, values));
    } else {
      throw fail("Received an invalid operation in config.");
    }
  }
  ValidateOrdering(*ordering);
}

// ------------------------------------------------------------ validation

// Replays the ordering's patch bookkeeping and collects every violation in
// one message (thrown as a plain string, like the reference's
// src/contraction_utils.cpp:534-610, whose message texts are kept):
//  * a grid tensor is contracted at most once and a bond is cut at most once;
//  * a patch grown before a cut is frozen: after the cut nothing may be added
//    to it and nothing merged into it (it is reused across the cut values);
//  * a patch merged into another one is gone: it can be neither merged again
//    nor grown.
void ValidateOrdering(const std::list<ContractionOperation>& ordering) {
  enum : unsigned { kGrown = 1, kFrozen = 2, kGone = 4 };
  std::map<std::string, unsigned> patch;  // patch id -> state bits
  std::set<std::pair<std::size_t, std::size_t>> contracted;
  std::set<std::string> cut_bonds;
  std::ostringstream problems;
  for (const auto& op : ordering) {
    if (op.op_type == ContractionOperation::CUT) {
      for (auto& p : patch)
        if (p.second & kGrown) p.second |= kFrozen;
      std::string bond;
      try {
        bond = index_name(op.cut.tensors);
      } catch (const std::string& err_msg) {
        throw ERROR_MSG("Failed to during CUT. Error:\n\t[", err_msg, "]");
      }
      if (!cut_bonds.insert(bond).second)
        problems << "\nIndex " << bond << " is cut multiple times.";
    } else if (op.op_type == ContractionOperation::EXPAND) {
      const std::size_t i = op.expand.tensor[0], j = op.expand.tensor[1];
      unsigned& st = patch[op.expand.id];
      for (unsigned bit : {unsigned(kFrozen), unsigned(kGone)})
        if (st & bit)
          problems << "\nTensor at (" << i << "," << j << ") is added to non-empty patch "
                   << op.expand.id << " after a cut.";
      if (!contracted.insert({i, j}).second)
        problems << "\nTensor (" << i << "," << j << ") is contracted multiple times.";
      st |= kGrown;
    } else {
      const std::string& from = op.merge.source_id;
      const std::string& into = op.merge.target_id;
      if (patch[from] & kGone) problems << "\nPatch " << from << " is merged multiple times.";
      if (patch[into] & kFrozen)
        problems << "\nPatch " << from << " is merged into non-empty patch " << into
                 << " after a cut.";
      patch[from] |= kGone;
      patch[into] |= kGrown;
    }
  }
  if (!problems.str().empty()) throw problems.str();
}

// ------------------------------------------------------------ planning

std::unordered_map<std::string, std::size_t> bond_cut_batches(
    const std::list<ContractionOperation>& ordering,
    const std::vector<std::vector<Tensor>>& tensor_grid) {
  std::unordered_map<std::string, std::size_t> out;
  std::size_t ordinal = 0;
  for (const auto& op : ordering) {
    if (op.op_type != ContractionOperation::CUT || op.cut.tensors.size() != 2) continue;
    const std::size_t want =
        global::slice_batch.empty()
            ? 1
            : global::slice_batch[std::min(ordinal, global::slice_batch.size() - 1)];
    ++ordinal;
    std::string name;
    try {
      name = index_name(op.cut.tensors);
    } catch (const std::string&) {
      continue;  // reported by the planner and the interpreter
    }
    std::size_t count = op.cut.values.size();
    if (count == 0) {
      const auto& t = op.cut.tensors[0];
      if (t.size() < 2 || t[0] >= tensor_grid.size() || t[1] >= tensor_grid[t[0]].size())
        continue;
      const auto& dims = tensor_grid[t[0]][t[1]].get_index_to_dimension();
      auto it = dims.find(name);
      count = it == dims.end() ? 1 : it->second;
    }
    out[name] = std::max<std::size_t>(1, std::min(want, count));
  }
  return out;
}

SymbolicPlan plan_contraction(const std::list<ContractionOperation>& ordering,
                              const std::vector<std::vector<Tensor>>& tensor_grid,
                              const std::vector<std::string>& projected) {
  SymbolicPlan plan;
  std::unordered_map<std::string, std::size_t> index_dimension;
  std::vector<std::vector<std::vector<std::string>>> grid_indices(tensor_grid.size());
  for (std::size_t i = 0; i < tensor_grid.size(); ++i) {
    grid_indices[i].resize(tensor_grid[i].size());
    for (std::size_t j = 0; j < tensor_grid[i].size(); ++j) {
      const Tensor& t = tensor_grid[i][j];
      grid_indices[i][j] = _vector_subtraction(t.get_indices(), projected);
      for (const auto& index : t.get_indices())
        index_dimension[index] = t.get_index_to_dimension().at(index);
      if (t.data()) plan.max_grid = std::max(plan.max_grid, t.size());
    }
  }

  // Bond cuts remove their index (and multiply the work after them by
  // their value count); terminal cuts keep the output index open with one
  // entry per requested value.  A bond cut taken a batch of values at a
  // time keeps its index, of the batch's size, as a batch index: present in
  // both operands it is a batch dimension of the product (never summed), and
  // the work after the cut runs once per batch.
  const auto batches = bond_cut_batches(ordering, tensor_grid);
  std::vector<std::string> batch;
  std::unordered_map<std::string, std::vector<std::string>> patch_indices;
  std::size_t multiplicity = 1;
  auto batch_common = [&](const std::vector<std::string>& a, const std::vector<std::string>& b) {
    std::vector<std::string> out;
    for (const auto& x : _vector_intersection(a, b))
      if (std::find(batch.begin(), batch.end(), x) != batch.end()) out.push_back(x);
    return out;
  };
  // indices of the product of a and b: batch indices of both, then the
  // free indices of each
  auto sym_diff = [&](const std::vector<std::string>& a, const std::vector<std::string>& b) {
    return _vector_union(batch_common(a, b),
                         _vector_subtraction(_vector_union(a, b), _vector_intersection(a, b)));
  };
  auto size_of = [&](const std::vector<std::string>& idx) {
    std::size_t s = 1;
    for (const auto& x : idx) s *= index_dimension.at(x);
    return s;
  };
  // C = A . B as contract_pair executes it: left = A \ B, common = A n B,
  // right = B \ A.
  auto record = [&](int op, const std::vector<std::string>& a, const std::vector<std::string>& b) {
    if (a.empty() || b.empty()) return;
    ContractionStep st;
    st.op = op;
    const auto bc = batch_common(a, b);
    st.m = size_of(_vector_subtraction(a, b));
    st.k = size_of(_vector_subtraction(_vector_intersection(a, b), bc));
    st.n = size_of(_vector_subtraction(b, a));
    st.batch = size_of(bc);
    st.multiplicity = multiplicity;
    const auto c = sym_diff(a, b);
    st.out_rank = c.size();
    st.out_size = size_of(c);
    plan.steps.push_back(st);
  };
  for (const auto& op : ordering) {
    switch (op.op_type) {
      case ContractionOperation::EXPAND: {
        auto& pi = patch_indices[op.expand.id];
        const auto& site = grid_indices.at(op.expand.tensor[0]).at(op.expand.tensor[1]);
        record(ContractionOperation::EXPAND, pi, site);
        pi = sym_diff(pi, site);
        auto& mx = plan.patch_max_size[op.expand.id];
        mx = std::max(mx, size_of(pi));
        break;
      }
      case ContractionOperation::CUT: {
        try {
          const std::string cut_index = index_name(op.cut.tensors);
          const std::size_t dim = index_dimension.count(cut_index) ? index_dimension[cut_index] : 1;
          const std::size_t count = op.cut.values.empty() ? dim : op.cut.values.size();
          const std::size_t group =
              op.cut.tensors.size() > 1 && batches.count(cut_index) ? batches.at(cut_index) : 1;
          int side = 0;
          for (const auto& tensor : op.cut.tensors) {
            auto& indices = grid_indices.at(tensor[0]).at(tensor[1]);
            const std::size_t full = tensor_grid[tensor[0]][tensor[1]].size();
            if (op.cut.tensors.size() > 1) {
              if (group == 1) indices = _vector_subtraction(indices, {cut_index});
              plan.side_buffers[ContractionData::cut_copy_name(op.cut.tensors, side)] =
                  full / std::max<std::size_t>(dim, 1) * group;
            } else {
              index_dimension[cut_index] = count;
              plan.side_buffers[ContractionData::cut_copy_name(op.cut.tensors, side)] =
                  full / std::max<std::size_t>(dim, 1) * count;
            }
            ++side;
          }
          if (op.cut.tensors.size() > 1) {
            multiplicity *= (count + group - 1) / group;
            if (group > 1) {
              index_dimension[cut_index] = group;
              batch.push_back(cut_index);
            }
          }
        } catch (const std::string& err_msg) {
          throw ERROR_MSG("Failed during CUT. Error:\n\t[", err_msg, "]");
        }
        break;
      }
      case ContractionOperation::MERGE: {
        auto& pi = patch_indices[op.merge.target_id];
        const auto src = patch_indices[op.merge.source_id];
        record(ContractionOperation::MERGE, src, pi);
        pi = sym_diff(pi, src);
        auto& mx = plan.patch_max_size[op.merge.target_id];
        mx = std::max(mx, size_of(pi));
        break;
      }
    }
  }
  return plan;
}

ContractionData ContractionData::Initialize(const std::list<ContractionOperation>& ordering,
                                            std::vector<std::vector<Tensor>>* tensor_grid,
                                            std::vector<std::complex<double>>* amplitudes,
                                            const std::vector<std::string>& projected) {
  if (tensor_grid == nullptr) throw ERROR_MSG("Tensor grid must be non-null.");
  if (amplitudes == nullptr) throw ERROR_MSG("Amplitude return vector must be non-null.");
  ContractionData data;
  data.tensor_grid_ = tensor_grid;
  data.amplitudes_ = amplitudes;

  SymbolicPlan plan = plan_contraction(ordering, *tensor_grid, projected);
  data.patch_max_size_ = plan.patch_max_size;
  const std::size_t max_grid = plan.max_grid;
  const auto& side_buffers = plan.side_buffers;

  std::size_t max_size = max_grid;
  std::vector<std::size_t> unique_sizes;
  for (const auto& p : data.patch_max_size_) {
    max_size = std::max(max_size, p.second);
    if (std::find(unique_sizes.begin(), unique_sizes.end(), p.second) == unique_sizes.end())
      unique_sizes.push_back(p.second);
  }
  std::sort(unique_sizes.begin(), unique_sizes.end());
  // Two general buffers (out-of-place reorders of both operands), one
  // result buffer per size class, one buffer per patch, the cut / gather
  // buffers and the FP64 amplitude vector.
  std::size_t allocated = 2 * max_size;
  for (auto s : unique_sizes) allocated += s;
  for (const auto& p : data.patch_max_size_) allocated += p.second;
  for (const auto& p : side_buffers) allocated += p.second;
  allocated += 2 * amplitudes->size();
  const double alloc_size = double(allocated) * sizeof(s_type);
  if (alloc_size > double(global::memory_limit)) {
    throw ERROR_MSG("Required space (", utils::readable_memory_string(alloc_size),
                    ") exceeds memory limit (",
                    utils::readable_memory_string(double(global::memory_limit)),
                    "). Cancelling simulation.");
  }
  if (global::verbose > 0)
    std::cerr << WARN_MSG("Allocating ", utils::readable_memory_string(alloc_size),
                          " for this simulation.")
              << std::endl;
  data.stats_.planned_bytes = std::size_t(alloc_size);

  data.scratch_.reserve(3 + unique_sizes.size() + data.patch_max_size_.size() +
                        side_buffers.size());
  auto add = [&](const std::string& name, std::size_t elements) {
    try {
      data.scratch_.push_back(Tensor({""}, {std::max<std::size_t>(elements, 1)}));
    } catch (const std::string& err_msg) {
      throw ERROR_MSG("Failed to call Tensor(). Error:\n\t[", err_msg, "]");
    }
    data.scratch_map_[name] = data.scratch_.size() - 1;
  };
  add(kGeneralSpace, max_size);
  add(kGeneralSpaceB, max_size);
  for (auto s : unique_sizes) add(result_space(s), s);
  for (const auto& p : data.patch_max_size_) add(p.first, p.second);
  for (const auto& p : side_buffers) add(p.first, p.second);
  add(kAmplitudeSpace, 2 * amplitudes->size());
  // Partial-maximum arrays beside every buffer a GEMM writes a patch into
  // (one float per output block; a block holds at least 2^14 outputs).
  std::vector<std::string> gemm_outputs;
  for (auto s : unique_sizes) gemm_outputs.push_back(result_space(s));
  for (const auto& p : data.patch_max_size_) gemm_outputs.push_back(p.first);
  data.amax_store_.reserve(gemm_outputs.size());
  for (const auto& name : gemm_outputs) {
    const Tensor& buf = data.scratch_[data.scratch_map_.at(name)];
    const int cap = int(buf.capacity() / 2048 + 64);
    data.amax_store_.push_back(Tensor({""}, {std::size_t(cap + 1) / 2}));
    data.amax_slots_[buf.data()] = {reinterpret_cast<float*>(data.amax_store_.back().data()), cap};
  }
  return data;
}

// ------------------------------------------------------------ execution

// Digit map (see qfx_index_map) of the matrix index over the named axes
// `axes` (most significant first) of a row-major tensor; false when a
// dimension is not a power of two or there are too many digits.
static bool build_index_map(const std::vector<std::string>& indices,
                            const std::vector<std::size_t>& dims,
                            const std::vector<std::string>& axes, qfx_index_map* map) {
  const std::size_t r = indices.size();
  std::vector<int> lg(r), shift(r);
  for (std::size_t i = 0; i < r; ++i) {
    if (dims[i] == 0 || (dims[i] & (dims[i] - 1)) != 0) return false;
    int b = 0;
    while ((std::size_t(1) << b) < dims[i]) ++b;
    lg[i] = b;
  }
  int acc = 0;
  for (std::size_t i = r; i-- > 0;) {
    shift[i] = acc;
    acc += lg[i];
  }
  *map = qfx_index_map{};
  for (auto it = axes.rbegin(); it != axes.rend(); ++it) {
    const std::size_t p = std::size_t(std::find(indices.begin(), indices.end(), *it) - indices.begin());
    if (p == r) return false;
    if (lg[p] == 0) continue;
    if (map->ndigits == 16) return false;
    map->bits[map->ndigits] = static_cast<unsigned char>(lg[p]);
    map->shift[map->ndigits] = static_cast<unsigned char>(shift[p]);
    map->ndigits++;
  }
  return true;
}

bool ContractionData::try_mapped(const s_type* a, const std::vector<std::string>& a_indices,
                                 const std::vector<std::size_t>& a_dims, const s_type* b,
                                 const std::vector<std::string>& b_indices,
                                 const std::vector<std::size_t>& b_dims,
                                 const std::vector<std::string>& left,
                                 const std::vector<std::string>& common,
                                 const std::vector<std::string>& right, s_type* out,
                                 qfx_scale_io* io) {
  qfx_index_map ar, ac, br, bc;
  if (!build_index_map(a_indices, a_dims, left, &ar) ||
      !build_index_map(a_indices, a_dims, common, &ac) ||
      !build_index_map(b_indices, b_dims, common, &br) ||
      !build_index_map(b_indices, b_dims, right, &bc))
    return false;
  const int st = qfx_cgemm_mapped_ex_c64(device(), global::precision,
                                         reinterpret_cast<const qfx_c64*>(a), &ar, &ac,
                                         reinterpret_cast<const qfx_c64*>(b), &br, &bc,
                                         reinterpret_cast<qfx_c64*>(out), io);
  if (st == QFX_ERR_UNSUPPORTED) return false;
  qfx_check(st, "qfx_cgemm_mapped_c64");
  if (global::verbose > 1) {
    std::size_t m = std::size_t(1) << 0, n = 1, k = 1;
    for (int j = 0; j < ar.ndigits; ++j) m <<= ar.bits[j];
    for (int j = 0; j < bc.ndigits; ++j) n <<= bc.bits[j];
    for (int j = 0; j < ac.ndigits; ++j) k <<= ac.bits[j];
    std::cerr << "multiply m=" << m << " n=" << n << " k=" << k << " mapped" << std::endl;
  }
  stats_.mapped_multiplies++;
  return true;
}

void ContractionData::grid_operand_maxima(const View& v, const float** out, int* count) {
  bool in_grid = false;
  for (const auto& r : grid_ranges_)
    if (v.ptr >= r.first && v.ptr + v.size() <= r.second) in_grid = true;
  if (!in_grid) return;  // patches change from slice to slice: no caching
  const auto key = std::make_pair(static_cast<const s_type*>(v.ptr), v.size());
  auto it = grid_amax_.find(key);
  if (it == grid_amax_.end()) {
    const int cap = 256;
    grid_amax_store_.emplace_back(std::vector<std::string>{""},
                                  std::vector<std::size_t>{std::size_t(cap / 2)});
    float* arr = reinterpret_cast<float*>(grid_amax_store_.back().data());
    int n = 0;
    qfx_check(qfx_amax_partials_c64(device(), reinterpret_cast<const qfx_c64*>(v.ptr), v.size(),
                                    arr, cap, &n),
              "qfx_amax_partials_c64");
    it = grid_amax_.emplace(key, std::make_pair(static_cast<const float*>(arr), n)).first;
  }
  *out = it->second.first;
  *count = it->second.second;
}

bool ContractionData::is_batch(const std::string& index) const {
  return std::find(batch_.begin(), batch_.end(), index) != batch_.end();
}

// C[batch][left][right] = sum_common A[batch, left, common] B[batch, common,
// right] for every value of the batch indices both operands carry: one
// product per batch value, each read in the operands' stored orders where a
// kernel supports it (else both operands are reordered to (batch, left,
// common) / (batch, common, right) once and the products read contiguous
// blocks).  The split-FP16 scale of every product comes from the maxima of
// the whole operand.
ContractionData::View ContractionData::contract_batched(const View& A, const View& B, s_type* out,
                                                        const std::vector<std::string>& bc) {
  std::unordered_map<std::string, std::size_t> adim, bdim;
  for (std::size_t i = 0; i < A.indices.size(); ++i) adim[A.indices[i]] = A.dims[i];
  for (std::size_t i = 0; i < B.indices.size(); ++i) bdim[B.indices[i]] = B.dims[i];
  const std::vector<std::string> left = _vector_subtraction(A.indices, B.indices);
  const std::vector<std::string> right = _vector_subtraction(B.indices, A.indices);
  const std::vector<std::string> common =
      _vector_subtraction(_vector_intersection(A.indices, B.indices), bc);
  std::size_t m = 1, n = 1, k = 1, nb = 1;
  for (const auto& x : left) m *= adim[x];
  for (const auto& x : right) n *= bdim[x];
  for (const auto& x : common) {
    if (adim[x] != bdim[x])
      throw ERROR_MSG("Common indices must have matching dimensions, but at index: ", x,
                      ", the dimensions are A: ", adim[x], ", B: ", bdim[x], ".");
    k *= adim[x];
  }
  for (const auto& x : bc) {
    if (adim[x] != bdim[x])
      throw ERROR_MSG("Batch index ", x, " has dimension ", adim[x], " in A and ", bdim[x],
                      " in B.");
    nb *= adim[x];
  }
  // element strides of the batch indices in A and B
  auto strides = [](const View& v, const std::vector<std::string>& names) {
    std::vector<std::size_t> out;
    for (const auto& x : names) {
      std::size_t st = 1;
      const std::size_t p = std::size_t(std::find(v.indices.begin(), v.indices.end(), x) -
                                        v.indices.begin());
      for (std::size_t a = p + 1; a < v.dims.size(); ++a) st *= v.dims[a];
      out.push_back(st);
    }
    return out;
  };
  const auto sa = strides(A, bc), sb = strides(B, bc);
  std::vector<std::size_t> bdims;
  for (const auto& x : bc) bdims.push_back(adim[x]);
  // offsets of batch value b (row-major over bc, first slowest)
  auto offsets = [&](std::size_t b, std::size_t* oa, std::size_t* ob) {
    *oa = *ob = 0;
    for (std::size_t j = bc.size(); j-- > 0;) {
      const std::size_t d = b % bdims[j];
      b /= bdims[j];
      *oa += d * sa[j];
      *ob += d * sb[j];
    }
  };
  qfx_scale_io io{};
  io.a_max = A.amax;
  io.a_max_n = A.amax_n;
  io.b_max = B.amax;
  io.b_max_n = B.amax_n;
  if (global::precision == QFX_PREC_SPLIT_FP16) {
    if (batch_amax_.empty())
      for (int i = 0; i < 2; ++i)
        batch_amax_.emplace_back(std::vector<std::string>{""}, std::vector<std::size_t>{128});
    auto whole = [&](const View& v, int slot, const float** arr, int* cnt) {
      if (*cnt > 0) return;
      float* dst = reinterpret_cast<float*>(batch_amax_[slot].data());
      qfx_check(qfx_amax_partials_c64(device(), reinterpret_cast<const qfx_c64*>(v.ptr), v.size(),
                                      dst, 256, cnt),
                "qfx_amax_partials_c64");
      *arr = dst;
    };
    whole(A, 0, &io.a_max, &io.a_max_n);
    whole(B, 1, &io.b_max, &io.b_max_n);
  }
  std::size_t done = 0;
  if (!left.empty() && !right.empty() && global::mapped_operands) {
    for (; done < nb; ++done) {
      std::size_t oa, ob;
      offsets(done, &oa, &ob);
      qfx_scale_io one = io;  // no output maxima: each product writes a part of C
      if (!try_mapped(A.ptr + oa, A.indices, A.dims, B.ptr + ob, B.indices, B.dims, left, common,
                      right, out + done * m * n, &one))
        break;
    }
  }
  if (done < nb) {
    const s_type* pa = A.ptr;
    const s_type* pb = B.ptr;
    const auto a_order = _vector_union(bc, _vector_union(left, common));
    const auto b_order = _vector_union(bc, _vector_union(common, right));
    if (a_order != A.indices) {
      s_type* buf = get_scratch(kGeneralSpace).data();
      device_reorder(A.ptr, buf, A.indices, A.dims, a_order);
      pa = buf;
      stats_.reorders++;
      stats_.reorder_elements += double(A.size());
    }
    if (b_order != B.indices) {
      s_type* buf = get_scratch(kGeneralSpaceB).data();
      device_reorder(B.ptr, buf, B.indices, B.dims, b_order);
      pb = buf;
      stats_.reorders++;
      stats_.reorder_elements += double(B.size());
    }
    for (; done < nb; ++done)
      device_multiply(pa + done * m * k, pb + done * k * n, out + done * m * n, m, n, k,
                      !left.empty(), !right.empty());
  }
  stats_.multiplies++;
  stats_.flops += 8.0 * double(m) * double(n) * double(k) * double(nb);
  View C;
  C.ptr = out;
  C.indices = _vector_union(bc, _vector_union(left, right));
  for (const auto& x : bc) C.dims.push_back(adim[x]);
  for (const auto& x : left) C.dims.push_back(adim[x]);
  for (const auto& x : right) C.dims.push_back(bdim[x]);
  return C;
}

ContractionData::View ContractionData::contract_pair(const View& A, const View& B, s_type* out) {
  if (!batch_.empty()) {
    std::vector<std::string> bc;
    for (const auto& x : _vector_intersection(A.indices, B.indices))
      if (is_batch(x)) bc.push_back(x);
    if (!bc.empty()) return contract_batched(A, B, out, bc);
  }
  std::vector<std::string> left = _vector_subtraction(A.indices, B.indices);
  std::vector<std::string> right = _vector_subtraction(B.indices, A.indices);
  std::vector<std::string> common = _vector_intersection(A.indices, B.indices);
  std::unordered_map<std::string, std::size_t> adim, bdim;
  for (std::size_t i = 0; i < A.indices.size(); ++i) adim[A.indices[i]] = A.dims[i];
  for (std::size_t i = 0; i < B.indices.size(); ++i) bdim[B.indices[i]] = B.dims[i];
  std::size_t left_dim = 1, right_dim = 1, common_dim = 1;
  for (const auto& x : left) left_dim *= adim[x];
  for (const auto& x : right) right_dim *= bdim[x];
  for (std::size_t i = 0; i < common.size(); ++i) {
    if (adim[common[i]] != bdim[common[i]])
      throw ERROR_MSG("Common indices must have matching dimensions, but at index: ", i,
                      ", the dimensions are A: ", adim[common[i]], ", B: ", bdim[common[i]], ".");
    common_dim *= adim[common[i]];
  }
  const s_type* pa = A.ptr;
  const s_type* pb = B.ptr;
  auto a_order = _vector_union(left, common);
  auto b_order = _vector_union(common, right);
  std::vector<std::string> b_indices = B.indices;
  std::vector<std::size_t> b_dims = B.dims;
  bool done = false;
  // split-FP16 scales: the operands' partial maxima if their producing GEMM
  // wrote them, and the output's if `out` has an array for them
  qfx_scale_io io{};
  io.a_max = A.amax;
  io.a_max_n = A.amax_n;
  io.b_max = B.amax;
  io.b_max_n = B.amax_n;
  if (global::precision == QFX_PREC_SPLIT_FP16) {
    if (io.a_max_n == 0) grid_operand_maxima(A, &io.a_max, &io.a_max_n);
    if (io.b_max_n == 0) grid_operand_maxima(B, &io.b_max, &io.b_max_n);
  }
  auto slot = amax_slots_.find(out);
  if (slot != amax_slots_.end()) {
    io.c_max = slot->second.first;
    io.c_max_cap = slot->second.second;
  }
  if (!left.empty() && !right.empty() && global::mapped_operands) {
    // Read A (and B) in stored order; if B's layout is the obstacle and B is
    // the smaller operand, reorder only B and try again.
    done = try_mapped(A.ptr, A.indices, A.dims, pb, b_indices, b_dims, left, common, right, out,
                      &io);
    if (!done && b_order != B.indices && B.size() <= A.size()) {
      s_type* buf = get_scratch(kGeneralSpaceB).data();
      device_reorder(B.ptr, buf, B.indices, B.dims, b_order);
      pb = buf;
      b_indices = b_order;
      b_dims.clear();
      for (const auto& x : b_order) b_dims.push_back(bdim[x]);
      stats_.reorders++;
      stats_.reorder_elements += double(B.size());
      done = try_mapped(A.ptr, A.indices, A.dims, pb, b_indices, b_dims, left, common, right, out,
                        &io);
    }
  }
  if (!done) io.c_max_n = 0;
  if (!done) {
    if (a_order != A.indices) {
      s_type* buf = get_scratch(kGeneralSpace).data();
      device_reorder(A.ptr, buf, A.indices, A.dims, a_order);
      pa = buf;
      stats_.reorders++;
      stats_.reorder_elements += double(A.size());
    }
    if (b_order != b_indices) {
      s_type* buf = get_scratch(kGeneralSpaceB).data();
      device_reorder(B.ptr, buf, B.indices, B.dims, b_order);
      pb = buf;
      stats_.reorders++;
      stats_.reorder_elements += double(B.size());
    }
    if (global::verbose > 1)
      std::cerr << "multiply m=" << left_dim << " n=" << right_dim << " k=" << common_dim
                << " reorder_a=" << (pa != A.ptr ? A.size() : 0)
                << " reorder_b=" << (pb != B.ptr ? B.size() : 0) << std::endl;
    device_multiply(pa, pb, out, left_dim, right_dim, common_dim, !left.empty(), !right.empty());
  }
  stats_.multiplies++;
  stats_.flops += 8.0 * double(left_dim) * double(right_dim) * double(common_dim);
  View C;
  C.ptr = out;
  C.amax = io.c_max_n > 0 ? io.c_max : nullptr;
  C.amax_n = io.c_max_n;
  C.indices = _vector_union(left, right);
  for (const auto& x : left) C.dims.push_back(adim[x]);
  for (const auto& x : right) C.dims.push_back(bdim[x]);
  return C;
}

ContractionData::View ContractionData::project_view(const View& v, const std::string& index,
                                                    std::size_t value,
                                                    const std::vector<std::size_t>* gather_values,
                                                    s_type* buffer, bool copy) {
  auto it = std::find(v.indices.begin(), v.indices.end(), index);
  if (it == v.indices.end())
    throw ERROR_MSG("Index: '", index, "' has to be equal to indices[0]: '",
                    v.indices.empty() ? std::string() : v.indices[0], "'.");
  const std::size_t p = std::size_t(it - v.indices.begin());
  const std::size_t dim = v.dims[p];
  std::size_t pre = 1, post = 1;
  for (std::size_t a = 0; a < p; ++a) pre *= v.dims[a];
  for (std::size_t a = p + 1; a < v.dims.size(); ++a) post *= v.dims[a];
  std::vector<std::size_t> picks;
  if (gather_values)
    picks = *gather_values;
  else
    picks = {value};
  for (auto x : picks)
    if (x >= dim)
      throw ERROR_MSG("index_value: ", x, " must be contained in [0, ", dim, ").");
  View out;
  out.indices = v.indices;
  out.dims = v.dims;
  if (!gather_values) {
    out.indices.erase(out.indices.begin() + p);
    out.dims.erase(out.dims.begin() + p);
  } else {
    out.dims[p] = picks.size();
  }
  bool consecutive = true;
  for (std::size_t j = 1; j < picks.size(); ++j) consecutive = consecutive && picks[j] == picks[0] + j;
  if (pre == 1 && consecutive && !picks.empty() && !copy) {
    out.ptr = v.ptr + picks[0] * post;  // contiguous slice: zero copy
    return out;
  }
  if (buffer == nullptr) throw ERROR_MSG("No buffer for projection of index ", index, ".");
  // out[a][k][rest] = v[a][picks[k]][rest], one launch
  std::vector<uint64_t> pk(picks.begin(), picks.end());
  qfx_check(qfx_project_c64(device(), reinterpret_cast<const qfx_c64*>(v.ptr),
                            reinterpret_cast<qfx_c64*>(buffer), pre, dim, post, pk.size(),
                            pk.data()),
            "qfx_project_c64");
  out.ptr = buffer;
  return out;
}

void ContractionData::emit_output(const View& output, std::size_t output_index,
                                  const std::vector<OpenOutput>& open_outputs) {
  std::vector<std::string> open_names;
  std::size_t P = 1;
  for (const auto& o : open_outputs) {
    open_names.push_back(o.index);
    P *= o.count;
  }
  // batched cut indices still open: one amplitude vector per batch value
  std::vector<std::string> batch_names;
  std::size_t nb = 1;
  for (std::size_t i = 0; i < output.indices.size(); ++i)
    if (is_batch(output.indices[i])) {
      batch_names.push_back(output.indices[i]);
      nb *= output.dims[i];
    }
  if (!_vector_subtraction(_vector_subtraction(output.indices, open_names), batch_names).empty() ||
      output.size() != P * nb) {
    std::string desc = "Tensor of rank " + std::to_string(output.indices.size()) + ": ";
    for (std::size_t i = 0; i < output.indices.size(); ++i)
      desc += output.indices[i] + " -> " + std::to_string(output.dims[i]) +
              (i + 1 < output.indices.size() ? ", " : "");
    throw ERROR_MSG("Contraction did not complete; final tensor is: ", desc);
  }
  const std::size_t offset = output_index * P;
  if (offset + P > amplitudes_->size())
    throw ERROR_MSG("Output index ", offset + P - 1, " is outside the amplitude vector of size ",
                    amplitudes_->size(), ".");
  const s_type* src = output.ptr;
  const std::vector<std::string> order = _vector_union(batch_names, open_names);
  if (order != output.indices) {
    s_type* buf = get_scratch(kGeneralSpaceB).data();
    if (buf == output.ptr) buf = get_scratch(kGeneralSpace).data();
    device_reorder(output.ptr, buf, output.indices, output.dims, order);
    src = buf;
  }
  qfx_c128* amps = reinterpret_cast<qfx_c128*>(get_scratch(kAmplitudeSpace).data());
  if (nb == 1)
    qfx_check(qfx_accumulate_c128(device(), reinterpret_cast<const qfx_c64*>(src), P, amps,
                                  offset),
              "qfx_accumulate_c128");
  else
    qfx_check(qfx_accumulate_batched_c128(device(), reinterpret_cast<const qfx_c64*>(src), P, nb,
                                          amps, offset),
              "qfx_accumulate_batched_c128");
  stats_.slices += nb;
}

// A bond cut's passes.  A batched pass keeps the cut index, restricted to
// the pass's values, as a batch index.  With global::record_slices (and no
// further bond cut in the rest of the ordering, no profiling, at least three
// passes) the projections go into the cut's fixed buffers, the first pass
// runs eagerly (it grows the workspace and tunes the tiles), the second is
// recorded as one graph, and every later pass is the projection copies
// plus one launch of that graph; the host's buffer assignment is reset
// before every pass so that all passes use the same buffers.
void ContractionData::cut_loop(const std::string& index, const std::vector<std::size_t>& pa,
                               const std::vector<std::size_t>& pb, const View& saved_a,
                               const View& saved_b, s_type* buf_a, s_type* buf_b,
                               const std::vector<std::vector<std::size_t>>& groups, bool batched,
                               const std::list<ContractionOperation>& ordering,
                               std::size_t output_index,
                               const std::unordered_map<std::string, bool>& active_patches,
                               const std::vector<OpenOutput>& open_outputs) {
  bool innermost = true;
  for (const auto& op : ordering)
    if (op.op_type == ContractionOperation::CUT && op.cut.tensors.size() > 1) innermost = false;
  const bool record = global::record_slices && !global::profiling && innermost && !recording_ &&
                      groups.size() >= 3;
  const auto map_before = scratch_map_;
  qfx_graph* graph = nullptr;
  ContractionStats per_pass{};
  if (batched) batch_.push_back(index);
  auto cleanup = [&]() {
    if (batched) batch_.pop_back();
    if (graph) qfx_graph_destroy(graph);
    graph = nullptr;
    recording_ = false;
  };
  try {
    for (std::size_t g = 0; g < groups.size(); ++g) {
      const auto& picks = groups[g];
      try {
        grid_views_[pa[0]][pa[1]] =
            project_view(saved_a, index, picks[0], batched ? &picks : nullptr, buf_a, record);
        grid_views_[pb[0]][pb[1]] =
            project_view(saved_b, index, picks[0], batched ? &picks : nullptr, buf_b, record);
      } catch (const std::string& err_msg) {
        throw ERROR_MSG("Failed to call project(). Error:\n\t[", err_msg, "]");
      }
      if (record) scratch_map_ = map_before;
      // the last pass may hold fewer values than the recorded one: eager
      const bool same_shape = picks.size() == groups[0].size();
      if (record && g >= 2 && graph && same_shape) {
        qfx_check(qfx_graph_launch(device(), graph), "qfx_graph_launch");
        stats_.multiplies += per_pass.multiplies;
        stats_.flops += per_pass.flops;
        stats_.reorders += per_pass.reorders;
        stats_.reorder_elements += per_pass.reorder_elements;
        stats_.mapped_multiplies += per_pass.mapped_multiplies;
        stats_.slices += per_pass.slices;
        stats_.replayed_passes++;
        continue;
      }
      std::list<ContractionOperation> rest(ordering);
      auto active = active_patches;
      if (record && g == 1 && same_shape) {
        const ContractionStats before = stats_;
        qfx_check(qfx_graph_begin(device(), 0), "qfx_graph_begin");
        recording_ = true;
        try {
          run(rest, output_index, active, open_outputs);
        } catch (...) {
          qfx_graph* dropped = nullptr;
          if (qfx_graph_end(device(), &dropped) == QFX_OK && dropped) qfx_graph_destroy(dropped);
          throw;
        }
        recording_ = false;
        qfx_check(qfx_graph_end(device(), &graph), "qfx_graph_end");
        qfx_check(qfx_graph_launch(device(), graph), "qfx_graph_launch");
        per_pass.multiplies = stats_.multiplies - before.multiplies;
        per_pass.flops = stats_.flops - before.flops;
        per_pass.reorders = stats_.reorders - before.reorders;
        per_pass.reorder_elements = stats_.reorder_elements - before.reorder_elements;
        per_pass.mapped_multiplies = stats_.mapped_multiplies - before.mapped_multiplies;
        per_pass.slices = stats_.slices - before.slices;
        continue;
      }
      run(rest, output_index, active, open_outputs);
    }
  } catch (...) {
    cleanup();
    throw;
  }
  cleanup();
}

void ContractionData::run(std::list<ContractionOperation>& ordering, std::size_t output_index,
                          std::unordered_map<std::string, bool>& active_patches,
                          std::vector<OpenOutput> open_outputs) {
  std::vector<std::pair<std::vector<std::size_t>, View>> restore;
  bool have_output = false;
  View output;
  while (!ordering.empty()) {
    const ContractionOperation op = ordering.front();
    ordering.pop_front();
    switch (op.op_type) {
      case ContractionOperation::EXPAND: {
        const View& next = grid_views_.at(op.expand.tensor[0]).at(op.expand.tensor[1]);
        if (!active_patches[op.expand.id]) {
          patches_[op.expand.id] = next;
          active_patches[op.expand.id] = true;
          continue;
        }
        const std::string rid = result_space(patch_max_size_[op.expand.id]);
        View res;
        try {
          res = contract_pair(patches_[op.expand.id], next, get_scratch(rid).data());
        } catch (const std::string& err_msg) {
          throw ERROR_MSG("Failed to call multiply(). Error:\n\t[", err_msg, "]");
        }
        if (ordering.empty()) {
          output = res;
          have_output = true;
          continue;
        }
        std::swap(scratch_map_[op.expand.id], scratch_map_[rid]);
        patches_[op.expand.id] = res;
        continue;
      }
      case ContractionOperation::CUT: {
        const std::string index = index_name(op.cut.tensors);
        const auto& pa = op.cut.tensors[0];
        View& va = grid_views_.at(pa[0]).at(pa[1]);
        const View saved_a = va;
        std::size_t dim = 0;
        for (std::size_t a = 0; a < saved_a.indices.size(); ++a)
          if (saved_a.indices[a] == index) dim = saved_a.dims[a];
        std::vector<std::size_t> values = op.cut.values;
        if (values.empty())
          for (std::size_t v = 0; v < dim; ++v) values.push_back(v);
        if (op.cut.tensors.size() > 1) {
          const auto& pb = op.cut.tensors[1];
          View& vb = grid_views_.at(pb[0]).at(pb[1]);
          const View saved_b = vb;
          s_type* buf_a = get_scratch(cut_copy_name(op.cut.tensors, 0)).data();
          s_type* buf_b = get_scratch(cut_copy_name(op.cut.tensors, 1)).data();
          auto batch_it = cut_batches_.find(index);
          const std::size_t group = batch_it == cut_batches_.end() ? 1 : batch_it->second;
          std::vector<std::vector<std::size_t>> groups;
          for (std::size_t first = 0; first < values.size(); first += std::max<std::size_t>(group, 1))
            groups.emplace_back(values.begin() + first,
                                values.begin() + std::min(first + std::max<std::size_t>(group, 1),
                                                          values.size()));
          cut_loop(index, pa, pb, saved_a, saved_b, buf_a, buf_b, groups, group > 1, ordering,
                   output_index, active_patches, open_outputs);
          grid_views_[pa[0]][pa[1]] = saved_a;
          grid_views_[pb[0]][pb[1]] = saved_b;
        } else {
          // Terminal cut: keep the output index open (one entry per value).
          bool full = values.size() == dim;
          for (std::size_t v = 0; full && v < values.size(); ++v) full = values[v] == v;
          if (!full) {
            try {
              grid_views_[pa[0]][pa[1]] = project_view(
                  saved_a, index, 0, &values, get_scratch(cut_copy_name(op.cut.tensors, 0)).data());
            } catch (const std::string& err_msg) {
              throw ERROR_MSG("Failed to call project(). Error:\n\t[", err_msg, "]");
            }
            restore.push_back({pa, saved_a});
          }
          open_outputs.push_back({index, values.size()});
          continue;
        }
        for (auto it = restore.rbegin(); it != restore.rend(); ++it)
          grid_views_[it->first[0]][it->first[1]] = it->second;
        return;  // the rest of the ordering ran inside the cut loop
      }
      case ContractionOperation::MERGE: {
        if (!active_patches[op.merge.target_id]) {
          patches_[op.merge.target_id] = patches_[op.merge.source_id];
          active_patches[op.merge.target_id] = true;
          continue;
        }
        const std::string rid = result_space(patch_max_size_[op.merge.target_id]);
        View res;
        try {
          res = contract_pair(patches_[op.merge.source_id], patches_[op.merge.target_id],
                              get_scratch(rid).data());
        } catch (const std::string& err_msg) {
          throw ERROR_MSG("Failed to call multiply(). Error:\n\t[", err_msg, "]");
        }
        if (ordering.empty()) {
          output = res;
          have_output = true;
          continue;
        }
        std::swap(scratch_map_[op.merge.target_id], scratch_map_[rid]);
        patches_[op.merge.target_id] = res;
        continue;
      }
    }
  }
  if (!have_output) throw ERROR_MSG("Contraction did not complete; no final tensor was produced.");
  emit_output(output, output_index, open_outputs);
  for (auto it = restore.rbegin(); it != restore.rend(); ++it)
    grid_views_[it->first[0]][it->first[1]] = it->second;
}

// Start of a top-level contraction: grid views of the grid tensors, fresh
// state and counters, the device drained (the network build and flatten
// before it are not counted), the FP64 amplitudes zeroed.
void ContractionData::begin_contraction() {
  grid_views_.assign(tensor_grid_->size(), {});
  for (std::size_t i = 0; i < tensor_grid_->size(); ++i) {
    grid_views_[i].resize((*tensor_grid_)[i].size());
    for (std::size_t j = 0; j < (*tensor_grid_)[i].size(); ++j) {
      const Tensor& t = (*tensor_grid_)[i][j];
      grid_views_[i][j].ptr = t.data();
      grid_views_[i][j].indices = t.get_indices();
      grid_views_[i][j].dims = t.get_dimensions();
    }
  }
  patches_.clear();
  batch_.clear();
  grid_amax_.clear();
  grid_amax_store_.clear();
  grid_ranges_.clear();
  for (const auto& row : *tensor_grid_)
    for (const auto& t : row) grid_ranges_.push_back({t.data(), t.data() + t.capacity()});
  ContractionStats fresh;
  fresh.planned_bytes = stats_.planned_bytes;
  stats_ = fresh;
  qfx_check(qfx_synchronize(device()), "qfx_synchronize");
  start_ = std::chrono::steady_clock::now();
  Tensor& amps = get_scratch(kAmplitudeSpace);
  qfx_check(qfx_zero(device(), amps.data(), amps.capacity() * sizeof(s_type)), "qfx_zero");
}

// End of a top-level contraction: the FP64 amplitudes downloaded once and
// added to the caller's vector; wall time and counters recorded.
void ContractionData::end_contraction() {
  std::vector<std::complex<double>> device_amps(amplitudes_->size());
  if (!device_amps.empty())
    qfx_check(qfx_copy_d2h(device(), device_amps.data(), get_scratch(kAmplitudeSpace).data(),
                           device_amps.size() * sizeof(std::complex<double>)),
              "qfx_copy_d2h");
  for (std::size_t i = 0; i < device_amps.size(); ++i) (*amplitudes_)[i] += device_amps[i];
  qfx_check(qfx_synchronize(device()), "qfx_synchronize");
  stats_.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - start_).count();
  g_last_stats = stats_;
}

void ContractionData::ContractGrid(std::list<ContractionOperation> ordering,
                                   std::size_t output_index,
                                   std::unordered_map<std::string, bool> active_patches) {
  const bool top = depth_ == 0;
  if (top) {
    cut_batches_ = bond_cut_batches(ordering, *tensor_grid_);
    begin_contraction();
  }
  ++depth_;
  try {
    run(ordering, output_index, active_patches, {});
  } catch (...) {
    --depth_;
    throw;
  }
  --depth_;
  if (top) end_contraction();
}

void ContractionData::ContractGridStates(const std::list<ContractionOperation>& ordering,
                                         const std::unordered_map<std::string, bool>& active_patches,
                                         const std::vector<std::vector<SiteProjection>>& states) {
  if (depth_ != 0) throw ERROR_MSG("ContractGridStates cannot be nested.");
  cut_batches_ = bond_cut_batches(ordering, *tensor_grid_);
  begin_contraction();
  const auto base = grid_views_;
  for (std::size_t s = 0; s < states.size(); ++s) {
    grid_views_ = base;
    for (const auto& p : states[s]) {
      if (p.i >= grid_views_.size() || p.j >= grid_views_[p.i].size())
        throw ERROR_MSG("Projection of a tensor outside the grid: (", p.i, ",", p.j, ").");
      try {
        grid_views_[p.i][p.j] = project_view(base[p.i][p.j], p.index, p.value, nullptr, nullptr);
      } catch (const std::string& err_msg) {
        throw ERROR_MSG("Failed to call project(). Error:\n\t[", err_msg, "]");
      }
    }
    // the operand maxima cached for grid views belong to this state's views
    grid_amax_.clear();
    grid_amax_store_.clear();
    patches_.clear();
    std::list<ContractionOperation> rest(ordering);
    auto active = active_patches;
    ++depth_;
    try {
      run(rest, s, active, {});
    } catch (...) {
      --depth_;
      throw;
    }
    --depth_;
  }
  end_contraction();
}

void ContractGrid(const std::list<ContractionOperation>& ordering,
                  std::vector<std::vector<Tensor>>* tensor_grid,
                  std::vector<std::complex<double>>* amplitudes) {
  if (tensor_grid == nullptr) throw ERROR_MSG("Tensor grid must be non-null.");
  if (amplitudes == nullptr) throw ERROR_MSG("Amplitude return vector must be non-null.");
  ContractionData data;
  try {
    data = ContractionData::Initialize(ordering, tensor_grid, amplitudes);
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call Initialize(). Error:\n\t[", err_msg, "]");
  }
  std::unordered_map<std::string, bool> active_patches;
  for (const auto& patch : data.scratch_list()) active_patches[patch] = false;
  try {
    data.ContractGrid(ordering, /*output_index=*/0, active_patches);
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call ContractGrid(). Error:\n\t[", err_msg, "]");
  }
}

}  // namespace qflex
