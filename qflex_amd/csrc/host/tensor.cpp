// Tensor over the device arena; every data operation goes through the qfx
// C-ABI.  Semantics and error texts follow the reference's src/tensor.cpp.
#include "tensor.h"

#include <algorithm>
#include <iostream>
#include <numeric>

#include "device.h"
#include "errors.h"

namespace qflex {

namespace {

qfx_c64* c64(s_type* p) { return reinterpret_cast<qfx_c64*>(p); }
const qfx_c64* c64(const s_type* p) { return reinterpret_cast<const qfx_c64*>(p); }

}  // namespace

s_type* device_alloc(std::size_t elements) {
  void* p = nullptr;
  qfx_check(qfx_alloc(device(), std::max<std::size_t>(elements, 1) * sizeof(s_type), &p),
            "qfx_alloc");
  return static_cast<s_type*>(p);
}

void device_free(s_type* p) {
  if (p) qfx_free(device(), p);
}

///////////////////////////// CLASS FUNCTIONS /////////////////////////////

void Tensor::_init(const std::vector<std::string>& indices,
                   const std::vector<std::size_t>& dimensions) {
  if (indices.size() != dimensions.size()) {
    throw ERROR_MSG("The number of indices: ", indices.size(),
                    ", and number of dimensions: ", dimensions.size(), ", should be equal.");
  }
  _indices = indices;
  _dimensions = dimensions;
  _index_to_dimension.clear();
  for (std::size_t i = 0; i < _dimensions.size(); ++i)
    _index_to_dimension[indices[i]] = dimensions[i];
}

void Tensor::_clear() {
  if (_data != nullptr) {
    device_free(_data);
    _data = nullptr;
  }
}

void Tensor::_copy(const Tensor& other) {
  if (_capacity == 0) {
    if (_data != nullptr) throw ERROR_MSG("Potential memory leak");
    _capacity = other._capacity;
    _data = device_alloc(_capacity);
  } else {
    try {
      set_dimensions(other.get_dimensions());
    } catch (const std::string& err_msg) {
      throw ERROR_MSG("Failed to call set_dimensions(). Error:\n\t[", err_msg, "]");
    }
  }
  try {
    _init(other.get_indices(), other.get_dimensions());
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call _init(). Error:\n\t[", err_msg, "]");
  }
  if (other._data != nullptr)
    qfx_check(qfx_copy_d2d(device(), _data, other._data, other.size() * sizeof(s_type)),
              "qfx_copy_d2d");
}

void Tensor::_move(Tensor&& other) noexcept {
  _clear();
  _indices = std::move(other._indices);
  _dimensions = std::move(other._dimensions);
  _index_to_dimension = std::move(other._index_to_dimension);
  _data = other._data;
  _capacity = other._capacity;
  other._data = nullptr;
  other._capacity = 0;
}

Tensor::Tensor() {}

Tensor::Tensor(const std::vector<std::string>& indices,
               const std::vector<std::size_t>& dimensions) {
  try {
    _init(indices, dimensions);
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call _init(). Error:\n\t[", err_msg, "]");
  }
  _capacity = size();
  _data = device_alloc(_capacity);
  qfx_check(qfx_zero(device(), _data, _capacity * sizeof(s_type)), "qfx_zero");
}

Tensor::Tensor(const std::vector<std::string>& indices,
               const std::vector<std::size_t>& dimensions, for_overwrite_t) {
  try {
    _init(indices, dimensions);
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call _init(). Error:\n\t[", err_msg, "]");
  }
  _capacity = size();
  _data = device_alloc(_capacity);
}

Tensor::Tensor(const std::vector<std::string>& indices,
               const std::vector<std::size_t>& dimensions, const std::vector<s_type>& data) {
  try {
    _init(indices, dimensions);
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call _init(). Error:\n\t[", err_msg, "]");
  }
  std::size_t this_size = size();
  if (this_size != data.size()) {
    throw ERROR_MSG("The vector data size: ", data.size(),
                    ", has to match the size of the Tensor: ", this_size);
  }
  _capacity = this_size;
  _data = device_alloc(_capacity);
  qfx_check(qfx_copy_h2d_async(device(), _data, data.data(), this_size * sizeof(s_type)),
            "qfx_copy_h2d_async");
}

Tensor::Tensor(const Tensor& other) { _copy(other); }

Tensor::Tensor(Tensor&& other) noexcept { _move(std::move(other)); }

Tensor::~Tensor() { _clear(); }

const Tensor& Tensor::operator=(const Tensor& other) {
  if (other._data != nullptr && this != &other)
    _copy(other);
  else
    throw ERROR_MSG("Cannot copy tensor.");
  return *this;
}

const Tensor& Tensor::operator=(Tensor&& other) {
  if (other._data != nullptr && this != &other)
    _move(std::move(other));
  else
    throw ERROR_MSG("Cannot move tensor.");
  return *this;
}

const std::vector<std::string>& Tensor::get_indices() const { return _indices; }

void Tensor::set_indices(const std::vector<std::string>& indices) { _indices = indices; }

const std::vector<std::size_t>& Tensor::get_dimensions() const { return _dimensions; }

void Tensor::set_dimensions(const std::vector<std::size_t>& dimensions) {
  if (_data) {
    std::size_t total_dim = 1;
    for (auto d : dimensions) total_dim *= d;
    if (capacity() < total_dim) {
      throw ERROR_MSG("The total allocated space: ", capacity(),
                      ", is insufficient for the requested tensor dimensions: ", total_dim,
                      ".");
    }
  }
  _dimensions = dimensions;
}

void Tensor::set_indices_and_dimensions(const std::vector<std::string>& indices,
                                        const std::vector<std::size_t>& dimensions) {
  try {
    set_dimensions(dimensions);
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call set_dimensions(). Error:\n\t[", err_msg, "]");
  }
  try {
    _init(indices, dimensions);
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call _init(). Error:\n\t[", err_msg, "]");
  }
}

const std::unordered_map<std::string, std::size_t>& Tensor::get_index_to_dimension() const {
  return _index_to_dimension;
}

void Tensor::generate_index_to_dimension() {
  _index_to_dimension.clear();
  for (std::size_t i = 0; i < _indices.size(); ++i)
    _index_to_dimension[_indices[i]] = _dimensions[i];
}

std::size_t Tensor::size() const {
  std::size_t total_dim = 1;
  for (auto d : _dimensions) total_dim *= d;
  return total_dim;
}

std::size_t Tensor::capacity() const { return _capacity; }

s_type* Tensor::data() { return _data; }

const s_type* Tensor::data() const { return _data; }

std::vector<s_type> Tensor::host_data() const {
  std::vector<s_type> out(_data ? size() : 0);
  if (!out.empty())
    qfx_check(qfx_copy_d2h(device(), out.data(), _data, out.size() * sizeof(s_type)),
              "qfx_copy_d2h");
  return out;
}

void Tensor::set_host_data(const std::vector<s_type>& values) {
  if (values.size() > _capacity)
    throw ERROR_MSG("The total allocated space: ", _capacity,
                    ", is insufficient for the requested data size: ", values.size(), ".");
  if (!values.empty())
    qfx_check(qfx_copy_h2d(device(), _data, values.data(), values.size() * sizeof(s_type)),
              "qfx_copy_h2d");
}

void Tensor::project(std::string index, std::size_t index_value,
                     Tensor& projection_tensor) const {
  if (_indices.empty() || index != _indices[0]) {
    throw ERROR_MSG("Index: '", index, "' has to be equal to indices[0]: '",
                    _indices.empty() ? std::string() : _indices[0], "'.");
  }
  if (index_value >= _dimensions[0]) {
    throw ERROR_MSG("index_value: ", index_value, " must be contained in [0, ", _dimensions[0],
                    ").");
  }
  std::vector<std::string> projection_indices(_indices.begin() + 1, _indices.end());
  std::vector<std::size_t> projection_dimensions(_dimensions.begin() + 1, _dimensions.end());
  projection_tensor.set_indices(projection_indices);
  try {
    projection_tensor.set_dimensions(projection_dimensions);
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call set_dimensions(). Error:\n\t[", err_msg, "]");
  }
  projection_tensor.generate_index_to_dimension();
  std::size_t projection_size = projection_tensor.size();
  qfx_check(qfx_copy_d2d(device(), projection_tensor.data(),
                         _data + projection_size * index_value,
                         projection_size * sizeof(s_type)),
            "qfx_copy_d2d");
}

void Tensor::rename_index(std::string old_name, std::string new_name) {
  auto it = std::find(_indices.begin(), _indices.end(), old_name);
  if (it == _indices.end()) throw ERROR_MSG("old_name: ", old_name, ", has to be a valid index.");
  if (std::find(_indices.begin(), _indices.end(), new_name) != _indices.end())
    throw ERROR_MSG("new_name: ", new_name, ", cannot be an existing index.");
  *it = new_name;
  _index_to_dimension[new_name] = _index_to_dimension[old_name];
  _index_to_dimension.erase(old_name);
}

void Tensor::bundle(std::vector<std::string> indices_to_bundle, std::string bundled_index) {
  if (!_vector_s_in_vector_s(indices_to_bundle, _indices)) {
    throw ERROR_MSG("indices_to_bundle: ", _string_vector_to_string(indices_to_bundle),
                    " has to be contained in indices: ", _string_vector_to_string(_indices),
                    ".");
  }
  std::vector<std::string> subtracted(_vector_subtraction(_indices, indices_to_bundle));
  std::vector<std::string> original_order(_vector_subtraction(_indices, subtracted));
  if (original_order != indices_to_bundle) {
    throw ERROR_MSG("indices_to_bundle: ", _string_vector_to_string(indices_to_bundle),
                    " must be in its original order: ", _string_vector_to_string(original_order),
                    ".");
  }
  std::size_t bundled_dim = 1;
  for (const auto& idx : indices_to_bundle) {
    bundled_dim *= _index_to_dimension[idx];
    _index_to_dimension.erase(idx);
  }
  _index_to_dimension[bundled_index] = bundled_dim;
  std::size_t pos = 0;
  for (std::size_t i = 0; i < _indices.size(); ++i)
    if (_string_in_vector(_indices[i], indices_to_bundle)) {
      pos = i;
      break;
    }
  std::vector<std::string> new_indices(subtracted);
  new_indices.insert(new_indices.begin() + pos, bundled_index);
  std::vector<std::size_t> new_dimensions(new_indices.size());
  for (std::size_t i = 0; i < new_dimensions.size(); ++i)
    new_dimensions[i] = _index_to_dimension[new_indices[i]];
  _indices = new_indices;
  _dimensions = new_dimensions;
}

void device_reorder(const s_type* src, s_type* dst, const std::vector<std::string>& old_ordering,
                    const std::vector<std::size_t>& old_dimensions,
                    const std::vector<std::string>& new_ordering) {
  const int r = int(old_ordering.size());
  std::vector<uint64_t> dims(old_dimensions.begin(), old_dimensions.end());
  std::vector<int> perm(r);
  for (int k = 0; k < r; ++k)
    perm[k] = int(std::find(old_ordering.begin(), old_ordering.end(), new_ordering[k]) -
                  old_ordering.begin());
  if (global::verbose > 1) {
    std::cerr << "reorder dims=";
    for (int k = 0; k < r; ++k) std::cerr << (k ? "," : "") << dims[k];
    std::cerr << " perm=";
    for (int k = 0; k < r; ++k) std::cerr << (k ? "," : "") << perm[k];
    std::cerr << std::endl;
  }
  qfx_check(qfx_reorder_c64(device(), c64(src), c64(dst), r, dims.data(), perm.data()),
            "qfx_reorder_c64");
}

void Tensor::reorder(std::vector<std::string> new_ordering, s_type* scratch_copy) {
  if (scratch_copy == nullptr) throw ERROR_MSG("Scratch copy must be non-null.");
  if (new_ordering.empty() && _indices.empty()) {
    std::cerr << WARN_MSG(
                     "Warning: encountered rank-zero tensor during reorder step. This suggests "
                     "that a qubit is fully disconnected from the rest of the device.")
              << std::endl;
    return;
  }
  if (!_vector_s_in_vector_s(new_ordering, _indices) ||
      !_vector_s_in_vector_s(_indices, new_ordering) ||
      new_ordering.size() != _indices.size()) {
    throw ERROR_MSG("new_ordering: ", _string_vector_to_string(new_ordering),
                    " must be a reordering of current indices: ",
                    _string_vector_to_string(_indices), ".");
  }
  if (new_ordering == _indices) return;
  // The permutation goes into a fresh arena buffer of the same capacity,
  // which replaces _data (the reference permutes into scratch_copy and
  // copies back; the copy is saved here).  The old buffer returns to the
  // arena; work on the stream that reads it is ordered before any reuse.
  s_type* fresh = device_alloc(_capacity);
  device_reorder(_data, fresh, _indices, _dimensions, new_ordering);
  device_free(_data);
  _data = fresh;
  std::vector<std::size_t> new_dimensions(new_ordering.size());
  for (std::size_t i = 0; i < new_ordering.size(); ++i)
    new_dimensions[i] = _index_to_dimension.at(new_ordering[i]);
  _indices = new_ordering;
  _dimensions = new_dimensions;
}

void Tensor::scalar_multiply(s_type scalar) {
  auto v = host_data();
  for (auto& x : v) x *= scalar;
  set_host_data(v);
}

double Tensor::tensor_norm() const {
  double total = 0.0;
  for (const auto& x : host_data()) total += std::norm(x);
  return total;
}

std::size_t Tensor::num_zeros() const {
  std::size_t count = 0;
  for (const auto& x : host_data())
    if (x == s_type(0.)) ++count;
  return count;
}

std::string Tensor::tensor_to_string() const {
  std::string out("Tensor of rank ");
  out += std::to_string(_indices.size());
  out += ": ";
  for (std::size_t i = 0; i < _indices.size(); ++i) {
    out += _indices[i];
    out += " -> ";
    out += std::to_string(_dimensions[i]);
    if (i + 1 < _indices.size()) out += ", ";
  }
  return out;
}

void Tensor::print_data() const {
  for (const auto& x : host_data()) std::cout << x << " ";
  std::cout << std::endl;
}

/////////////////////////// EXTERNAL FUNCTIONS ////////////////////////////////

void device_multiply(const s_type* A, const s_type* B, s_type* C, std::size_t left_dim,
                     std::size_t right_dim, std::size_t common_dim, bool has_left,
                     bool has_right) {
  qfx_context* ctx = device();
  if (has_left && has_right) {
    qfx_check(qfx_cgemm_c64(ctx, global::precision, left_dim, right_dim, common_dim, c64(A),
                            c64(B), c64(C)),
              "qfx_cgemm_c64");
  } else if (has_left) {
    qfx_check(qfx_cgemv_c64(ctx, 0, left_dim, common_dim, c64(A), c64(B), c64(C)),
              "qfx_cgemv_c64");
  } else if (has_right) {
    // vM: y[right] = sum_common a[common] B[common][right]  (B, A swapped
    // as in the reference's transposed cgemv).
    qfx_check(qfx_cgemv_c64(ctx, 1, common_dim, right_dim, c64(B), c64(A), c64(C)),
              "qfx_cgemv_c64");
  } else {
    qfx_check(qfx_cdotu_c64(ctx, common_dim, c64(A), c64(B), c64(C)), "qfx_cdotu_c64");
  }
}

ProductShape prepare_multiply(Tensor& A, Tensor& B, Tensor& C, s_type* scratch_copy) {
  if (scratch_copy == nullptr) throw ERROR_MSG("Scratch copy must be non-null.");
  if (A.data() == C.data())
    throw ERROR_MSG("A and C cannot be the same tensor: ", C.tensor_to_string());
  if (B.data() == C.data())
    throw ERROR_MSG("B and C cannot be the same tensor: ", C.tensor_to_string());

  std::vector<std::string> left = _vector_subtraction(A.get_indices(), B.get_indices());
  std::vector<std::string> right = _vector_subtraction(B.get_indices(), A.get_indices());
  std::vector<std::string> common = _vector_intersection(A.get_indices(), B.get_indices());
  ProductShape p;
  for (const auto& i : left) p.left_dim *= A.get_index_to_dimension().at(i);
  for (const auto& i : right) p.right_dim *= B.get_index_to_dimension().at(i);
  for (std::size_t i = 0; i < common.size(); ++i) {
    std::size_t a_dim = A.get_index_to_dimension().at(common[i]);
    std::size_t b_dim = B.get_index_to_dimension().at(common[i]);
    if (a_dim != b_dim) {
      throw ERROR_MSG("Common indices must have matching dimensions, but at index: ", i,
                      ", the dimensions are A: ", a_dim, ", B: ", b_dim, ".");
    }
    p.common_dim *= a_dim;
  }
  if (p.left_dim * p.right_dim > C.capacity()) {
    throw ERROR_MSG("C: ", C.capacity(), " doesn't have enough space for the product of A*B: ",
                    (p.left_dim * p.right_dim), ".");
  }
  try {
    A.reorder(_vector_union(left, common), scratch_copy);
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call reorder(). Error:\n\t[", err_msg, "]");
  }
  try {
    B.reorder(_vector_union(common, right), scratch_copy);
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call reorder(). Error:\n\t[", err_msg, "]");
  }
  p.has_left = !left.empty();
  p.has_right = !right.empty();

  std::vector<std::string> C_indices = _vector_union(left, right);
  std::vector<std::size_t> C_dimensions(C_indices.size());
  for (std::size_t i = 0; i < left.size(); ++i)
    C_dimensions[i] = A.get_index_to_dimension().at(left[i]);
  for (std::size_t i = 0; i < right.size(); ++i)
    C_dimensions[i + left.size()] = B.get_index_to_dimension().at(right[i]);
  C.set_indices(C_indices);
  try {
    C.set_dimensions(C_dimensions);
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call set_dimensions(). Error:\n\t[", err_msg, "]");
  }
  C.generate_index_to_dimension();
  return p;
}

void multiply(Tensor& A, Tensor& B, Tensor& C, s_type* scratch_copy) {
  const ProductShape p = prepare_multiply(A, B, C, scratch_copy);
  device_multiply(A.data(), B.data(), C.data(), p.left_dim, p.right_dim, p.common_dim,
                  p.has_left, p.has_right);
}

std::size_t result_size(Tensor& A, Tensor& B) {
  std::size_t left_dim = 1, right_dim = 1;
  for (const auto& i : _vector_subtraction(A.get_indices(), B.get_indices()))
    left_dim *= A.get_index_to_dimension().at(i);
  for (const auto& i : _vector_subtraction(B.get_indices(), A.get_indices()))
    right_dim *= B.get_index_to_dimension().at(i);
  return left_dim * right_dim;
}

void bundle_between(Tensor& A, Tensor& B, std::string bundled_index, s_type* scratch_copy) {
  std::vector<std::string> left = _vector_subtraction(A.get_indices(), B.get_indices());
  std::vector<std::string> right = _vector_subtraction(B.get_indices(), A.get_indices());
  std::vector<std::string> common = _vector_intersection(A.get_indices(), B.get_indices());
  try {
    A.reorder(_vector_union(left, common), scratch_copy);
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call reorder(). Error:\n\t[", err_msg, "]");
  }
  try {
    B.reorder(_vector_union(common, right), scratch_copy);
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call reorder(). Error:\n\t[", err_msg, "]");
  }
  try {
    A.bundle(common, bundled_index);
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call bundle(). Error:\n\t[", err_msg, "]");
  }
  try {
    B.bundle(common, bundled_index);
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call bundle(). Error:\n\t[", err_msg, "]");
  }
}

std::string _string_vector_to_string(std::vector<std::string> input) {
  std::string out = "{";
  for (std::size_t i = 0; i < input.size(); ++i) {
    out += input[i];
    if (i + 1 < input.size()) out += ", ";
  }
  return out + "}";
}

bool _string_in_vector(const std::string& s, const std::vector<std::string>& v) {
  return std::find(v.cbegin(), v.cend(), s) != v.cend();
}

bool _vector_s_in_vector_s(const std::vector<std::string>& v, const std::vector<std::string>& w) {
  for (const auto& x : v)
    if (!_string_in_vector(x, w)) return false;
  return true;
}

std::vector<std::string> _vector_intersection(const std::vector<std::string>& v,
                                              const std::vector<std::string>& w) {
  std::vector<std::string> out;
  for (const auto& x : v)
    if (_string_in_vector(x, w)) out.push_back(x);
  return out;
}

std::vector<std::string> _vector_union(const std::vector<std::string>& v,
                                       const std::vector<std::string>& w) {
  std::vector<std::string> out(v);
  for (const auto& x : w)
    if (!_string_in_vector(x, v)) out.push_back(x);
  return out;
}

std::vector<std::string> _vector_subtraction(const std::vector<std::string>& v,
                                             const std::vector<std::string>& w) {
  std::vector<std::string> out;
  for (const auto& x : v)
    if (!_string_in_vector(x, w)) out.push_back(x);
  return out;
}

std::vector<std::string> _vector_concatenation(const std::vector<std::string>& v,
                                               const std::vector<std::string>& w) {
  std::vector<std::string> out(v);
  out.insert(out.end(), w.begin(), w.end());
  return out;
}

}  // namespace qflex
