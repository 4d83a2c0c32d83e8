// Tensor with device-resident storage.
//
// Same interface as the reference's class Tensor (src/tensor.h:34-344) and
// its free functions multiply / result_size / bundle_between
// (src/tensor.h:414-433).  Differences a caller can observe:
//   * data() returns a DEVICE pointer into the engine's arena; use
//     host_data() / set_host_data() to move values to and from the host;
//   * `scratch_copy` arguments are device buffers (at least size() elements);
//   * every data-moving or arithmetic method runs on the GPU through the
//     C-ABI of include/qfx.h.
#ifndef QFLEX_AMD_TENSOR_H_
#define QFLEX_AMD_TENSOR_H_

#include <complex>
#include <string>
#include <unordered_map>
#include <vector>

namespace qflex {

typedef std::complex<float> s_type;

class Tensor {
 public:
  Tensor();
  // Allocates size() elements of device memory, zero-filled.
  Tensor(const std::vector<std::string>& indices, const std::vector<std::size_t>& dimensions);
  // Allocates size() elements and leaves them unset: for a result that the
  // next call overwrites whole (the C of multiply), which then needs no fill.
  struct for_overwrite_t {};
  static constexpr for_overwrite_t for_overwrite{};
  Tensor(const std::vector<std::string>& indices, const std::vector<std::size_t>& dimensions,
         for_overwrite_t);
  // Allocates and uploads `data` (host values, row-major, first index slowest).
  Tensor(const std::vector<std::string>& indices, const std::vector<std::size_t>& dimensions,
         const std::vector<s_type>& data);
  Tensor(const Tensor& other);
  // noexcept, so that std::vector<Tensor> moves (instead of copying device
  // buffers) when it grows.
  Tensor(Tensor&& other) noexcept;
  ~Tensor();

  // Copies into existing capacity when this tensor has storage (throws if
  // too small), as the reference's operator= (src/tensor.cpp:189-205).
  const Tensor& operator=(const Tensor& other);
  const Tensor& operator=(Tensor&& other);

  const std::vector<std::string>& get_indices() const;
  void set_indices(const std::vector<std::string>& indices);
  const std::vector<std::size_t>& get_dimensions() const;
  void set_dimensions(const std::vector<std::size_t>& dimensions);
  void set_indices_and_dimensions(const std::vector<std::string>& indices,
                                  const std::vector<std::size_t>& dimensions);
  const std::unordered_map<std::string, std::size_t>& get_index_to_dimension() const;
  void generate_index_to_dimension();
  std::size_t size() const;
  std::size_t capacity() const;

  s_type* data();
  const s_type* data() const;
  std::vector<s_type> host_data() const;
  void set_host_data(const std::vector<s_type>& values);

  // Copies the slice index_value of indices[0] into projection_tensor.
  void project(std::string index, std::size_t index_value, Tensor& projection_tensor) const;
  void rename_index(std::string old_name, std::string new_name);
  void bundle(std::vector<std::string> indices_to_bundle, std::string bundled_index);
  // Permutes the data in place (one permutation pass into scratch_copy and
  // one copy back).
  void reorder(std::vector<std::string> new_ordering, s_type* scratch_copy);
  void scalar_multiply(s_type scalar);
  double tensor_norm() const;
  std::size_t num_zeros() const;
  std::string tensor_to_string() const;
  void print_data() const;

 private:
  std::vector<std::string> _indices;
  std::vector<std::size_t> _dimensions;
  std::unordered_map<std::string, std::size_t> _index_to_dimension;
  s_type* _data{nullptr};
  std::size_t _capacity{0};

  void _init(const std::vector<std::string>& indices, const std::vector<std::size_t>& dimensions);
  void _clear();
  void _copy(const Tensor& other);
  void _move(Tensor&& other) noexcept;
};

// C = A . B over the shared indices; A and B are reordered in place to
// (left, common) and (common, right); C gets indices (left, right).
// scratch_copy: device buffer of at least max(A.size(), B.size()) elements.
void multiply(Tensor& A, Tensor& B, Tensor& C, s_type* scratch_copy);

// The host half of multiply: the checks, the index algebra, the reorders of
// A to (left, common) and B to (common, right), and C's indices and
// dimensions; the product itself is left to the caller (device_multiply, or
// a grouped launch of several products).
struct ProductShape {
  std::size_t left_dim = 1, right_dim = 1, common_dim = 1;
  bool has_left = false, has_right = false;
};
ProductShape prepare_multiply(Tensor& A, Tensor& B, Tensor& C, s_type* scratch_copy);

std::size_t result_size(Tensor& A, Tensor& B);

void bundle_between(Tensor& A, Tensor& B, std::string bundled_index, s_type* scratch_copy);

// Device allocation helpers (engine arena).
s_type* device_alloc(std::size_t elements);
void device_free(s_type* p);

// Issues C[m x n] = A[m x k] . B[k x n] with the BLAS-case dispatch of the
// reference (MM / Mv / vM / vv, src/tensor.cpp:1127-1158) on device pointers.
void device_multiply(const s_type* A, const s_type* B, s_type* C, std::size_t left_dim,
                     std::size_t right_dim, std::size_t common_dim, bool has_left,
                     bool has_right);

// dst = src permuted from old_ordering to new_ordering (device pointers).
void device_reorder(const s_type* src, s_type* dst, const std::vector<std::string>& old_ordering,
                    const std::vector<std::size_t>& old_dimensions,
                    const std::vector<std::string>& new_ordering);

std::string _string_vector_to_string(std::vector<std::string> input);
bool _string_in_vector(const std::string& s, const std::vector<std::string>& v);
bool _vector_s_in_vector_s(const std::vector<std::string>& v, const std::vector<std::string>& w);
std::vector<std::string> _vector_intersection(const std::vector<std::string>& v,
                                              const std::vector<std::string>& w);
std::vector<std::string> _vector_union(const std::vector<std::string>& v,
                                       const std::vector<std::string>& w);
std::vector<std::string> _vector_subtraction(const std::vector<std::string>& v,
                                             const std::vector<std::string>& w);
std::vector<std::string> _vector_concatenation(const std::vector<std::string>& v,
                                               const std::vector<std::string>& w);

}  // namespace qflex

#endif
