// Circuit -> 3D tensor network -> 2D grid of site tensors (interface of
// reference src/read_circuit.h:24-103).  Tensors live on the device; the
// per-qubit time-line products, the edge bundling and the final site
// reorders run through the same C-ABI as the contraction.
#ifndef QFLEX_AMD_READ_CIRCUIT_H_
#define QFLEX_AMD_READ_CIRCUIT_H_

#include <functional>
#include <list>
#include <optional>
#include <string>
#include <tuple>
#include <vector>

#include "circuit.h"
#include "contraction_utils.h"
#include "tensor.h"

namespace qflex {

// Gate data in (input, output) order (one-qubit) or as the two rank-3
// (input, virtual, output) factors of a two-qubit gate.
std::vector<s_type> gate_array(const std::string& gate_name, const std::vector<double>& params);
std::tuple<std::vector<s_type>, std::vector<s_type>, std::vector<std::size_t>> gate_arrays(
    const std::string& gate_name, const std::vector<double>& params);

std::function<bool(std::vector<std::size_t>, std::vector<std::size_t>)> order_func(
    const std::list<ContractionOperation>& ordering, std::vector<std::size_t> local);

void circuit_data_to_tensor_network(
    const QflexCircuit& circuit, std::size_t I, std::size_t J, const std::string initial_conf,
    const std::string final_conf,
    const std::optional<std::vector<std::vector<std::size_t>>>& final_qubit_region,
    const std::optional<std::vector<std::vector<std::size_t>>>& off,
    std::vector<std::vector<std::vector<Tensor>>>& grid_of_tensors, s_type* scratch);

void flatten_grid_of_tensors(
    std::vector<std::vector<std::vector<Tensor>>>& grid_of_tensors,
    std::vector<std::vector<Tensor>>& grid_of_tensors_2D,
    const std::optional<std::vector<std::vector<std::size_t>>>& final_qubit_region,
    const std::optional<std::vector<std::vector<std::size_t>>>& off,
    const std::list<ContractionOperation>& ordering, s_type* scratch);

}  // namespace qflex

#endif
