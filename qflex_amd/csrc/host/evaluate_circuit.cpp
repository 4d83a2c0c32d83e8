// EvaluateCircuit (reference src/evaluate_circuit.cpp:22-253).
#include "evaluate_circuit.h"

#include <algorithm>
#include <chrono>
#include <iostream>
#include <unordered_map>

#include "errors.h"

namespace qflex {

std::size_t find_output_pos(const QflexInput* input, std::vector<std::size_t> tensor_pos) {
  if (input == nullptr) throw ERROR_MSG("Input must be non-null.");
  std::size_t pos = tensor_pos[0] * input->grid.J + tensor_pos[1];
  for (const auto& off : input->grid.qubits_off)
    if (off[0] < tensor_pos[0] || (off[0] == tensor_pos[0] && off[1] < tensor_pos[1])) --pos;
  return pos;
}

std::string get_output_states(const QflexInput* input,
                              const std::list<ContractionOperation>& ordering,
                              std::vector<std::vector<std::size_t>>* final_qubits,
                              std::vector<std::string>* output_states) {
  if (input == nullptr) throw ERROR_MSG("Input must be non-null.");
  if (final_qubits == nullptr) throw ERROR_MSG("Final qubits must be non-null.");
  if (output_states == nullptr) throw ERROR_MSG("Output states must be non-null");
  std::vector<std::size_t> pos_map;
  std::vector<std::vector<std::size_t>> values_map;
  std::string base_state = input->final_state;
  const bool unspecified = input->final_state.empty();
  if (unspecified) base_state = std::string(input->initial_state.length(), '0');
  for (const auto& op : ordering) {
    if (op.op_type != ContractionOperation::CUT || op.cut.tensors.size() != 1) continue;
    const std::size_t output_pos = find_output_pos(input, op.cut.tensors[0]);
    if (unspecified) base_state.at(output_pos) = 'x';
    pos_map.push_back(output_pos);
    values_map.push_back(op.cut.values.empty() ? std::vector<std::size_t>{0, 1} : op.cut.values);
    final_qubits->push_back(op.cut.tensors[0]);
  }
  output_states->push_back(base_state);
  for (std::size_t i = 0; i < final_qubits->size(); ++i) {
    std::vector<std::string> next;
    for (const auto& state : *output_states)
      for (const auto val : values_map[i]) {
        std::string s = state;
        s.at(pos_map[i]) = std::to_string(val).at(0);
        next.push_back(s);
      }
    *output_states = next;
  }
  for (std::size_t i = 0; i < output_states->at(0).length(); ++i) {
    const char c = output_states->at(0)[i];
    if (c != '0' && c != '1')
      throw ERROR_MSG("Final state has non-binary character ", c, " at index ", i,
                      "despite having no terminal cut there.");
  }
  return base_state;
}

namespace {

// Ordering, output strings and the flattened 2D grid of an input (the part
// of EvaluateCircuit before ContractGrid).
void build_network(QflexInput* input, std::list<ContractionOperation>* ordering,
                   std::vector<std::string>* output_states,
                   std::vector<std::vector<Tensor>>* tensor_grid_out) {
  using clock = std::chrono::steady_clock;
  auto t0 = clock::now();
  auto lap = [&t0](const char* what) {
    if (global::verbose > 0) {
      auto t1 = clock::now();
      std::cerr << WARN_MSG(what, std::chrono::duration<double>(t1 - t0).count(), "s")
                << std::endl;
      t0 = t1;
    }
  };
  auto& tensor_grid = *tensor_grid_out;
  ordering_data_to_contraction_ordering(*input, ordering);
  lap("Time spent making contraction ordering: ");

  const std::size_t init_length = input->grid.I * input->grid.J - input->grid.qubits_off.size();
  if (input->initial_state.empty()) input->initial_state = std::string(init_length, '0');

  std::vector<std::vector<std::size_t>> final_qubits;
  try {
    input->final_state = get_output_states(input, *ordering, &final_qubits, output_states);
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call get_output_states(). Error:\n\t[", err_msg, "]");
  }

  tensor_grid.assign(input->grid.I, std::vector<Tensor>(input->grid.J));
  {
    std::vector<std::vector<std::vector<Tensor>>> tensor_grid_3D;
    // Gate tensors have at most 16 elements.
    s_type* scratch_3D = device_alloc(16);
    try {
      circuit_data_to_tensor_network(input->circuit, input->grid.I, input->grid.J,
                                     input->initial_state, input->final_state, final_qubits,
                                     input->grid.qubits_off, tensor_grid_3D, scratch_3D);
    } catch (...) {
      device_free(scratch_3D);
      throw;
    }
    device_free(scratch_3D);
    lap("Time spent creating 3D grid of tensors from file: ");

    // Largest site tensor after the time lines are contracted.
    std::size_t max_size = 1;
    for (auto& row : tensor_grid_3D)
      for (auto& line : row) {
        std::unordered_map<std::string, std::size_t> index_dim;
        for (const auto& t : line)
          for (const auto& [index, dim] : t.get_index_to_dimension()) {
            if (index_dim.count(index))
              index_dim.erase(index);
            else
              index_dim[index] = dim;
          }
        std::size_t s = 1;
        for (const auto& p : index_dim) s *= p.second;
        max_size = std::max(max_size, s);
        // intermediate products along the time line can be larger
        std::size_t run = 1;
        for (const auto& t : line) run = std::max(run, t.size());
        max_size = std::max(max_size, run);
      }
    s_type* scratch_2D = device_alloc(max_size * 2);
    try {
      flatten_grid_of_tensors(tensor_grid_3D, tensor_grid, final_qubits, input->grid.qubits_off,
                              *ordering, scratch_2D);
    } catch (...) {
      device_free(scratch_2D);
      throw;
    }
    device_free(scratch_2D);
    lap("Time spent creating 2D grid of tensors from 3D one: ");
  }

}

}  // namespace

std::vector<std::pair<std::string, std::complex<double>>> EvaluateCircuit(QflexInput* input) {
  if (input == nullptr) throw ERROR_MSG("Input must be non-null.");
  using clock = std::chrono::steady_clock;
  auto t0 = clock::now();
  auto lap = [&t0](const char* what) {
    if (global::verbose > 0) {
      auto t1 = clock::now();
      std::cerr << WARN_MSG(what, std::chrono::duration<double>(t1 - t0).count(), "s")
                << std::endl;
      t0 = t1;
    }
  };
  std::list<ContractionOperation> ordering;
  std::vector<std::string> output_states;
  std::vector<std::vector<Tensor>> tensor_grid;
  build_network(input, &ordering, &output_states, &tensor_grid);
  t0 = clock::now();

  std::vector<std::complex<double>> amplitudes(output_states.size());
  try {
    ContractGrid(ordering, &tensor_grid, &amplitudes);
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call ContractGrid(). Error:\n\t[", err_msg, "]");
  }
  lap("Time spent contracting tensors: ");
  std::vector<std::pair<std::string, std::complex<double>>> result;
  for (std::size_t c = 0; c < amplitudes.size(); ++c)
    result.push_back({output_states[c], amplitudes[c]});
  return result;
}

std::vector<std::pair<std::string, std::complex<double>>> EvaluateCircuitStates(
    QflexInput* input, const std::vector<std::string>& final_states) {
  if (input == nullptr) throw ERROR_MSG("Input must be non-null.");
  if (final_states.empty()) throw ERROR_MSG("At least one final state is needed.");
  if (final_states.size() == 1) {
    input->final_state = final_states[0];
    return EvaluateCircuit(input);
  }
  std::list<ContractionOperation> ordering;
  ordering_data_to_contraction_ordering(*input, &ordering);
  const std::size_t active = input->grid.I * input->grid.J - input->grid.qubits_off.size();
  if (input->initial_state.empty()) input->initial_state = std::string(active, '0');

  // Each state's output strings (terminal cuts enumerated) and base string
  // ('x' at the terminal positions).
  std::vector<std::vector<std::size_t>> terminal;
  std::vector<std::string> bases;
  std::vector<std::vector<std::string>> outputs;
  for (const auto& state : final_states) {
    QflexInput one = *input;
    one.final_state = state;
    std::vector<std::vector<std::size_t>> fq;
    std::vector<std::string> outs;
    std::string base;
    try {
      base = get_output_states(&one, ordering, &fq, &outs);
    } catch (const std::string& err_msg) {
      throw ERROR_MSG("Failed to call get_output_states(). Error:\n\t[", err_msg, "]");
    }
    if (base.size() != active)
      throw ERROR_MSG("Final state '", state, "' has ", base.size(), " qubits instead of ", active,
                      ".");
    terminal = fq;
    bases.push_back(base);
    outputs.push_back(outs);
  }
  // Qubits whose bit differs between the states keep an open output index
  // (the terminal-cut qubits are open already).
  std::vector<std::size_t> terminal_pos;
  for (const auto& q : terminal) terminal_pos.push_back(find_output_pos(input, q));
  std::vector<std::vector<std::size_t>> open_qubits = terminal;
  std::vector<std::vector<std::size_t>> varying;  // grid coordinates
  std::vector<std::size_t> varying_pos;           // positions in the bitstrings
  {
    std::size_t pos = 0;
    for (std::size_t i = 0; i < input->grid.I; ++i)
      for (std::size_t j = 0; j < input->grid.J; ++j) {
        if (find_grid_coord_in_list(input->grid.qubits_off, i, j)) continue;
        const char c0 = bases[0][pos];
        bool differs = false;
        for (const auto& b : bases) differs = differs || b[pos] != c0;
        const bool is_terminal =
            std::find(terminal_pos.begin(), terminal_pos.end(), pos) != terminal_pos.end();
        if (differs && !is_terminal) {
          varying.push_back({i, j});
          varying_pos.push_back(pos);
          open_qubits.push_back({i, j});
        }
        ++pos;
      }
  }
  std::vector<std::vector<Tensor>> tensor_grid(input->grid.I, std::vector<Tensor>(input->grid.J));
  {
    std::vector<std::vector<std::vector<Tensor>>> grid3D;
    s_type* scratch_3D = device_alloc(16);
    try {
      circuit_data_to_tensor_network(input->circuit, input->grid.I, input->grid.J,
                                     input->initial_state, bases[0], open_qubits,
                                     input->grid.qubits_off, grid3D, scratch_3D);
    } catch (...) {
      device_free(scratch_3D);
      throw;
    }
    device_free(scratch_3D);
    std::size_t max_size = 1;
    for (auto& row : grid3D)
      for (auto& line : row) {
        std::size_t run = 1;
        for (const auto& t : line) run = std::max(run, t.size());
        max_size = std::max(max_size, run);
      }
    // a site's product along its time line: at most 2^4 per bond index and
    // 2 per open output, times the largest single gate tensor
    std::size_t site_max = 1;
    for (auto& row : grid3D)
      for (auto& line : row) {
        std::unordered_map<std::string, std::size_t> index_dim;
        for (const auto& t : line)
          for (const auto& [index, dim] : t.get_index_to_dimension()) {
            if (index_dim.count(index))
              index_dim.erase(index);
            else
              index_dim[index] = dim;
          }
        std::size_t sz = 1;
        for (const auto& kv : index_dim) sz *= kv.second;
        site_max = std::max(site_max, sz);
      }
    s_type* scratch_2D = device_alloc(std::max(max_size, site_max) * 2);
    try {
      flatten_grid_of_tensors(grid3D, tensor_grid, open_qubits, input->grid.qubits_off, ordering,
                              scratch_2D);
    } catch (...) {
      device_free(scratch_2D);
      throw;
    }
    device_free(scratch_2D);
  }
  // Per state: the varying sites' output indices fixed to the state's bits
  // (the open output index is each such tensor's first, so the projections
  // are views).
  std::vector<std::string> projected;
  for (const auto& q : varying) projected.push_back(index_name(q, {}));
  std::vector<std::vector<ContractionData::SiteProjection>> states;
  for (const auto& base : bases) {
    std::vector<ContractionData::SiteProjection> ps;
    for (std::size_t v = 0; v < varying.size(); ++v)
      ps.push_back({varying[v][0], varying[v][1], projected[v],
                    std::size_t(base[varying_pos[v]] == '1' ? 1 : 0)});
    states.push_back(ps);
  }
  const std::size_t per_state = outputs[0].size();
  std::vector<std::complex<double>> amplitudes(per_state * final_states.size());
  ContractionData data;
  try {
    data = ContractionData::Initialize(ordering, &tensor_grid, &amplitudes, projected);
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call Initialize(). Error:\n\t[", err_msg, "]");
  }
  std::unordered_map<std::string, bool> active_patches;
  for (const auto& patch : data.scratch_list()) active_patches[patch] = false;
  try {
    data.ContractGridStates(ordering, active_patches, states);
  } catch (const std::string& err_msg) {
    throw ERROR_MSG("Failed to call ContractGrid(). Error:\n\t[", err_msg, "]");
  }
  std::vector<std::pair<std::string, std::complex<double>>> result;
  for (std::size_t s = 0; s < final_states.size(); ++s)
    for (std::size_t c = 0; c < per_state; ++c)
      result.push_back({outputs[s][c], amplitudes[s * per_state + c]});
  return result;
}

SymbolicPlan PlanCircuit(QflexInput* input) {
  if (input == nullptr) throw ERROR_MSG("Input must be non-null.");
  std::list<ContractionOperation> ordering;
  std::vector<std::string> output_states;
  std::vector<std::vector<Tensor>> tensor_grid;
  build_network(input, &ordering, &output_states, &tensor_grid);
  return plan_contraction(ordering, tensor_grid);
}

}  // namespace qflex
