// Circuit -> tensor network (reference src/read_circuit.cpp:27-689).
#include "read_circuit.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <unordered_map>
#include <unordered_set>

#include "device.h"
#include "errors.h"
#include "global.h"

namespace qflex {

namespace {

using f = s_type::value_type;
const f kR = static_cast<f>(M_SQRT1_2);

// Same constants as the reference's gate table (src/read_circuit.cpp:27-79):
// identical values keep the networks bit-compatible.  cz/cx are split into
// two rank-3 factors with a dimension-2 virtual index ("balanced" SVD).
const std::unordered_map<std::string, std::vector<s_type>>& gate_table() {
  static const std::unordered_map<std::string, std::vector<s_type>> t = {
      {"delta_0", {1.0, 0.0}},
      {"delta_1", {0.0, 1.0}},
      {"h", {kR, kR, kR, -kR}},
      {"hz_1_2", {{0.5, 0.5}, {kR, 0}, {0., -kR}, {0.5, 0.5}}},
      {"t", {1.0, 0., 0., {kR, kR}}},
      {"x_1_2", {{0.5, 0.5}, {0.5, -0.5}, {0.5, -0.5}, {0.5, 0.5}}},
      {"y_1_2", {{0.5, 0.5}, {0.5, 0.5}, {-0.5, -0.5}, {0.5, 0.5}}},
      {"cz_q1",
       {-0.3446133714352872, 0., 1.1381806476131544, 0., 0., -1.1381806476131544, 0.,
        -0.3446133714352872}},
      {"cz_q2",
       {-1.0484937059720079, 0., 0.5611368023131075, 0., 0., 0.5611368023131075, 0.,
        1.0484937059720079}},
      {"cz", {1., 0., 0., 0., 0., 1., 0., 0., 0., 0., 1., 0., 0., 0., 0., -1.}},
      {"cx_q1",
       {0.8408964152537143, 0., 0.8408964152537143, 0., 0., -0.8408964152537143, 0.,
        0.8408964152537143}},
      {"cx_q2",
       {0.5946035575013604, -0.5946035575013604, 0.5946035575013604, 0.5946035575013604,
        -0.5946035575013604, 0.5946035575013604, 0.5946035575013604, 0.5946035575013604}},
      {"cx", {1., 0., 0., 0., 0., 1., 0., 0., 0., 0., 0., 1., 0., 0., 1., 0.}}};
  return t;
}

// fSim(theta, phi) split into two identical rank-3 (input, virtual(4),
// output) factors; theta, phi in radians (src/read_circuit.cpp:105-139).
std::vector<s_type> fsim_factor(f theta, f phi) {
  const s_type half(0.5);
  const s_type e = std::exp(s_type(0.0, -phi / 2));
  const s_type c1 = half * (e + s_type(std::cos(theta), 0.0));
  const s_type c2 = half * (e - s_type(std::cos(theta), 0.0));
  const s_type c3 = s_type(0.0, -0.5) * s_type(std::sin(theta), 0.0);
  const s_type s1 = std::sqrt(c1), s2 = std::sqrt(c2), s3 = std::sqrt(c3), s4 = std::sqrt(c3);
  const s_type d = std::exp(s_type(0.0, phi / 4));
  const s_type z(0.0, 0.0), i1(0.0, 1.0);
  return {d * s1, z, d * s2, z, z, s3, z, i1 * s4,
          z, std::conj(d) * s1, z, -std::conj(d) * s2, s3, z, -i1 * s4, z};
}

std::string site_index(std::size_t i, std::size_t j, std::size_t k) {
  return "(" + std::to_string(i) + "," + std::to_string(j) + "),(" + std::to_string(k) + ")";
}

}  // namespace

std::vector<s_type> gate_array(const std::string& gate_name, const std::vector<double>& params) {
  if (gate_name == "rz") {
    if (params.empty()) throw ERROR_MSG("Gate rz requires one parameter.");
    const double angle_rads = M_PI * params[0];
    const s_type phase = std::exp(s_type(0., angle_rads / 2));
    return {std::conj(phase), 0., 0., phase};
  }
  const auto& t = gate_table();
  auto it = t.find(gate_name);
  if (it == t.end()) throw ERROR_MSG("Invalid gate name provided: ", gate_name);
  return it->second;
}

std::tuple<std::vector<s_type>, std::vector<s_type>, std::vector<std::size_t>> gate_arrays(
    const std::string& gate_name, const std::vector<double>& params) {
  if (gate_name == "cz") return {gate_array("cz_q1", params), gate_array("cz_q2", params), {2, 2, 2}};
  if (gate_name == "cx") return {gate_array("cx_q1", params), gate_array("cx_q2", params), {2, 2, 2}};
  if (gate_name == "fsim") {
    if (params.size() < 2) throw ERROR_MSG("Gate fsim requires two parameters.");
    auto a = fsim_factor(f(M_PI * params[0]), f(M_PI * params[1]));
    return {a, a, {2, 4, 2}};
  }
  throw ERROR_MSG("Invalid gate name provided: ", gate_name);
}

// Order in which the neighbours of `local` meet it during the contraction:
// cut bonds first (in cut order), then by patch and expansion position,
// then by the order in which patches merge (src/read_circuit.cpp:191-291).
std::function<bool(std::vector<std::size_t>, std::vector<std::size_t>)> order_func(
    const std::list<ContractionOperation>& ordering, std::vector<std::size_t> local) {
  return [&ordering, local](const std::vector<std::size_t> lhs,
                            const std::vector<std::size_t> rhs) {
    auto pair_of = [&local](const std::vector<std::size_t>& x) {
      if (local[0] < x[0] || local[1] < x[1])
        return std::vector<std::vector<std::size_t>>{local, x};
      return std::vector<std::vector<std::size_t>>{x, local};
    };
    const auto lhs_pair = pair_of(lhs), rhs_pair = pair_of(rhs);
    for (const auto& op : ordering) {
      if (op.op_type != ContractionOperation::CUT) continue;
      if (lhs_pair == op.cut.tensors) return true;
      if (rhs_pair == op.cut.tensors) return false;
    }
    auto find_expand = [&ordering](const std::vector<std::size_t>& pos, std::string* patch) {
      std::size_t n = 0;
      for (const auto& op : ordering) {
        if (op.op_type != ContractionOperation::EXPAND) continue;
        if (op.expand.tensor == pos) {
          *patch = op.expand.id;
          return std::optional<std::size_t>(n);
        }
        ++n;
      }
      return std::optional<std::size_t>();
    };
    std::string lpatch = "null", rpatch = "null";
    auto lpos = find_expand(lhs, &lpatch);
    if (!lpos.has_value())
      throw ERROR_MSG("Left hand side of pair not found: (", local[0], ',', local[1], "),(",
                      lhs[0], ',', lhs[1], ").");
    auto rpos = find_expand(rhs, &rpatch);
    if (!rpos.has_value())
      throw ERROR_MSG("Right hand side of pair not found: (", local[0], ',', local[1], "),(",
                      rhs[0], ',', rhs[1], ").");
    if (lpatch == rpatch) return lpos < rpos;
    std::string local_patch;
    for (const auto& op : ordering)
      if (op.op_type == ContractionOperation::EXPAND && op.expand.tensor == local) {
        local_patch = op.expand.id;
        break;
      }
    if (lpatch == local_patch) return true;
    if (rpatch == local_patch) return false;
    for (const auto& op : ordering) {
      if (op.op_type != ContractionOperation::MERGE) continue;
      const auto& src = op.merge.source_id;
      const auto& tgt = op.merge.target_id;
      if (local_patch == src) {
        if (lpatch == tgt) return true;
        if (rpatch == tgt) return false;
        local_patch = tgt;
      } else if (local_patch == tgt) {
        if (lpatch == src) return true;
        if (rpatch == src) return false;
      } else if (lpatch == src) {
        lpatch = tgt;
      } else if (rpatch == src) {
        rpatch = tgt;
      }
    }
    throw ERROR_MSG("Failed to compare (", lhs[0], ',', lhs[1], ") and (", rhs[0], ',', rhs[1],
                    ") for local (", local[0], ',', local[1], ").");
  };
}

void circuit_data_to_tensor_network(
    const QflexCircuit& circuit, std::size_t I, std::size_t J, const std::string initial_conf,
    const std::string final_conf,
    const std::optional<std::vector<std::vector<std::size_t>>>& final_qubit_region,
    const std::optional<std::vector<std::vector<std::size_t>>>& off,
    std::vector<std::vector<std::vector<Tensor>>>& grid_of_tensors, s_type* scratch) {
  if (scratch == nullptr) throw ERROR_MSG("Scratch must be non-null.");
  const std::size_t grid_size = I * J;
  const std::size_t off_size = off.has_value() ? off.value().size() : 0;
  const std::size_t num_active = grid_size - off_size;
  if (circuit.num_active_qubits != num_active)
    throw ERROR_MSG("The number of active qubits read from the file: ", circuit.num_active_qubits,
                    ", does not match the number of active qubits read from the grid: ",
                    num_active, ".");
  if (initial_conf.size() != num_active)
    throw ERROR_MSG("Size of initial_conf: ", initial_conf.size(),
                    ", must be equal to the number of qubits: ", num_active, ".");
  if (final_conf.size() != initial_conf.size())
    throw ERROR_MSG("Size of final_conf: ", final_conf.size(),
                    ", must be equal to size of initial_conf: ", initial_conf.size(), ".");

  grid_of_tensors.assign(I, std::vector<std::vector<Tensor>>(J));
  std::vector<std::vector<std::size_t>> counters(I, std::vector<std::size_t>(J, 0));
  std::unordered_map<std::string, std::size_t> link_counters;

  std::size_t idx = 0;
  for (std::size_t q = 0; q < grid_size; ++q) {
    const std::size_t i = q / J, j = q % J;
    if (find_grid_coord_in_list(off, i, j)) continue;
    const std::string delta = (initial_conf[idx] == '0') ? "delta_0" : "delta_1";
    try {
      grid_of_tensors[i][j].push_back(Tensor({site_index(i, j, counters[i][j])}, {2},
                                             gate_array(delta, {})));
    } catch (const std::string& err_msg) {
      throw ERROR_MSG("Failed to call Tensor(). Error:\n\t[", err_msg, "]");
    }
    ++idx;
  }
  if (off.has_value())
    for (const auto& w : off.value())
      if (w[0] >= I || w[1] >= J)
        throw ERROR_MSG("Off qubit '(", w[0], ", ", w[1], ")' is outside the grid.");

  std::unordered_set<std::size_t> used_qubits;
  std::size_t last_cycle = 0;
  for (const auto& gate : circuit.gates) {
    if (last_cycle != gate.cycle) {
      last_cycle = gate.cycle;
      used_qubits.clear();
    }
    for (const auto& q : gate.qubits)
      if (!used_qubits.insert(q).second)
        throw ERROR_MSG("Qubit ", q, " has been used twice in the same cycle: ", gate.cycle);
    if (gate.qubits.size() == 1) {
      const std::size_t q1 = gate.qubits[0];
      const auto p = get_qubit_coords(q1, J);
      if (p[0] >= I || find_grid_coord_in_list(off, p[0], p[1]))
        throw ERROR_MSG("Qubit '", q1, "' in gate '", gate.raw,
                        "' must correspond to an active qubit.");
      const std::string in = site_index(p[0], p[1], counters[p[0]][p[1]]);
      ++counters[p[0]][p[1]];
      const std::string out = site_index(p[0], p[1], counters[p[0]][p[1]]);
      try {
        grid_of_tensors[p[0]][p[1]].push_back(
            Tensor({in, out}, {2, 2}, gate_array(gate.name, gate.params)));
      } catch (const std::string& err_msg) {
        throw ERROR_MSG("Failed to call Tensor(). Error:\n\t[", err_msg, "]");
      }
    } else if (gate.qubits.size() == 2) {
      const std::size_t q1 = gate.qubits[0], q2 = gate.qubits[1];
      const auto p1 = get_qubit_coords(q1, J), p2 = get_qubit_coords(q2, J);
      if (p1[0] >= I || find_grid_coord_in_list(off, p1[0], p1[1]))
        throw ERROR_MSG("Qubit '", q1, "' in gate '", gate.raw,
                        "' must correspond to an active qubit.");
      if (p2[0] >= I || find_grid_coord_in_list(off, p2[0], p2[1]))
        throw ERROR_MSG("Qubit '", q2, "' in gate '", gate.raw,
                        "' must correspond to an active qubit.");
      const std::size_t dx = std::max(p1[0], p2[0]) - std::min(p1[0], p2[0]);
      const std::size_t dy = std::max(p1[1], p2[1]) - std::min(p1[1], p2[1]);
      if (dx + dy != 1) throw ERROR_MSG("Qubits ", q1, " and ", q2, " are not nearest neighbors.");
      std::vector<s_type> g1, g2;
      std::vector<std::size_t> dims;
      std::tie(g1, g2, dims) = gate_arrays(gate.name, gate.params);
      const std::string link = index_name(p1, p2);
      const std::size_t counter = ++link_counters[link];
      const std::string virt =
          "(" + std::to_string(std::min(p1[0], p2[0])) + "," +
          std::to_string(std::min(p1[1], p2[1])) + "," + std::to_string(counter) + "),(" +
          std::to_string(std::max(p1[0], p2[0])) + "," + std::to_string(std::max(p1[1], p2[1])) +
          "," + std::to_string(counter) + ")";
      const std::string in1 = site_index(p1[0], p1[1], counters[p1[0]][p1[1]]);
      const std::string in2 = site_index(p2[0], p2[1], counters[p2[0]][p2[1]]);
      ++counters[p1[0]][p1[1]];
      ++counters[p2[0]][p2[1]];
      const std::string out1 = site_index(p1[0], p1[1], counters[p1[0]][p1[1]]);
      const std::string out2 = site_index(p2[0], p2[1], counters[p2[0]][p2[1]]);
      try {
        grid_of_tensors[p1[0]][p1[1]].push_back(Tensor({in1, virt, out1}, dims, g1));
        grid_of_tensors[p2[0]][p2[1]].push_back(Tensor({in2, virt, out2}, dims, g2));
      } catch (const std::string& err_msg) {
        throw ERROR_MSG("Failed to call Tensor(). Error:\n\t[", err_msg, "]");
      }
    } else {
      throw ERROR_MSG("k-qubit gates with k > 2 not yet implemented.");
    }
  }

  // Close each time line: a delta for the final bit, or an open "(i,j),(o)"
  // index for qubits of the final region.
  idx = 0;
  for (std::size_t q = 0; q < grid_size; ++q) {
    const std::size_t i = q / J, j = q % J;
    if (find_grid_coord_in_list(off, i, j)) continue;
    const std::string last = site_index(i, j, counters[i][j]);
    if (find_grid_coord_in_list(final_qubit_region, i, j)) {
      try {
        grid_of_tensors[i][j].back().rename_index(
            last, "(" + std::to_string(i) + "," + std::to_string(j) + "),(o)");
      } catch (const std::string& err_msg) {
        throw ERROR_MSG("Failed to call rename_index(). Error:\n\t[", err_msg, "]");
      }
    } else {
      const std::string delta = (final_conf[idx] == '0') ? "delta_0" : "delta_1";
      try {
        grid_of_tensors[i][j].push_back(Tensor({last}, {2}, gate_array(delta, {})));
      } catch (const std::string& err_msg) {
        throw ERROR_MSG("Failed to call Tensor(). Error:\n\t[", err_msg, "]");
      }
    }
    ++idx;
  }
}

void flatten_grid_of_tensors(
    std::vector<std::vector<std::vector<Tensor>>>& grid_of_tensors,
    std::vector<std::vector<Tensor>>& grid_of_tensors_2D,
    const std::optional<std::vector<std::vector<std::size_t>>>& final_qubit_region,
    const std::optional<std::vector<std::vector<std::size_t>>>& off,
    const std::list<ContractionOperation>& ordering, s_type* scratch) {
  if (scratch == nullptr) throw ERROR_MSG("Scratch must be non-null.");
  const std::size_t I = grid_of_tensors.size();
  // 1. Contract each qubit's time line into one site tensor, one gate layer
  //    at a time over all sites (the reference runs each site's line to the
  //    end; the products are the same): the matrix products of a layer go
  //    out as one grouped launch, products with a vector operand one by one.
  std::vector<std::array<std::size_t, 2>> sites;
  std::vector<Tensor> line_product;
  std::size_t depth = 0;
  for (std::size_t i = 0; i < I; ++i) {
    const std::size_t J = grid_of_tensors[i].size();
    for (std::size_t j = 0; j < J; ++j) {
      if (find_grid_coord_in_list(off, i, j)) continue;
      if (grid_of_tensors[i][j].empty()) throw ERROR_MSG("Time-direction cannot be empty");
      sites.push_back({i, j});
      line_product.push_back(grid_of_tensors[i][j][0]);
      depth = std::max(depth, grid_of_tensors[i][j].size());
    }
  }
  std::vector<Tensor> next(sites.size());
  std::vector<qfx_gemm_desc> layer;
  for (std::size_t k = 1; k < depth; ++k) {
    layer.clear();
    std::vector<std::size_t> stepped;
    for (std::size_t s = 0; s < sites.size(); ++s) {
      auto& line = grid_of_tensors[sites[s][0]][sites[s][1]];
      if (k >= line.size()) continue;
      Tensor& A = line_product[s];
      Tensor& B = line[k];
      Tensor C({""}, {result_size(A, B)}, Tensor::for_overwrite);  // written whole below
      try {
        const ProductShape p = prepare_multiply(A, B, C, scratch);
        if (p.has_left && p.has_right) {
          layer.push_back(qfx_gemm_desc{p.left_dim, p.right_dim, p.common_dim,
                                        reinterpret_cast<const qfx_c64*>(A.data()),
                                        reinterpret_cast<const qfx_c64*>(B.data()),
                                        reinterpret_cast<qfx_c64*>(C.data())});
        } else {
          device_multiply(A.data(), B.data(), C.data(), p.left_dim, p.right_dim, p.common_dim,
                          p.has_left, p.has_right);
        }
      } catch (const std::string& err_msg) {
        throw ERROR_MSG("Failed to call multiply(). Error:\n\t[", err_msg, "]");
      }
      next[s] = std::move(C);
      stepped.push_back(s);
    }
    if (!layer.empty())
      qfx_check(qfx_cgemm_grouped_c64(device(), global::precision, int(layer.size()), layer.data()),
                "qfx_cgemm_grouped_c64");
    for (std::size_t s : stepped) line_product[s] = std::move(next[s]);
  }
  for (std::size_t s = 0; s < sites.size(); ++s)
    grid_of_tensors_2D[sites[s][0]][sites[s][1]] = std::move(line_product[s]);
  // 2. Bundle the virtual indices shared by neighbouring sites into one bond.
  for (std::size_t i = 0; i < I; ++i) {
    const std::size_t J = grid_of_tensors[i].size();
    for (std::size_t j = 0; j < J; ++j) {
      if (find_grid_coord_in_list(off, i, j)) continue;
      std::vector<std::vector<std::size_t>> pairs;
      if (i < I - 1 && !find_grid_coord_in_list(off, i + 1, j)) pairs.push_back({i + 1, j});
      if (j < J - 1 && !find_grid_coord_in_list(off, i, j + 1)) pairs.push_back({i, j + 1});
      for (const auto& pair : pairs)
        bundle_between(grid_of_tensors_2D[i][j], grid_of_tensors_2D[pair[0]][pair[1]],
                       index_name({i, j}, pair), scratch);
    }
  }
  // 3. Put each site's indices in the order the contraction consumes them.
  for (std::size_t i = 0; i < I; ++i) {
    const std::size_t J = grid_of_tensors[i].size();
    for (std::size_t j = 0; j < J; ++j) {
      if (find_grid_coord_in_list(off, i, j)) continue;
      std::vector<std::vector<std::size_t>> pairs;
      if (i > 0 && !find_grid_coord_in_list(off, i - 1, j)) pairs.push_back({i - 1, j});
      if (j > 0 && !find_grid_coord_in_list(off, i, j - 1)) pairs.push_back({i, j - 1});
      if (i < I - 1 && !find_grid_coord_in_list(off, i + 1, j)) pairs.push_back({i + 1, j});
      if (j < J - 1 && !find_grid_coord_in_list(off, i, j + 1)) pairs.push_back({i, j + 1});
      std::vector<std::string> order;
      if (find_grid_coord_in_list(final_qubit_region, i, j)) order.push_back(index_name({i, j}, {}));
      std::vector<std::size_t> local = {i, j};
      // small insertion sort: the comparator is only consistent on the
      // neighbours of one site
      auto less = order_func(ordering, local);
      for (std::size_t a = 1; a < pairs.size(); ++a) {
        auto x = pairs[a];
        std::size_t b = a;
        while (b > 0 && less(x, pairs[b - 1])) {
          pairs[b] = pairs[b - 1];
          --b;
        }
        pairs[b] = x;
      }
      for (const auto& pair : pairs) {
        if (pair[0] < i || pair[1] < j)
          order.push_back(index_name(pair, local));
        else
          order.push_back(index_name(local, pair));
      }
      try {
        grid_of_tensors_2D[i][j].reorder(order, scratch);
      } catch (const std::string& err_msg) {
        throw ERROR_MSG("Failed to call reorder(). Error:\n\t[", err_msg, "]");
      }
    }
  }
}

}  // namespace qflex
