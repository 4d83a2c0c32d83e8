// Python module `qflex`.
//
// simulate(options: dict) -> [(bitstring, complex)] is the reference's entry
// (src/pybind_main.cpp:75-138, src/pybind_main.h:19-23): the same keys
// (grid/circuit/ordering as lists of lines or *_filename, initial_state(s),
// final_state(s), memory_limit), the same single-state restriction, and
// errors printed to stderr with [] returned.  Extra optional keys:
// "precision" ("fp32" | "split_fp16" | "split_bf16x3" | "split_bf16" |
// "fp32_exact"), "verbose" (int) and "autotune" (bool: time the GEMM tiles
// on first use and keep the fastest, for this and later calls) and
// "slice_batch" (int, or one int per bond cut in ordering order: the number
// of cut values contracted together, the cut index kept as a batch index;
// default 1, the reference's one slice at a time).  "final_states" may list
// several bitstrings (the reference accepts one): the network is then built
// and flattened once and contracted once per state.  "record_slices" (bool):
// the innermost bond cut's passes after the first replay one recorded graph.
//
// The remaining functions expose the host's Tensor / contraction API to the
// tests; they raise RuntimeError with the reference's message texts.
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <iostream>
#include <sstream>

#include "device.h"
#include "errors.h"
#include "evaluate_circuit.h"
#include "global.h"

namespace py = pybind11;

namespace {

// The lines of a Python list of strings, newline-terminated, as one stream.
std::stringstream lines_of(const py::handle& seq, const std::string& key) {
  std::stringstream text;
  for (const auto& item : seq.cast<py::iterable>()) {
    if (!py::isinstance<py::str>(item)) throw ERROR_MSG(key, " must be a list of strings.");
    text << item.cast<std::string>() << '\n';
  }
  return text;
}

// Input `key` ("grid", "circuit" or "ordering") of a simulate() dict: the
// text itself as a list of lines under `key`, else a file path under
// "<key>_filename"; with neither, `target` stays empty.  A bare string under
// `key` is refused (it would otherwise be read one character per line).
template <typename Target>
void read_input(const py::dict& options, const std::string& key, Target& target) {
  if (options.contains(key)) {
    const py::handle value = options[key.c_str()];
    if (py::isinstance<py::str>(value) || !py::isinstance<py::iterable>(value))
      throw ERROR_MSG("'", key, "' must be a list of strings.");
    target.load(lines_of(value, key));
    return;
  }
  const std::string file_key = key + "_filename";
  if (!options.contains(file_key)) return;
  const py::handle path = options[file_key.c_str()];
  if (!py::isinstance<py::str>(path)) throw ERROR_MSG("'", file_key, "' must be a string.");
  target.load(path.cast<std::string>());
}

// Bitstrings of one side ("initial" or "final"): "<side>_states" holds one
// string or a list of them, "<side>_state" one string; neither gives the
// empty string (no fixed bits).
std::vector<std::string> read_states(const py::dict& options, const std::string& side) {
  auto one_string = [](const py::handle& h) {
    if (!py::isinstance<py::str>(h)) throw ERROR_MSG("states must be strings.");
    return h.cast<std::string>();
  };
  const std::string plural = side + "_states", singular = side + "_state";
  if (options.contains(plural)) {
    const py::handle value = options[plural.c_str()];
    if (py::isinstance<py::str>(value)) return {value.cast<std::string>()};
    if (!py::isinstance<py::iterable>(value))
      throw ERROR_MSG(plural, " must be a list of strings.");
    std::vector<std::string> states;
    for (const auto& item : value.cast<py::iterable>()) states.push_back(one_string(item));
    return states;
  }
  if (options.contains(singular)) return {one_string(options[singular.c_str()])};
  return {""};
}

int parse_precision(const std::string& p) {
  if (p == "fp32") return QFX_PREC_FP32;
  if (p == "split_bf16") return QFX_PREC_SPLIT_BF16;
  if (p == "split_bf16x3") return QFX_PREC_SPLIT_BF16X3;
  if (p == "fp32_exact") return QFX_PREC_FP32_EXACT;
  if (p == "split_fp16") return QFX_PREC_SPLIT_FP16;
  // The FP8 precision of the C-ABI (QFX_PREC_FP8) is not offered to whole
  // contractions: its first run over the 7x7, depth 1+32+1 workload ended in
  // a GPU memory fault whose cause is not found; the engine refuses it.
  if (p == "fp8")
    throw ERROR_MSG("Precision 'fp8' is disabled in the engine (the C-ABI's FP8 GEMM is "
                    "available through qfx_cgemm_c64 / qfx_cgemm_mapped_c64).");
  throw ERROR_MSG("Unknown precision '", p,
                  "' (expected fp32, split_fp16, split_bf16x3, split_bf16 or fp32_exact).");
}

// Fills a QflexInput (and the process-wide settings) from a simulate() dict.
void load_input(const py::dict& options, qflex::QflexInput& input,
                std::vector<std::string>* final_states = nullptr) {
  if (options.contains("memory_limit")) {
    const auto& ml = options["memory_limit"];
    if (py::isinstance<py::int_>(ml)) {
      const long v = ml.cast<long>();
      if (v <= 0) throw ERROR_MSG("'memory_limit' must be positive.");
      qflex::global::memory_limit = std::size_t(v);
    } else if (py::isinstance<py::str>(ml)) {
      qflex::global::memory_limit =
          qflex::utils::from_readable_memory_string(ml.cast<std::string>());
    } else {
      throw ERROR_MSG("'memory_limit' must be valid.");
    }
  } else {
    qflex::global::memory_limit = std::size_t(1) << 30;
  }
  qflex::global::precision = QFX_PREC_FP32;
  if (options.contains("precision"))
    qflex::global::precision = parse_precision(options["precision"].cast<std::string>());
  if (options.contains("verbose")) qflex::global::verbose = options["verbose"].cast<int>();
  qflex::global::record_slices =
      options.contains("record_slices") && options["record_slices"].cast<bool>();
  qflex::global::slice_batch.clear();
  if (options.contains("slice_batch")) {
    const auto& sb = options["slice_batch"];
    std::vector<long> sizes;
    if (py::isinstance<py::int_>(sb))
      sizes.push_back(sb.cast<long>());
    else
      sizes = sb.cast<std::vector<long>>();
    for (long v : sizes) {
      if (v < 1) throw ERROR_MSG("'slice_batch' entries must be positive.");
      qflex::global::slice_batch.push_back(std::size_t(v));
    }
  }
  if (options.contains("autotune"))
    qflex::qfx_check(qfx_autotune_enable(qflex::device(), options["autotune"].cast<bool>() ? 1 : 0),
                     "qfx_autotune_enable");

  read_input(options, "grid", input.grid);
  read_input(options, "circuit", input.circuit);
  read_input(options, "ordering", input.ordering);
  const auto initial_states = read_states(options, "initial");
  const auto finals = read_states(options, "final");
  // One initial state, as the reference; several final states are an
  // extension (EvaluateCircuitStates: one network build for all of them).
  if (initial_states.size() != 1 || finals.empty()) throw ERROR_MSG("Not yet supported");
  input.initial_state = initial_states[0];
  input.final_state = finals[0];
  if (final_states) *final_states = finals;
}

std::vector<std::pair<std::string, std::complex<double>>> evaluate(const py::dict& options) {
  qflex::QflexInput input;
  std::vector<std::string> finals;
  load_input(options, input, &finals);
  if (finals.size() > 1) return qflex::EvaluateCircuitStates(&input, finals);
  return qflex::EvaluateCircuit(&input);
}

// Symbolic plan of a simulate() dict: one dict per pairwise contraction
// (op, m, n, k, multiplicity, out_rank, out_size) plus the patch sizes.
py::dict plan(const py::dict& options) {
  qflex::QflexInput input;
  load_input(options, input);
  const qflex::SymbolicPlan p = qflex::PlanCircuit(&input);
  py::list steps;
  for (const auto& st : p.steps) {
    py::dict d;
    d["op"] = st.op == qflex::ContractionOperation::EXPAND ? "expand" : "merge";
    d["m"] = st.m;
    d["n"] = st.n;
    d["k"] = st.k;
    d["multiplicity"] = st.multiplicity;
    d["batch"] = st.batch;
    d["out_rank"] = st.out_rank;
    d["out_size"] = st.out_size;
    steps.append(d);
  }
  py::dict patches;
  for (const auto& kv : p.patch_max_size) patches[py::str(kv.first)] = kv.second;
  py::dict out;
  out["steps"] = steps;
  out["patch_max_size"] = patches;
  out["max_grid"] = p.max_grid;
  return out;
}

std::vector<std::pair<std::string, std::complex<double>>> simulate(const py::dict& options) {
  try {
    return evaluate(options);
  } catch (std::string err_msg) {
    std::cerr << err_msg << std::endl;
  } catch (const char* err_msg) {
    std::cerr << err_msg << std::endl;
  } catch (const std::exception& e) {
    std::cerr << e.what() << std::endl;
  }
  return {};
}

// Runs f, turning the reference's string throws into RuntimeError.
template <typename F>
auto wrap(F&& f) -> decltype(f()) {
  try {
    return f();
  } catch (const std::string& msg) {
    throw std::runtime_error(msg);
  }
}

qflex::s_type* scratch_for(std::size_t n) {
  static qflex::s_type* buf = nullptr;
  static std::size_t cap = 0;
  if (n > cap) {
    if (buf) qflex::device_free(buf);
    buf = qflex::device_alloc(n);
    cap = n;
  }
  return buf;
}

std::vector<qflex::s_type> to_values(py::array_t<std::complex<float>, py::array::c_style |
                                                                        py::array::forcecast>
                                         a) {
  auto r = a.unchecked();
  std::vector<qflex::s_type> v(a.size());
  const std::complex<float>* p = a.data();
  std::copy(p, p + a.size(), v.begin());
  return v;
}

// ("expand", id, (i, j)) | ("cut", ((i, j), ...), [values]) | ("merge", src, tgt)
std::list<qflex::ContractionOperation> to_ops(const py::list& ops) {
  std::list<qflex::ContractionOperation> out;
  for (const auto& o : ops) {
    auto t = o.cast<py::tuple>();
    const std::string kind = t[0].cast<std::string>();
    if (kind == "expand") {
      out.emplace_back(qflex::ExpandPatch(t[1].cast<std::string>(),
                                          t[2].cast<std::vector<std::size_t>>()));
    } else if (kind == "cut") {
      std::vector<std::size_t> values;
      if (t.size() > 2) values = t[2].cast<std::vector<std::size_t>>();
      out.emplace_back(
          qflex::CutIndex(t[1].cast<std::vector<std::vector<std::size_t>>>(), values));
    } else if (kind == "merge") {
      out.emplace_back(qflex::MergePatches(t[1].cast<std::string>(), t[2].cast<std::string>()));
    } else {
      throw std::runtime_error("unknown operation " + kind);
    }
  }
  return out;
}

py::list from_ops(const std::list<qflex::ContractionOperation>& ops) {
  py::list out;
  for (const auto& op : ops) {
    if (op.op_type == qflex::ContractionOperation::EXPAND)
      out.append(py::make_tuple("expand", op.expand.id, op.expand.tensor));
    else if (op.op_type == qflex::ContractionOperation::CUT)
      out.append(py::make_tuple("cut", op.cut.tensors, op.cut.values));
    else
      out.append(py::make_tuple("merge", op.merge.source_id, op.merge.target_id));
  }
  return out;
}

}  // namespace

PYBIND11_MODULE(qflex, m) {
  m.doc() = "MI355X contraction engine for qFlex grid circuits";
  m.def("simulate", &simulate, "Call the simulator with the normal parameters");
  m.def("plan", [](const py::dict& o) { return wrap([&] { return plan(o); }); },
        "Symbolic plan (pairwise-contraction shapes) of a simulate() input");
  m.def("evaluate", [](const py::dict& o) { return wrap([&] { return evaluate(o); }); },
        "simulate() that raises RuntimeError instead of printing errors");

  py::class_<qflex::Tensor>(m, "Tensor")
      .def(py::init<>())
      .def(py::init([](std::vector<std::string> idx, std::vector<std::size_t> dims) {
             return wrap([&] { return new qflex::Tensor(idx, dims); });
           }))
      .def(py::init([](std::vector<std::string> idx, std::vector<std::size_t> dims,
                       py::array_t<std::complex<float>, py::array::c_style | py::array::forcecast>
                           data) {
             return wrap([&] { return new qflex::Tensor(idx, dims, to_values(data)); });
           }))
      .def("get_indices", &qflex::Tensor::get_indices)
      .def("get_dimensions", &qflex::Tensor::get_dimensions)
      .def("size", &qflex::Tensor::size)
      .def("capacity", &qflex::Tensor::capacity)
      .def("data",
           [](const qflex::Tensor& t) {
             auto v = wrap([&] { return t.host_data(); });
             return py::array_t<std::complex<float>>(v.size(), v.data());
           })
      .def("set_indices_and_dimensions",
           [](qflex::Tensor& t, std::vector<std::string> i, std::vector<std::size_t> d) {
             wrap([&] {
               t.set_indices_and_dimensions(i, d);
               return 0;
             });
           })
      .def("rename_index",
           [](qflex::Tensor& t, std::string a, std::string b) {
             wrap([&] {
               t.rename_index(a, b);
               return 0;
             });
           })
      .def("bundle",
           [](qflex::Tensor& t, std::vector<std::string> i, std::string n) {
             wrap([&] {
               t.bundle(i, n);
               return 0;
             });
           })
      .def("project",
           [](const qflex::Tensor& t, std::string i, std::size_t v, qflex::Tensor& out) {
             wrap([&] {
               t.project(i, v, out);
               return 0;
             });
           })
      .def("reorder",
           [](qflex::Tensor& t, std::vector<std::string> order, bool null_scratch) {
             wrap([&] {
               t.reorder(order, null_scratch ? nullptr : scratch_for(t.size()));
               return 0;
             });
           },
           py::arg("new_ordering"), py::arg("null_scratch") = false)
      .def("assign",
           [](qflex::Tensor& t, const qflex::Tensor& o) {
             wrap([&] {
               t = o;
               return 0;
             });
           })
      .def("tensor_norm", [](const qflex::Tensor& t) { return wrap([&] { return t.tensor_norm(); }); })
      .def("num_zeros", [](const qflex::Tensor& t) { return wrap([&] { return t.num_zeros(); }); })
      .def("tensor_to_string", &qflex::Tensor::tensor_to_string);

  m.def("multiply",
        [](qflex::Tensor& A, qflex::Tensor& B, qflex::Tensor& C, bool null_scratch) {
          wrap([&] {
            qflex::multiply(A, B, C,
                            null_scratch ? nullptr : scratch_for(std::max(A.size(), B.size())));
            return 0;
          });
        },
        py::arg("A"), py::arg("B"), py::arg("C"), py::arg("null_scratch") = false);
  m.def("bundle_between", [](qflex::Tensor& A, qflex::Tensor& B, std::string name) {
    wrap([&] {
      qflex::bundle_between(A, B, name, scratch_for(std::max(A.size(), B.size())));
      return 0;
    });
  });

  m.def("contract_grid",
        [](py::list ops, py::list grid, std::size_t n_amplitudes) {
          return wrap([&] {
            auto ordering = to_ops(ops);
            std::vector<std::vector<qflex::Tensor>> tg;
            for (const auto& row : grid) {
              tg.emplace_back();
              for (const auto& t : row.cast<py::list>()) {
                if (t.is_none())
                  tg.back().emplace_back();
                else
                  tg.back().push_back(t.cast<qflex::Tensor&>());
              }
            }
            std::vector<std::complex<double>> amps(n_amplitudes);
            auto data = qflex::ContractionData::Initialize(ordering, &tg, &amps);
            std::unordered_map<std::string, bool> active;
            for (const auto& p : data.scratch_list()) active[p] = false;
            data.ContractGrid(ordering, 0, active);
            return amps;
          });
        },
        "ContractionData::Initialize + ContractGrid on a grid of Tensors");
  m.def("validate_ordering", [](py::list ops) {
    wrap([&] {
      qflex::ValidateOrdering(to_ops(ops));
      return 0;
    });
  });
  m.def("parse_ordering",
        [](std::vector<std::string> ordering_lines, std::vector<std::string> grid_lines) {
          return wrap([&] {
            qflex::QflexInput input;
            std::stringstream o, g;
            for (auto& l : ordering_lines) o << l << "\n";
            for (auto& l : grid_lines) g << l << "\n";
            input.ordering.load(o);
            input.grid.load(g);
            std::list<qflex::ContractionOperation> ops;
            qflex::ordering_data_to_contraction_ordering(input, &ops);
            return from_ops(ops);
          });
        });
  m.def("get_output_states",
        [](std::size_t I, std::size_t J, std::vector<std::vector<std::size_t>> off,
           std::string initial, std::string final_state, py::list ops) {
          return wrap([&] {
            qflex::QflexInput input;
            input.grid.I = I;
            input.grid.J = J;
            input.grid.qubits_off = off;
            input.initial_state = initial;
            input.final_state = final_state;
            std::vector<std::vector<std::size_t>> fq;
            std::vector<std::string> states;
            auto base = qflex::get_output_states(&input, to_ops(ops), &fq, &states);
            return py::make_tuple(base, fq, states);
          });
        });
  m.def("load_circuit", [](std::vector<std::string> lines) {
    return wrap([&] {
      std::stringstream s;
      for (auto& l : lines) s << l << "\n";
      qflex::QflexCircuit c;
      c.load(s);
      py::list gates;
      for (const auto& g : c.gates) gates.append(py::make_tuple(g.cycle, g.name, g.qubits, g.params));
      return py::make_tuple(c.num_active_qubits, c.depth, gates);
    });
  });
  m.def("canonical_lines", [](std::vector<std::string> lines, bool circuit) {
    std::vector<std::string> out;
    for (const auto& l : lines)
      out.push_back(circuit ? qflex::canonical_circuit_line(l) : qflex::strip_ordering_line(l));
    return out;
  });
  m.def("load_grid", [](std::vector<std::string> lines) {
    return wrap([&] {
      std::stringstream s;
      for (auto& l : lines) s << l << "\n";
      qflex::QflexGrid g;
      g.load(s);
      return py::make_tuple(g.I, g.J, g.qubits_off);
    });
  });
  m.def("flatten_indices",
        [](std::vector<std::string> circuit_lines, std::size_t I, std::size_t J,
           std::string initial, std::string final_state,
           std::vector<std::vector<std::size_t>> final_region,
           std::vector<std::vector<std::size_t>> off, py::list ops) {
          return wrap([&] {
            std::stringstream s;
            for (auto& l : circuit_lines) s << l << "\n";
            qflex::QflexCircuit c;
            c.load(s);
            auto ordering = to_ops(ops);
            std::vector<std::vector<std::vector<qflex::Tensor>>> g3;
            qflex::s_type* scratch = scratch_for(1 << 16);
            qflex::circuit_data_to_tensor_network(c, I, J, initial, final_state, final_region, off,
                                                  g3, scratch);
            std::vector<std::vector<qflex::Tensor>> g2(I, std::vector<qflex::Tensor>(J));
            qflex::flatten_grid_of_tensors(g3, g2, final_region, off, ordering, scratch);
            std::vector<std::vector<std::vector<std::string>>> idx(I);
            for (std::size_t i = 0; i < I; ++i)
              for (std::size_t j = 0; j < J; ++j) idx[i].push_back(g2[i][j].get_indices());
            return idx;
          });
        });
  m.def("last_stats", [] {
    const auto& s = qflex::last_contraction_stats();
    py::dict d;
    d["slices"] = s.slices;
    d["multiplies"] = s.multiplies;
    d["flops"] = s.flops;
    d["reorders"] = s.reorders;
    d["reorder_elements"] = s.reorder_elements;
    d["mapped_multiplies"] = s.mapped_multiplies;
    d["replayed_passes"] = s.replayed_passes;
    d["seconds"] = s.seconds;
    d["planned_bytes"] = s.planned_bytes;
    return d;
  });
  m.def("set_mapped_operands", [](bool on) { qflex::global::mapped_operands = on; });
  m.def("set_precision", [](std::string p) {
    wrap([&] {
      qflex::global::precision = parse_precision(p);
      return 0;
    });
  });
  m.def("profile_enable", [](bool on) {
    wrap([&] {
      qflex::qfx_check(qfx_profile_enable(qflex::device(), on ? 1 : 0), "qfx_profile_enable");
      qflex::global::profiling = on;
      return 0;
    });
  });
  m.def("profile_reset", [] {
    wrap([&] {
      qflex::qfx_check(qfx_profile_reset(qflex::device()), "qfx_profile_reset");
      return 0;
    });
  });
  m.def("profile_read", [] {
    return wrap([&] {
      uint64_t la[QFX_K_NUM_CLASSES];
      double ms[QFX_K_NUM_CLASSES], by[QFX_K_NUM_CLASSES], fl[QFX_K_NUM_CLASSES];
      qflex::qfx_check(qfx_profile_read(qflex::device(), la, ms, by, fl), "qfx_profile_read");
      const char* names[] = {"permute", "cgemm",  "gemv",       "dot",
                             "accumulate", "copy", "cgemm_small"};
      py::dict d;
      for (int i = 0; i < QFX_K_NUM_CLASSES; ++i) {
        py::dict e;
        e["launches"] = la[i];
        e["ms"] = ms[i];
        e["bytes"] = by[i];
        e["flops"] = fl[i];
        d[names[i]] = e;
      }
      return d;
    });
  });
  m.def("synchronize", [] {
    wrap([&] {
      qflex::qfx_check(qfx_synchronize(qflex::device()), "qfx_synchronize");
      return 0;
    });
  });
  // Tile autotuner of the device context (see qfx_autotune_* in qfx.h).
  m.def("set_autotune", [](bool on) {
    wrap([&] {
      qflex::qfx_check(qfx_autotune_enable(qflex::device(), on ? 1 : 0), "qfx_autotune_enable");
      return 0;
    });
  });
  m.def("autotune_table", [] {
    return wrap([&] {
      int count = 0;
      qflex::qfx_check(qfx_autotune_entries(qflex::device(), nullptr, 0, &count),
                       "qfx_autotune_entries");
      std::vector<uint64_t> buf(6 * std::size_t(count));
      qflex::qfx_check(qfx_autotune_entries(qflex::device(), buf.data(), count, &count),
                       "qfx_autotune_entries");
      py::list out;
      for (int i = 0; i < count; ++i)
        out.append(py::make_tuple(buf[6 * i], buf[6 * i + 1], buf[6 * i + 2], buf[6 * i + 3],
                                  buf[6 * i + 4], buf[6 * i + 5]));
      return out;
    });
  });
  m.def("set_autotune_entry", [](int precision, int form, uint64_t m, uint64_t n, uint64_t k,
                                 int tile) {
    wrap([&] {
      qflex::qfx_check(qfx_autotune_set(qflex::device(), precision, form, m, n, k, tile),
                       "qfx_autotune_set");
      return 0;
    });
  });
  m.def("arena_stats", [] {
    return wrap([&] {
      size_t reserved = 0, in_use = 0;
      qflex::qfx_check(qfx_arena_stats(qflex::device(), &reserved, &in_use), "qfx_arena_stats");
      return py::make_tuple(reserved, in_use);
    });
  });
}
