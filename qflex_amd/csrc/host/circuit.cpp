// Circuit parser; accepted format and error texts follow reference
// src/circuit.cpp:30-206.
#include "circuit.h"

#include <algorithm>
#include <fstream>
#include <regex>
#include <unordered_map>
#include <unordered_set>

#include "errors.h"

namespace qflex {

std::ostream& QflexGate::operator<<(std::ostream& out) const {
  out << "gate_name: " << name << std::endl << "qubits: ";
  for (const auto& q : qubits) out << q << " ";
  out << std::endl;
  if (!params.empty()) {
    out << "params: ";
    for (const auto& p : params) out << p << " ";
    out << std::endl;
  }
  return out;
}

std::ostream& operator<<(std::ostream& out, const QflexGate& gate) { return gate.operator<<(out); }

void QflexCircuit::clear() {
  num_active_qubits = 0;
  depth = 0;
  gates.clear();
}

void QflexCircuit::load(std::istream& istream) { load(std::move(istream)); }

namespace {

bool is_number(const std::string& token) {
  try {
    std::stol(token);
  } catch (...) {
    return false;
  }
  return true;
}

bool is_integer(const std::string& token) {
  return is_number(token) && double(std::stol(token)) == std::stod(token);
}

// Normalises a line to "cycle gate(p1,p2) q1 q2".
std::string strip_line(std::string line, bool join_paren) {
  static const std::regex comment("#.*"), special("[^)(\\s\\ta-zA-Z0-9_.,-]"), tab("[\\t]"),
      spaces("[\\s]{2,}"), trailing("\\s+$"), before_paren("[\\s]+[(]"),
      inside_paren("\\s+(?=[^()]*\\))");
  line = std::regex_replace(line, comment, "");
  line = std::regex_replace(line, special, "");
  line = std::regex_replace(line, tab, " ");
  line = std::regex_replace(line, spaces, " ");
  line = std::regex_replace(line, trailing, "");
  if (join_paren) line = std::regex_replace(line, before_paren, "(");
  line = std::regex_replace(line, inside_paren, "");
  return line;
}

// Token patterns, compiled once: words of a line, and the parameters of
// "gate(p1,p2)".
const std::regex& word_pattern() {
  static const std::regex word("[^\\s]+");
  return word;
}
const std::regex& param_pattern() {
  static const std::regex param("[^,]+");
  return param;
}

std::vector<std::string> tokenize(const std::string& line, const std::regex& word = word_pattern()) {
  std::vector<std::string> tokens;
  for (auto w = std::sregex_iterator(line.begin(), line.end(), word); w != std::sregex_iterator();
       ++w)
    tokens.push_back(w->str());
  return tokens;
}

}  // namespace

std::string strip_ordering_line(const std::string& line) { return strip_line(line, false); }

void QflexCircuit::load(std::istream&& istream) {
  clear();
  std::size_t line_counter = 0, last_cycle_number = 0;
  std::unordered_set<std::size_t> used_qubits;
  std::string line;
  auto error_msg = [&line, &line_counter](const std::string& msg) {
    return ERROR_MSG("[", std::to_string(line_counter + 1), ": ", line, "] ", msg);
  };
  while (std::getline(istream, line)) {
    if (!(line = strip_line(line, true)).empty()) {
      const auto n_open = std::count(line.begin(), line.end(), '(');
      const auto n_close = std::count(line.begin(), line.end(), ')');
      if (n_open != n_close || n_open > 1) throw error_msg("Not-matching parentheses.");
      auto tokens = tokenize(line);
      if (line_counter == 0) {
        if (tokens.size() != 1 || !is_number(tokens[0]) || std::stol(tokens[0]) <= 0)
          throw error_msg("First line in circuit must be the number of active qubits.");
        num_active_qubits = std::stol(tokens[0]);
      } else {
        if (tokens.size() < 3)
          throw error_msg(
              "Gate must be specified as: cycle gate_name[(p1[,p2,...])] q1 [q2, ...]");
        gates.emplace_back();
        auto& gate = gates.back();
        gate.raw = line;
        if (!is_integer(tokens[0])) throw error_msg("First token must be a valid cycle number.");
        gate.cycle = std::stol(tokens[0]);
        if (gate.cycle < last_cycle_number) throw error_msg("Cycle number can only increase.");
        if (gate.cycle != last_cycle_number) {
          used_qubits.clear();
          last_cycle_number = gate.cycle;
        }
        if (auto beg = tokens[1].find_first_of('('); beg != std::string::npos) {
          gate.name = tokens[1].substr(0, beg);
          std::string params = tokens[1].substr(beg + 1, tokens[1].find_first_of(')') - beg - 1);
          for (const auto& p : tokenize(params, param_pattern())) {
            if (!is_number(p)) throw ERROR_MSG("Params must be valid numbers.");
            gate.params.push_back(std::stod(p));
          }
        } else {
          gate.name = tokens[1];
        }
        for (std::size_t i = 2; i < tokens.size(); ++i) {
          if (!is_integer(tokens[i])) throw error_msg("Qubit must be a valid number.");
          gate.qubits.push_back(std::stol(tokens[i]));
        }
        for (const auto& q : gate.qubits)
          if (!used_qubits.insert(q).second)
            throw error_msg("Qubits can only used once per cycle.");
      }
    }
    ++line_counter;
  }
  // Depth: largest number of two-qubit gates on one pair (fsim counts 2).
  std::unordered_map<std::size_t, std::unordered_map<std::size_t, std::size_t>> layers;
  for (const auto& gate : gates) {
    if (gate.qubits.size() > 2)
      throw error_msg("Depth calculation does not handle k-qubit gates with k > 2.");
    if (gate.qubits.size() == 2) {
      auto q1 = gate.qubits[0], q2 = gate.qubits[1];
      if (q2 > q1) std::swap(q1, q2);
      layers[q1][q2] += (gate.name == "fsim") ? 2 : 1;
      depth = std::max(depth, layers[q1][q2]);
    }
  }
}

void QflexCircuit::load(const std::string& filename) {
  if (auto in = std::ifstream(filename); in.good())
    load(in);
  else
    throw ERROR_MSG("Cannot open circuit file ", filename, ".");
}

}  // namespace qflex
