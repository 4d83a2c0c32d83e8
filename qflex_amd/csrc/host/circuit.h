// Circuit text format (reference src/circuit.h:14-35, src/circuit.cpp):
//   line 1: number of active qubits
//   then:   cycle gate_name[(p1,p2,...)] q1 [q2]
// '#' starts a comment; each qubit at most once per cycle; cycles never
// decrease.
#ifndef QFLEX_AMD_CIRCUIT_H_
#define QFLEX_AMD_CIRCUIT_H_

#include <iostream>
#include <string>
#include <vector>

namespace qflex {

struct QflexGate {
  std::string name;
  std::size_t cycle;
  std::vector<std::size_t> qubits;
  std::vector<double> params;
  std::string raw;

  std::ostream& operator<<(std::ostream& out) const;
  friend std::ostream& operator<<(std::ostream& out, const QflexGate& gate);
};

struct QflexCircuit {
  std::size_t num_active_qubits{0};
  std::size_t depth{0};
  std::vector<QflexGate> gates;

  void clear();
  void load(std::istream& istream);
  void load(std::istream&& istream);
  void load(const std::string& filename);
};

// Canonical form of one line of circuit text: what the parser splits into
// words and what its error messages quote.
std::string canonical_circuit_line(const std::string& line);
// The same for ordering text (no gluing of "name (" into "name(").
std::string strip_ordering_line(const std::string& line);

}  // namespace qflex

#endif
