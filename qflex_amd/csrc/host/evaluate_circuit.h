// Circuit evaluation entry (interface of reference src/evaluate_circuit.h:
// get_output_states :63-66, EvaluateCircuit :81-82).
#ifndef QFLEX_AMD_EVALUATE_CIRCUIT_H_
#define QFLEX_AMD_EVALUATE_CIRCUIT_H_

#include <complex>
#include <list>
#include <string>
#include <utility>
#include <vector>

#include "contraction_utils.h"
#include "input.h"
#include "read_circuit.h"
#include "tensor.h"

namespace qflex {

std::size_t find_output_pos(const QflexInput* input, std::vector<std::size_t> tensor_pos);

// Output bitstrings of an ordering's terminal cuts: the base final state with
// terminal-cut positions enumerated in cut order, first cut slowest.
std::string get_output_states(const QflexInput* input,
                              const std::list<ContractionOperation>& ordering,
                              std::vector<std::vector<std::size_t>>* final_qubits,
                              std::vector<std::string>* output_states);

// Parses the ordering, builds and flattens the network on the device,
// contracts it and returns (bitstring, amplitude) pairs.
std::vector<std::pair<std::string, std::complex<double>>> EvaluateCircuit(QflexInput* input);

// Amplitudes of several final states of one circuit (no counterpart in the
// reference, whose simulate() takes one final state,
// src/pybind_main.cpp:112-113).  The network is built and flattened once,
// with an open output index on every qubit whose bit differs between the
// states (besides the ordering's terminal cuts); each state then contracts
// zero-copy projections of those sites' tensors onto its bits.  Returns, per
// state in order, the (bitstring, amplitude) pairs EvaluateCircuit would
// return for it.
std::vector<std::pair<std::string, std::complex<double>>> EvaluateCircuitStates(
    QflexInput* input, const std::vector<std::string>& final_states);

// The same network build, then only the symbolic pass of the contraction
// (pairwise-contraction shapes and multiplicities, patch sizes): the
// features of the learned cost model.
SymbolicPlan PlanCircuit(QflexInput* input);

}  // namespace qflex

#endif
