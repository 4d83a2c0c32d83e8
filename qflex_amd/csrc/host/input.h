// Everything EvaluateCircuit needs (reference src/input.h:11-17).
#ifndef QFLEX_AMD_INPUT_H_
#define QFLEX_AMD_INPUT_H_

#include <string>

#include "circuit.h"
#include "grid.h"
#include "ordering.h"

namespace qflex {

struct QflexInput {
  QflexOrdering ordering;
  QflexCircuit circuit;
  QflexGrid grid;
  std::string initial_state;
  std::string final_state;
};

}  // namespace qflex

#endif
