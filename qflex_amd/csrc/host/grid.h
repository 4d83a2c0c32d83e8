// Grid layout format (reference src/grid.h:13-20): rows of 0/1 (0 = qubit
// off); anything else on a line is ignored; '#' lines are comments.
#ifndef QFLEX_AMD_GRID_H_
#define QFLEX_AMD_GRID_H_

#include <iostream>
#include <string>
#include <vector>

namespace qflex {

struct QflexGrid {
  std::size_t I{0}, J{0};
  std::vector<std::vector<std::size_t>> qubits_off;
  void clear();
  void load(std::istream& istream);
  void load(std::istream&& istream);
  void load(const std::string& filename);
};

}  // namespace qflex

#endif
