// Ordering text (reference src/ordering.h:13-19): one instruction per line
// (expand / cut / merge), comments stripped; parsed into operations by
// ordering_data_to_contraction_ordering (contraction_utils.h).
#ifndef QFLEX_AMD_ORDERING_H_
#define QFLEX_AMD_ORDERING_H_

#include <iostream>
#include <string>
#include <vector>

namespace qflex {

struct QflexOrdering {
  std::vector<std::string> instructions;
  void clear();
  void load(std::istream& istream);
  void load(std::istream&& istream);
  void load(const std::string& filename);
};

}  // namespace qflex

#endif
