// Ordering loader (reference src/ordering.cpp:10-45).
#include "ordering.h"

#include <fstream>

#include "circuit.h"

namespace qflex {

void QflexOrdering::clear() { instructions.clear(); }

void QflexOrdering::load(std::istream& istream) { load(std::move(istream)); }

void QflexOrdering::load(std::istream&& istream) {
  clear();
  std::string line;
  while (std::getline(istream, line))
    if (!(line = strip_ordering_line(line)).empty()) instructions.push_back(line);
}

void QflexOrdering::load(const std::string& filename) {
  if (auto in = std::ifstream(filename); in.good())
    load(in);
  else
    throw std::string("Cannot open ordering file: ") + filename;
}

}  // namespace qflex
