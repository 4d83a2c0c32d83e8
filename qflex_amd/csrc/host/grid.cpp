// Grid parser; behaviour and error texts follow reference src/grid.cpp:11-45.
#include "grid.h"

#include <algorithm>
#include <fstream>

#include "errors.h"

namespace qflex {

void QflexGrid::clear() {
  I = 0;
  J = 0;
  qubits_off.clear();
}

void QflexGrid::load(std::istream& istream) { load(std::move(istream)); }

void QflexGrid::load(std::istream&& istream) {
  clear();
  std::string line;
  std::size_t line_counter = 0;
  while (std::getline(istream, line)) {
    if (line.empty() || line[0] == '#') continue;
    ++line_counter;
    line.erase(std::remove_if(line.begin(), line.end(), [](char x) { return x != '0' && x != '1'; }),
               line.end());
    if (line.empty()) continue;
    if (J != 0 && J != line.size())
      throw ERROR_MSG("Grid size at line ", line_counter, " is inconsistent with a width of ",
                      line.size(), " instead of ", J, ".");
    J = line.size();
    for (std::size_t q = 0; q < J; ++q)
      if (line[q] == '0') qubits_off.push_back({I, q});
    ++I;
  }
}

void QflexGrid::load(const std::string& filename) {
  if (auto in = std::ifstream(filename); in.good())
    load(in);
  else
    throw ERROR_MSG("Cannot open grid file: ", filename, ".");
}

}  // namespace qflex
