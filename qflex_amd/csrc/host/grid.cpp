// Grid layout parser.  The accepted text and the error message are the
// reference's contract (src/grid.cpp:11-45, docs/input_formats.md): each
// non-comment line is one row of 0/1 flags (0 = qubit off), every other
// character on it is ignored, all rows have the same width.
#include "grid.h"

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
  std::string text;
  std::size_t counted = 0;  // non-empty, non-comment lines seen (numbers the error)
  while (std::getline(istream, text)) {
    if (text.empty() || text.front() == '#') continue;
    ++counted;
    std::vector<std::size_t> off_columns;
    std::size_t width = 0;
    for (char c : text) {
      if (c != '0' && c != '1') continue;
      if (c == '0') off_columns.push_back(width);
      ++width;
    }
    if (width == 0) continue;  // a line without flags is no row
    if (J == 0) J = width;
    if (width != J)
      throw ERROR_MSG("Grid size at line ", counted, " is inconsistent with a width of ", width,
                      " instead of ", J, ".");
    for (std::size_t col : off_columns) qubits_off.push_back({I, col});
    ++I;
  }
}

void QflexGrid::load(const std::string& filename) {
  std::ifstream in(filename);
  if (!in.good()) throw ERROR_MSG("Cannot open grid file: ", filename, ".");
  load(in);
}

}  // namespace qflex
