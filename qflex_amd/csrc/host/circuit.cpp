// Circuit text parser.  The accepted text and the error messages are the
// reference's contract (src/circuit.cpp:30-206, docs/input_formats.md); the
// scanner below is a single pass over each line instead of a chain of
// regular-expression rewrites.
#include "circuit.h"

#include <cerrno>
#include <cstdlib>
#include <fstream>
#include <map>
#include <set>

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

bool is_space(char c) {
  return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}

bool is_kept(char c) {
  return is_space(c) || c == '(' || c == ')' || c == '_' || c == '.' || c == ',' || c == '-' ||
         (c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9');
}

// Canonical form of one line of circuit or ordering text:
//  * everything from '#' on is a comment;
//  * characters outside [A-Za-z0-9_.,-()] and white space are dropped;
//  * a run of two or more white-space characters becomes one space (a lone
//    tab becomes a space), trailing white space goes;
//  * with `glue_parens`, white space right before '(' goes;
//  * white space whose next parenthesis is ')' (inside "( ... )") goes.
// So "  2\tfsim( 0.25 , 0.1 )  0 1  # c" reads "2 fsim(0.25,0.1) 0 1" (the
// leading space survives as one space; tokens ignore it).
std::string canonical_line(const std::string& raw, bool glue_parens) {
  std::string s;
  s.reserve(raw.size());
  for (char c : raw) {
    if (c == '#') break;
    if (is_kept(c)) s.push_back(c == '\t' ? ' ' : c);
  }
  // collapse white-space runs
  std::string t;
  t.reserve(s.size());
  for (std::size_t i = 0; i < s.size();) {
    if (!is_space(s[i])) {
      t.push_back(s[i++]);
      continue;
    }
    std::size_t j = i;
    while (j < s.size() && is_space(s[j])) ++j;
    t.push_back(j - i >= 2 ? ' ' : s[i]);
    i = j;
  }
  while (!t.empty() && is_space(t.back())) t.pop_back();
  // drop white space before '(' and white space inside parentheses: scan
  // from the right remembering the next parenthesis seen
  std::string out(t.size(), '\0');
  std::size_t w = t.size();
  char next_paren = 0;
  for (std::size_t i = t.size(); i-- > 0;) {
    const char c = t[i];
    // (white-space runs are single characters here)
    if (is_space(c) && ((glue_parens && i + 1 < t.size() && t[i + 1] == '(') ||
                        next_paren == ')'))
      continue;
    if (c == '(' || c == ')') next_paren = c;
    out[--w] = c;
  }
  return out.substr(w);
}

std::vector<std::string> split_words(const std::string& s, char sep_if_any = 0) {
  std::vector<std::string> words;
  std::string cur;
  for (char c : s) {
    const bool sep = sep_if_any ? c == sep_if_any : is_space(c);
    if (sep) {
      if (!cur.empty()) words.push_back(cur);
      cur.clear();
    } else {
      cur.push_back(c);
    }
  }
  if (!cur.empty()) words.push_back(cur);
  return words;
}

// A token "is a number" when it starts with a decimal integer (an optional
// sign, then digits) that fits a long; "is an integer" when that integer
// also equals the token read as a double ("3", "3.0", "3x"; not "3.5").
bool leading_long(const std::string& tok, long* value) {
  const char* s = tok.c_str();
  char* end = nullptr;
  errno = 0;
  const long v = std::strtol(s, &end, 10);
  if (end == s || errno == ERANGE) return false;
  if (value) *value = v;
  return true;
}

bool integer_token(const std::string& tok, long* value) {
  long v = 0;
  if (!leading_long(tok, &v)) return false;
  if (double(v) != std::strtod(tok.c_str(), nullptr)) return false;
  if (value) *value = v;
  return true;
}

}  // namespace

std::string strip_ordering_line(const std::string& line) { return canonical_line(line, false); }

std::string canonical_circuit_line(const std::string& line) { return canonical_line(line, true); }

void QflexCircuit::load(std::istream&& istream) {
  clear();
  std::string line;
  std::size_t line_no = 0;  // physical lines read so far, comments included
  auto fail = [&](const std::string& msg) {
    return ERROR_MSG("[", std::to_string(line_no + 1), ": ", line, "] ", msg);
  };
  std::size_t cycle_now = 0;
  std::set<std::size_t> busy;  // qubits used in cycle_now
  for (; std::getline(istream, line); ++line_no) {
    line = canonical_line(line, true);
    if (line.empty()) continue;
    std::size_t opens = 0, closes = 0;
    for (char c : line) opens += c == '(', closes += c == ')';
    if (opens != closes || opens > 1) throw fail("Not-matching parentheses.");
    const std::vector<std::string> words = split_words(line);
    if (line_no == 0) {  // the first physical line holds the qubit count
      long n = 0;
      if (words.size() != 1 || !leading_long(words[0], &n) || n <= 0)
        throw fail("First line in circuit must be the number of active qubits.");
      num_active_qubits = std::size_t(n);
      continue;
    }
    if (words.size() < 3)
      throw fail("Gate must be specified as: cycle gate_name[(p1[,p2,...])] q1 [q2, ...]");
    QflexGate gate;
    gate.raw = line;
    long cycle = 0;
    if (!integer_token(words[0], &cycle)) throw fail("First token must be a valid cycle number.");
    gate.cycle = std::size_t(cycle);
    if (gate.cycle < cycle_now) throw fail("Cycle number can only increase.");
    if (gate.cycle != cycle_now) {
      busy.clear();
      cycle_now = gate.cycle;
    }
    const std::string& head = words[1];
    const std::size_t open = head.find('(');
    gate.name = head.substr(0, open);
    if (open != std::string::npos) {
      const std::size_t close = head.find(')');
      for (const auto& p : split_words(head.substr(open + 1, close - open - 1), ',')) {
        if (!leading_long(p, nullptr)) throw ERROR_MSG("Params must be valid numbers.");
        gate.params.push_back(std::strtod(p.c_str(), nullptr));
      }
    }
    for (std::size_t w = 2; w < words.size(); ++w) {
      long q = 0;
      if (!integer_token(words[w], &q)) throw fail("Qubit must be a valid number.");
      gate.qubits.push_back(std::size_t(q));
    }
    gates.push_back(gate);
    for (auto q : gates.back().qubits)
      if (!busy.insert(q).second) throw fail("Qubits can only used once per cycle.");
  }
  // Depth = the largest number of two-qubit gates acting on one qubit pair
  // (an fsim counts twice: its virtual bond has dimension 4).
  std::map<std::pair<std::size_t, std::size_t>, std::size_t> per_pair;
  for (const auto& gate : gates) {
    if (gate.qubits.size() > 2)
      throw fail("Depth calculation does not handle k-qubit gates with k > 2.");
    if (gate.qubits.size() != 2) continue;
    const std::size_t a = gate.qubits[0], b = gate.qubits[1];
    std::size_t& d = per_pair[a < b ? std::make_pair(a, b) : std::make_pair(b, a)];
    d += gate.name == "fsim" ? 2 : 1;
    if (d > depth) depth = d;
  }
}

void QflexCircuit::load(const std::string& filename) {
  std::ifstream in(filename);
  if (!in.good()) throw ERROR_MSG("Cannot open circuit file ", filename, ".");
  load(in);
}

}  // namespace qflex
