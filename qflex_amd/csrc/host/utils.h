// String helpers and the error-message convention of the reference
// (src/errors.h:7-12, src/utils.h:14-36).
#ifndef QFLEX_AMD_UTILS_H_
#define QFLEX_AMD_UTILS_H_

#include <sstream>
#include <string>

namespace qflex::utils {

template <typename T>
std::string concat(const T& x) {
  std::stringstream ss;
  ss << x;
  return ss.str();
}

template <typename T, typename... Q>
std::string concat(const T& x, const Q&... y) {
  std::stringstream ss;
  ss << x;
  return ss.str() + concat(y...);
}

// "Thu Aug 23 14:55:02 2001"
std::string get_date_time();

// "1GB" / "512 MB" / "1024" -> bytes (suffixes GB, MB, KB, B; 1024-based).
std::size_t from_readable_memory_string(std::string memory);

// bytes -> "6.26 GB" (two decimals, 1024-based).
std::string readable_memory_string(double memory);

}  // namespace qflex::utils

#endif
