// Errors are thrown as std::string "ERROR (file: function:line) --> ...",
// nested failures wrap as "Failed to call X(). Error:\n\t[...]"; callers and
// tests match substrings (reference src/errors.h:7-12).
#ifndef QFLEX_AMD_ERRORS_H_
#define QFLEX_AMD_ERRORS_H_

#include "global.h"
#include "utils.h"

#define ERROR_MSG(...)                                                       \
  qflex::utils::concat("ERROR (" __FILE__ ": ", __FUNCTION__, ":", __LINE__, \
                       ") --> ", __VA_ARGS__)

#define WARN_MSG(...) \
  qflex::utils::concat("[", qflex::utils::get_date_time(), "] ", __VA_ARGS__)

#endif
