// utils:: helpers (reference src/utils.cpp:10-68) and the device context.
#include <algorithm>
#include <cctype>
#include <cmath>
#include <ctime>
#include <regex>

#include "device.h"
#include "errors.h"

namespace qflex::utils {

std::string get_date_time() {
  std::time_t t = std::time(nullptr);
  char buf[100];
  if (std::strftime(buf, sizeof(buf), "%c", std::localtime(&t))) return buf;
  return "[time-readout error]";
}

std::size_t from_readable_memory_string(std::string memory) {
  memory = std::regex_replace(memory, std::regex("[^)(\\s\\ta-zA-Z0-9_.,-]"), "");
  memory.erase(std::remove_if(memory.begin(), memory.end(),
                              [](unsigned char c) { return std::isspace(c); }),
               memory.end());
  for (auto& c : memory) c = char(std::toupper(static_cast<unsigned char>(c)));
  const std::size_t digits_end = std::min(memory.find_first_not_of("0123456789"), memory.size());
  if (digits_end == 0) throw ERROR_MSG("Memory string '", memory, "' has no leading number.");
  std::size_t m = std::stoull(memory.substr(0, digits_end));
  const std::string suffix = memory.substr(digits_end);
  if (suffix == "GB")
    m <<= 30;
  else if (suffix == "MB")
    m <<= 20;
  else if (suffix == "KB")
    m <<= 10;
  else if (suffix != "B" && !suffix.empty())
    throw ERROR_MSG("Suffix ", suffix, " not recognized.");
  return m;
}

std::string readable_memory_string(double memory) {
  const char* suffix[] = {" B", " kB", " MB", " GB"};
  std::size_t scale = 0;
  while (memory >= 1024 && scale < 3) {
    ++scale;
    memory /= 1024;
  }
  return concat(static_cast<std::size_t>(memory * 100) / 100., suffix[scale]);
}

}  // namespace qflex::utils

namespace qflex {

qfx_context* device() {
  static qfx_context* ctx = [] {
    int dev = 0;
    if (const char* e = std::getenv("QFX_DEVICE")) dev = std::atoi(e);
    qfx_context* c = nullptr;
    int st = qfx_context_create(dev, 0, &c);
    if (st != QFX_OK)
      throw ERROR_MSG("Cannot create the device context: ", qfx_last_error());
    return c;
  }();
  return ctx;
}

void qfx_check(int status, const char* what) {
  if (status != QFX_OK) throw ERROR_MSG(what, " failed: ", qfx_last_error());
}

}  // namespace qflex
