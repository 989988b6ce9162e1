// caviar2/log.hpp -- messages and the fatal-error convention of the caviar2
// API (reference src/log.cpp:190-194,337-365: "[ERR]" then exit(1)).
#pragma once
#include <string>

namespace caviar2 {

class Caviar2;

class Log {
 public:
  void set_caviar(Caviar2 *c) { caviar_ = c; }
  void info(const std::string &s, int level = 0);
  void warning(const std::string &s);
  [[noreturn]] void error_all(const std::string &s);
  [[noreturn]] void error_all(const char *file, int line, const char *func, const std::string &s);
  void set_verbosity(int v) { verbosity_ = v; }

 private:
  Caviar2 *caviar_ = nullptr;
  int verbosity_ = 1;
};

}  // namespace caviar2

#define FC_FILE_LINE_FUNC __FILE__, __LINE__, __func__
