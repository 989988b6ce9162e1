// caviar2/writer.hpp -- output plugin base (reference inc/caviar2/writer.hpp).
#pragma once
#include <cstdint>

namespace caviar2 {

class Caviar2;

class Writer {
 public:
  Writer() = default;
  explicit Writer(Caviar2 *c) : caviar_(c) {}
  virtual ~Writer() = default;
  virtual void write(int64_t step, double time) = 0;
  // Smallest step >= `from` at which write() produces output (-1: never);
  // the simulator brings the host mirror up to date only at those steps.
  // The default (every step) suits writers that do not report a cadence.
  virtual int64_t next_output_step(int64_t from) const { return from; }

 protected:
  Caviar2 *caviar_ = nullptr;
};

}  // namespace caviar2
