// Host build of the engine's elementary functions (qdeqcycle_fdm_amd/csrc/
// qd_math.h, the code the fault kernels run): reads FP64 bit patterns (hex,
// one per line) on stdin and prints, per input, the bit patterns of
// qd_exp, qd_sinh, qd_cosh, qd_log, qd_log1p, qd_asinh and the two results
// of qd_sinh_cosh.  tests/test_qd_math_host.py builds it with g++
// -ffp-contract=off (the flag the library and the oracle are built with).
#include <cinttypes>
#include <cstdio>
#include <cstring>

#include "qd_math.h"

static uint64_t bits(double v) {
  uint64_t b;
  std::memcpy(&b, &v, sizeof b);
  return b;
}

int main() {
  uint64_t in;
  while (std::scanf("%" SCNx64, &in) == 1) {
    double x;
    std::memcpy(&x, &in, sizeof x);
    double s, c;
    qdeq::qd_sinh_cosh(x, &s, &c);
    std::printf("%016" PRIx64 " %016" PRIx64 " %016" PRIx64 " %016" PRIx64
                " %016" PRIx64 " %016" PRIx64 " %016" PRIx64 " %016" PRIx64 "\n",
                bits(qdeq::qd_exp(x)), bits(qdeq::qd_sinh(x)), bits(qdeq::qd_cosh(x)),
                bits(qdeq::qd_log(x)), bits(qdeq::qd_log1p(x)), bits(qdeq::qd_asinh(x)),
                bits(s), bits(c));
  }
  return 0;
}
