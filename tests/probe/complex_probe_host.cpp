// Host build of the scalar complex API probe (tests/test_complex_api.py).
#include <cstdint>

#include "veig_complex.hpp"
#include "complex_probe.inc"

extern "C" void probe_host(int op, int64_t n, const double *a, double *out) {
  for (int64_t k = 0; k < n; ++k) probe_one(op, a + 4 * k, out + 2 * k);
}
