// Host build of the scalar complex API probe (tests/test_complex_api.py).
#include <cstdint>

#include "veig_complex.hpp"
#include "complex_probe.inc"

extern "C" void probe_host(int op, int64_t n, const double *a, double *out) {
  for (int64_t k = 0; k < n; ++k) probe_one(op, a + 4 * k, out + 2 * k);
}

// scale_rc on host planes (tests/test_complex_api.py)
extern "C" void probe_scale_rc_host(double *rV, double *iV, double zr, double zi, int I, int I1, int I2, int dim,
                                    int n) {
  scale_rc(rV, iV, zr, zi, I, I1, I2, dim, n);
}
