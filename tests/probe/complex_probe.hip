// Device build (gfx950) of the scalar complex API probe
// (tests/test_complex_api.py): the same element function in a kernel.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "veig_complex.hpp"
#include "complex_probe.inc"

__global__ void k_probe(int op, int64_t n, const double *a, double *out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) probe_one(op, a + 4 * k, out + 2 * k);
}

// a, out: device pointers (4n and 2n doubles); returns a hipError_t
extern "C" int probe_device(int op, int64_t n, const double *a, double *out) {
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_probe, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, 0, op, n, a, out);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  return (int)e;
}

__global__ void k_scale_rc(double *rV, double *iV, double zr, double zi, int I, int I1, int I2, int dim, int n) {
  if (blockIdx.x == 0 && threadIdx.x == 0) scale_rc(rV, iV, zr, zi, I, I1, I2, dim, n);
}

// scale_rc in device code on device planes; returns a hipError_t
extern "C" int probe_scale_rc_device(double *rV, double *iV, double zr, double zi, int I, int I1, int I2, int dim,
                                     int n) {
  hipLaunchKernelGGL(k_scale_rc, dim3(1), dim3(64), 0, 0, rV, iV, zr, zi, I, I1, I2, dim, n);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipDeviceSynchronize();
  return (int)e;
}
