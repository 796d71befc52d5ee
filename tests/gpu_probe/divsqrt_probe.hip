// Test infrastructure: runs the FP64 quotient and square-root helpers of the stage kernel
// (DivBy, sqrt_ieee, markstein_div in fv2d_amd/csrc/fv2d_math.h) on the GPU next to the
// compiler's own a / b and sqrt, for a host-side bitwise comparison. Built with the flags of the FP64 units.
#include <hip/hip_runtime.h>

#include "../../fv2d_amd/csrc/fv2d_math.h"

using namespace fv2d;

// Per i: out[6i + 0..5] = DivBy(b, a0).first(), DivBy(b, a0)(a1), a0 / b, a1 / b,
// sqrt_ieee(a0), sqrt(a0).
__global__ void probe_kernel(const double *a0, const double *a1, const double *b, int n, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double x0 = a0[i], x1 = a1[i], d = b[i];
  const DivBy<double> by(d, x0);
  out[6 * i + 0] = by.first();
  out[6 * i + 1] = by(x1);
  out[6 * i + 2] = x0 / d;
  out[6 * i + 3] = x1 / d;
  out[6 * i + 4] = sqrt_ieee(x0);
  out[6 * i + 5] = sqrt(x0);
}

// Per i: out[2i] = markstein_div(a[i], d, rd, fast), out[2i + 1] = a[i] / d.
__global__ void markstein_kernel(const double *a, int n, double d, double rd, int fast, double *out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  out[2 * i] = markstein_div(a[i], d, rd, fast);
  out[2 * i + 1] = a[i] / d;
}

extern "C" int markstein_probe(const double *a, int n, double d, double rd, int fast, double *out) {
  if (n <= 0) return -1;
  double *m = nullptr;
  const size_t in_bytes = sizeof(double) * (size_t)n;
  if (hipMalloc((void **)&m, 3 * in_bytes) != hipSuccess) return -2;
  int rc = 0;
  if (hipMemcpy(m, a, in_bytes, hipMemcpyHostToDevice) != hipSuccess) {
    rc = -3;
  } else {
    hipLaunchKernelGGL(markstein_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, m, n, d, rd, fast, m + n);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(out, m + n, 2 * in_bytes, hipMemcpyDeviceToHost) != hipSuccess)
      rc = -4;
  }
  (void)hipFree(m);
  return rc;
}

extern "C" int divsqrt_probe(const double *a0, const double *a1, const double *b, int n, double *out) {
  if (n <= 0) return -1;
  double *d = nullptr;
  const size_t in_bytes = sizeof(double) * (size_t)n, out_bytes = 6 * in_bytes;
  if (hipMalloc((void **)&d, 3 * in_bytes + out_bytes) != hipSuccess) return -2;
  int rc = 0;
  if (hipMemcpy(d, a0, in_bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d + n, a1, in_bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d + 2 * (size_t)n, b, in_bytes, hipMemcpyHostToDevice) != hipSuccess) {
    rc = -3;
  } else {
    hipLaunchKernelGGL(probe_kernel, dim3((n + 255) / 256), dim3(256), 0, 0, d, d + n, d + 2 * (size_t)n, n,
                       d + 3 * (size_t)n);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(out, d + 3 * (size_t)n, out_bytes, hipMemcpyDeviceToHost) != hipSuccess)
      rc = -4;
  }
  (void)hipFree(d);
  return rc;
}
