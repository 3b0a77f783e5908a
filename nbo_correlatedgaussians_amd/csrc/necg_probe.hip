// necg_probe.hip -- the FP64 vector FMA rate this device sustains: a
// measurement aid that sets the pair kernels' FLOP rates beside a number
// measured on the same device (MI355X publishes 78.6 TFLOP/s, 256 CUs x
// 128 FP64 FLOP per clock x 2.4 GHz).
#include <hip/hip_runtime.h>

#include "../../include/necg.h"

namespace {

// Eight independent FMA chains per lane, so that one wave per SIMD already
// issues an FMA every cycle it may; the chains' sums are stored so that no
// chain is dead code.
__global__ void __launch_bounds__(256) fma_rate_kernel(double m, double c, int iters,
                                                       double *__restrict__ out) {
  double a[8];
#pragma unroll
  for (int i = 0; i < 8; ++i) a[i] = 1.0 + 1e-3 * (double)(threadIdx.x + i);
#pragma unroll 4
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < 8; ++i) a[i] = __builtin_fma(a[i], m, c);
  }
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < 8; ++i) s += a[i];
  out[(size_t)blockIdx.x * blockDim.x + threadIdx.x] = s;
}

int hip_code(hipError_t e) {
  if (e == hipSuccess) return NECG_OK;
  return e == hipErrorOutOfMemory ? NECG_ERR_OOM : NECG_ERR_DEVICE;
}

}  // namespace

extern "C" int necg_fp64_fma_rate(int32_t device, int32_t iters, double *tflops) {
  if (!tflops || iters < 1) return NECG_ERR_ARG;
  int prev = 0, ncu = 0;
  hipError_t e = hipGetDevice(&prev);
  if (e == hipSuccess) e = hipSetDevice(device);
  if (e == hipSuccess) e = hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device);
  if (e != hipSuccess) return hip_code(e);
  const int grid = 8 * ncu, block = 256;   // eight workgroups (two waves per SIMD) per CU
  double *out = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float best = 0.0f;
  e = hipMalloc((void **)&out, sizeof(double) * grid * block);
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  for (int rep = 0; rep < 4 && e == hipSuccess; ++rep) {   // the first run warms up
    e = hipEventRecord(e0, nullptr);
    if (e == hipSuccess) {
      hipLaunchKernelGGL(fma_rate_kernel, dim3(grid), dim3(block), 0, nullptr, 1.0 - 1e-9, 1e-9,
                         iters, out);
      e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipEventRecord(e1, nullptr);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float ms = 0.0f;
    if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
    if (e == hipSuccess && rep > 0 && (best == 0.0f || ms < best)) best = ms;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (out) (void)hipFree(out);
  (void)hipSetDevice(prev);
  if (e != hipSuccess) return hip_code(e);
  const double flop = 2.0 * 8.0 * (double)iters * grid * block;
  *tflops = best > 0.0f ? flop / (best * 1e-3) / 1e12 : 0.0;
  return NECG_OK;
}
