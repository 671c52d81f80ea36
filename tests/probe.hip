// Lane-layout probe of the 16x16x4 MFMA builtins (test support, not part of
// the solver ABI).  One wave computes D = A B with A[i][k] = i + 16 k + 1 and
// B[k][j] = (k + 1) * 1000 + j, feeding lane l with A[l%16][l/16] and
// B[l/16][l%16]; every lane writes its four outputs to out[4*l + v].
#include <hip/hip_runtime.h>

typedef double d4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

__global__ void k_probe64(double *out) {
  const int l = threadIdx.x;
  const double a = (l % 16) + 16 * (l / 16) + 1;
  const double b = (l / 16 + 1) * 1000.0 + (l % 16);
  d4 c = {0, 0, 0, 0};
  c = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  for (int v = 0; v < 4; v++) out[4 * l + v] = c[v];
}

__global__ void k_probe32(double *out) {
  const int l = threadIdx.x;
  const float a = (l % 16) + 16 * (l / 16) + 1;
  const float b = (l / 16 + 1) * 1000.0f + (l % 16);
  f4 c = {0, 0, 0, 0};
  c = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  for (int v = 0; v < 4; v++) out[4 * l + v] = c[v];
}

extern "C" int fdmgpu_probe_mfma(int prec, double *host_out) {
  double *d = nullptr;
  if (hipMalloc(&d, sizeof(double) * 256) != hipSuccess) return -2;
  if (prec == 0)
    hipLaunchKernelGGL(k_probe64, dim3(1), dim3(64), 0, 0, d);
  else
    hipLaunchKernelGGL(k_probe32, dim3(1), dim3(64), 0, 0, d);
  hipError_t e = hipMemcpy(host_out, d, sizeof(double) * 256, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return e == hipSuccess ? 0 : -2;
}

// XCD placement probe: workgroup w of a grid of nwg workgroups (256 threads)
// writes the id of the XCD (accelerator complex die) it runs on, read from
// the XCC_ID hardware register (hwreg 20, bits 3:0), to out[w].
__global__ void k_probe_xcc(int *out) {
  if (threadIdx.x == 0) out[blockIdx.x] = __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 15;
}

extern "C" int fdmgpu_probe_xcc(int nwg, int *host_out) {
  if (nwg <= 0) return -1;
  int *d = nullptr;
  if (hipMalloc(&d, sizeof(int) * (size_t)nwg) != hipSuccess) return -2;
  hipLaunchKernelGGL(k_probe_xcc, dim3(nwg), dim3(256), 0, 0, d);
  hipError_t e = hipMemcpy(host_out, d, sizeof(int) * (size_t)nwg, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  return e == hipSuccess ? 0 : -2;
}
