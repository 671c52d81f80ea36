// FP64 iterative refinement around FP32 factors.
#include <algorithm>
#include <cmath>
#include <vector>

#include "refine.h"

namespace fdmgpu {

// One thread per (row, right-hand side): r = b - sum_q a_q x_{col_q}.  Rows
// longer than LONG_ROW are left to k_residual_long.
template <bool C>
__global__ void k_residual(int n, const int64_t *__restrict__ rp, const int *__restrict__ col,
                           const double *__restrict__ av, const double *__restrict__ B,
                           const double *__restrict__ X, double *__restrict__ R, int nrhs) {
  constexpr int NC = C ? 2 : 1;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)n * nrhs) return;
  const int row = (int)(t / nrhs), l = (int)(t % nrhs);
  if (rp[row + 1] - rp[row] > LONG_ROW) return;
  double sr = B[t * NC], si = C ? B[t * NC + 1] : 0.0;
  for (int64_t q = rp[row]; q < rp[row + 1]; q++) {
    const int64_t xi = ((int64_t)(col[q] - 1) * nrhs + l) * NC;
    const double xr = X[xi];
    if constexpr (C) {
      const double xim = X[xi + 1], ar = av[2 * q], ai = av[2 * q + 1];
      sr -= ar * xr - ai * xim;
      si -= ar * xim + ai * xr;
    } else {
      sr -= av[q] * xr;
    }
  }
  R[t * NC] = sr;
  if constexpr (C) R[t * NC + 1] = si;
}

// Long rows (the dense outer-sphere rows): workgroup (row, chunk of 16
// right-hand sides); thread (e, l) sums entries rp[row] + e, e + 64, ... for
// right-hand side l, then the 64 partial sums of each l are added in LDS.
constexpr int LONG_LANES = 64;
template <bool C>
__global__ __launch_bounds__(LONG_LANES * 16) void k_residual_long(const int *__restrict__ rows,
                                                                 const int64_t *__restrict__ rp,
                                                                 const int *__restrict__ col,
                                                                 const double *__restrict__ av,
                                                                 const double *__restrict__ B,
                                                                 const double *__restrict__ X,
                                                                 double *__restrict__ R, int nrhs) {
  constexpr int NC = C ? 2 : 1;
  __shared__ double sh[NC][LONG_LANES][17];
  const int row = rows[blockIdx.x];
  const int l = threadIdx.x & 15, e = threadIdx.x >> 4, lg = blockIdx.y * 16 + l;
  double sr = 0, si = 0;
  if (lg < nrhs) {
    for (int64_t q = rp[row] + e; q < rp[row + 1]; q += LONG_LANES) {
      const int64_t xi = ((int64_t)(col[q] - 1) * nrhs + lg) * NC;
      if constexpr (C) {
        const double xr = X[xi], xim = X[xi + 1], ar = av[2 * q], ai = av[2 * q + 1];
        sr += ar * xr - ai * xim;
        si += ar * xim + ai * xr;
      } else {
        sr += av[q] * X[xi];
      }
    }
  }
  sh[0][e][l] = sr;
  if constexpr (C) sh[NC - 1][e][l] = si;
  __syncthreads();
  if (e == 0 && lg < nrhs) {
    double tr = 0, ti = 0;
    for (int k = 0; k < LONG_LANES; k++) {
      tr += sh[0][k][l];
      if constexpr (C) ti += sh[NC - 1][k][l];
    }
    const int64_t t = ((int64_t)row * nrhs + lg) * NC;
    R[t] = B[t] - tr;
    if constexpr (C) R[t + 1] = B[t + 1] - ti;
  }
}

// Partial squared norms per right-hand side: block b sums rows b, b+grid, ...
template <bool C>
__global__ void k_norms2(const double *__restrict__ V, int n, int nrhs, double *__restrict__ part) {
  constexpr int NC = C ? 2 : 1;
  extern __shared__ double sh[];
  const int l = threadIdx.x % nrhs;
  const int rows_per_pass = blockDim.x / nrhs;
  const int r0 = threadIdx.x / nrhs;
  double s = 0;
  if (r0 < rows_per_pass) {
    for (int row = blockIdx.x * rows_per_pass + r0; row < n; row += gridDim.x * rows_per_pass) {
      const int64_t e = ((int64_t)row * nrhs + l) * NC;
      s += V[e] * V[e];
      if constexpr (C) s += V[e + 1] * V[e + 1];
    }
  }
  sh[threadIdx.x] = (r0 < rows_per_pass) ? s : 0.0;
  __syncthreads();
  if (threadIdx.x < nrhs) {
    double acc = 0;
    for (int k = threadIdx.x; k < rows_per_pass * nrhs; k += nrhs) acc += sh[k];
    part[(int64_t)blockIdx.x * nrhs + threadIdx.x] = acc;
  }
}

template <class R>
__global__ void k_convert(const double *__restrict__ s, R *__restrict__ d, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) d[i] = R(s[i]);
}

template <class R>
__global__ void k_accumulate(double *__restrict__ x, const R *__restrict__ d, size_t count) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) x[i] += double(d[i]);
}

template <bool C>
int residual(const DevCsr &a, const double *B, const double *X, double *R, int nrhs, hipStream_t s) {
  const int64_t tot = (int64_t)a.n * nrhs;
  hipLaunchKernelGGL((k_residual<C>), dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, s, a.n, a.row_ptr, a.col,
                     a.vals, B, X, R, nrhs);
  if (a.nlong > 0)
    hipLaunchKernelGGL((k_residual_long<C>), dim3(a.nlong, (nrhs + 15) / 16), dim3(LONG_LANES * 16), 0, s,
                       a.long_rows, a.row_ptr, a.col, a.vals, B, X, R, nrhs);
  FDM_HIP(hipGetLastError());
  return 0;
}

template <bool C>
int col_norms2(const double *V, int n, int nrhs, double *out, double *dscratch, int scratch_elems, hipStream_t s) {
  if (nrhs > 256) {
    set_error("col_norms2: at most 256 right-hand sides");
    return -1;
  }
  int grid = std::max(1, std::min(256, scratch_elems / nrhs));
  const int threads = 256;
  hipLaunchKernelGGL((k_norms2<C>), dim3(grid), dim3(threads), threads * sizeof(double), s, V, n, nrhs, dscratch);
  FDM_HIP(hipGetLastError());
  std::vector<double> h((size_t)grid * nrhs);
  FDM_HIP(hipMemcpyAsync(h.data(), dscratch, sizeof(double) * h.size(), hipMemcpyDeviceToHost, s));
  FDM_HIP(hipStreamSynchronize(s));
  for (int l = 0; l < nrhs; l++) {
    double acc = 0;
    for (int g = 0; g < grid; g++) acc += h[(size_t)g * nrhs + l];
    out[l] = acc;
  }
  return 0;
}

template <bool C>
int launch_norms2(const double *V, int n, int nrhs, double *dscratch, hipStream_t s) {
  if (nrhs > 256) {
    set_error("launch_norms2: at most 256 right-hand sides");
    return -1;
  }
  hipLaunchKernelGGL((k_norms2<C>), dim3(NORM_GRID), dim3(256), 256 * sizeof(double), s, V, n, nrhs, dscratch);
  FDM_HIP(hipGetLastError());
  return 0;
}

void sum_norm_partials(const double *h, int nrhs, double *out) {
  for (int l = 0; l < nrhs; l++) {
    double acc = 0;
    for (int g = 0; g < NORM_GRID; g++) acc += h[(size_t)g * nrhs + l];
    out[l] = acc;
  }
}

template <class R>
int convert(const double *src, R *dst, size_t count, hipStream_t s) {
  hipLaunchKernelGGL((k_convert<R>), dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, src, dst, count);
  FDM_HIP(hipGetLastError());
  return 0;
}

template <class R>
int accumulate(double *X, const R *D, size_t count, hipStream_t s) {
  hipLaunchKernelGGL((k_accumulate<R>), dim3((unsigned)((count + 255) / 256)), dim3(256), 0, s, X, D, count);
  FDM_HIP(hipGetLastError());
  return 0;
}

template int residual<true>(const DevCsr &, const double *, const double *, double *, int, hipStream_t);
template int residual<false>(const DevCsr &, const double *, const double *, double *, int, hipStream_t);
template int col_norms2<true>(const double *, int, int, double *, double *, int, hipStream_t);
template int col_norms2<false>(const double *, int, int, double *, double *, int, hipStream_t);
template int launch_norms2<true>(const double *, int, int, double *, hipStream_t);
template int launch_norms2<false>(const double *, int, int, double *, hipStream_t);
template int convert<float>(const double *, float *, size_t, hipStream_t);
template int accumulate<float>(double *, const float *, size_t, hipStream_t);
template int convert<double>(const double *, double *, size_t, hipStream_t);
template int accumulate<double>(double *, const double *, size_t, hipStream_t);

}  // namespace fdmgpu
