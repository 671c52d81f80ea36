// FP64 iterative refinement around FP32 factors, and the sigma reduction.
#pragma once
#include "solver.h"

namespace fdmgpu {

// Compressed-row matrix on the device (FP64, 1-based columns).
struct DevCsr {
  int n = 0;
  int64_t nnz = 0;
  const int64_t *row_ptr = nullptr;
  const int *col = nullptr;
  const double *vals = nullptr;  // interleaved (re, im) when complex
  const int *long_rows = nullptr;  // rows longer than LONG_ROW entries (the outer-sphere rows)
  int nlong = 0;
};

// Rows with more entries than this are summed by a whole workgroup.
constexpr int LONG_ROW = 256;

// R = B - A X over nrhs right-hand sides, all rows*nrhs interleaved when C.
template <bool C>
int residual(const DevCsr &a, const double *B, const double *X, double *R, int nrhs, hipStream_t s);

// Per right-hand side squared 2-norms of V (n rows), summed on the host.
template <bool C>
int col_norms2(const double *V, int n, int nrhs, double *out, double *dscratch, int scratch_elems,
               hipStream_t s);

// Launch-only variant (graph-capturable): partial sums of NORM_GRID blocks
// per right-hand side into dscratch[NORM_GRID * nrhs]; finish with
// sum_norm_partials on the host copy.
constexpr int NORM_GRID = 256;
template <bool C>
int launch_norms2(const double *V, int n, int nrhs, double *dscratch, hipStream_t s);
void sum_norm_partials(const double *h, int nrhs, double *out);

// dst (R) <- src (double), elementwise, count elements.
template <class R>
int convert(const double *src, R *dst, size_t count, hipStream_t s);
// X += D (D in precision R), count elements.
template <class R>
int accumulate(double *X, const R *D, size_t count, hipStream_t s);

}  // namespace fdmgpu
