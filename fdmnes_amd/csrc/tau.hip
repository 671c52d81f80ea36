// Post-solve reduction of energy points' solutions to the atom amplitude
// products tau (mat.f90:297-405: regular solution, nspino = 1, one
// representation), on the device, so that only tau (nlmsa^2 complex numbers
// per atom) leaves the GPU instead of the solution (nligne x nlmso).
//
// One workgroup per (energy point, atom):
//   1. real basis (Basereel): norm = I + i X, X the solution's outer-sphere
//      rows (mat.f90:314-328), inverted in LDS by invcomp's sequence
//      (mat.f90:1927-1987 with is = 0): ZGETF2 (pivot = first largest
//      |re| + |im|), ZTRTI2 of the upper factor, ZGETRI's unblocked column
//      sweep and its column interchanges (sub_util.f:15877, 18626, 11382);
//   2. amplitudes a(f, Lm) = sum_k norm^-1(f, k) x(ianew + Lm, k) (norm_t of
//      mat.f90:340, :368-376), or the solution's entries in the complex basis;
//   3. tau(Lm1, Lm2) (+)= -i sum_f a(f, Lm1) conj(a(f, Lm2)) (-i a1 a2 with
//      FDM_comp), mat.f90:392-399.
// The loops keep the reference BLAS order per entry and multiply-adds are not
// fused, so the results follow the reference LAPACK's arithmetic.  The
// matrices are tiny (nlmso = 16 for lmax 3): the kernel is latency-bound,
// a few tens of microseconds per launch.
#include <cfloat>
#include <string>

#include "solver.h"
#include "tau.h"

#pragma clang fp contract(off)

namespace fdmgpu {
namespace {

struct cz {
  double re, im;
};
__device__ inline cz add(cz a, cz b) { return {a.re + b.re, a.im + b.im}; }
__device__ inline cz sub(cz a, cz b) { return {a.re - b.re, a.im - b.im}; }
__device__ inline cz mul(cz a, cz b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline bool nonzero(cz a) { return a.re != 0 || a.im != 0; }
// ONE / z as the reference build evaluates it: (c - i d) / (c^2 + d^2)
__device__ inline cz recip(cz z) {
  const double den = z.re * z.re + z.im * z.im;
  return {(1.0 * z.re + 0.0 * z.im) / den, (0.0 * z.re - 1.0 * z.im) / den};
}
// a / b in the same form
__device__ inline cz cdiv(cz a, cz b) {
  const double den = b.re * b.re + b.im * b.im;
  return {(a.re * b.re + a.im * b.im) / den, (a.im * b.re - a.re * b.im) / den};
}

template <bool CAL>
__device__ inline cz xval(const double *x, int n, int ii, int k) {
  const size_t q = (size_t)ii * n + k;
  if (CAL) return {x[2 * q], x[2 * q + 1]};
  return {x[q], 0.0};
}

// LDS: M (n*n, column-major), work (n), amp (na_max*n, amp[l*n + f]), then ipiv (n ints)
template <bool CAL>
__global__ __launch_bounds__(TAU_NT) void k_tau(TauJob J, int n, int nso, int basereel, int fdm_comp,
                                                int accumulate, int na_max) {
  extern __shared__ cz lds[];
  cz *M = lds;
  cz *work = M + n * n;
  cz *amp = work + n;
  int *ipiv = (int *)(amp + (size_t)na_max * n);
  __shared__ int s_piv, s_nz, s_info;
  const int tid = threadIdx.x, p = blockIdx.x, at = blockIdx.y;
  const double *x = J.x[p];
  const int row0 = J.row[at], na = J.na[at];
#define MM(i, j) M[(j) * n + (i)]

  if (tid == 0) s_info = 0;
  if (basereel) {
    for (int e = tid; e < n * n; e += TAU_NT) {
      const int r = e % n, c = e / n;
      const cz v = xval<CAL>(x, n, nso + c, r);
      cz m = CAL ? cz{-v.im, v.re} : cz{0.0, v.re};  // i * smr (+ i * i smi)
      if (r == c) m.re = 1.0 + m.re;
      M[e] = m;
    }
    // ZGETF2
    for (int j = 0; j < n; j++) {
      __syncthreads();
      if (tid < 64) {
        double best = -1.0;
        int bi = n;
        for (int r = j + tid; r < n; r += 64) {
          const double v = fabs(MM(r, j).re) + fabs(MM(r, j).im);
          if (v > best) { best = v; bi = r; }
        }
        for (int o = 32; o > 0; o >>= 1) {
          const double ob = __shfl_xor(best, o, 64);
          const int oi = __shfl_xor(bi, o, 64);
          if (ob > best || (ob == best && oi < bi)) { best = ob; bi = oi; }
        }
        if (tid == 0) {
          if (bi >= n) bi = j;  // a column of NaNs: IZAMAX keeps its first entry
          s_piv = bi;
          s_nz = nonzero(MM(bi, j));
          ipiv[j] = bi;
        }
      }
      __syncthreads();
      // jp and the pivot test come from LDS words no thread changes before
      // the next step's first barrier: the branch below is uniform
      const int jp = s_piv;
      if (s_nz) {
        if (jp != j)
          for (int c = tid; c < n; c += TAU_NT) {
            const cz t = MM(j, c);
            MM(j, c) = MM(jp, c);
            MM(jp, c) = t;
          }
        __syncthreads();
        const cz d = MM(j, j);
        if (hypot(d.re, d.im) >= DBL_MIN) {
          const cz r = recip(d);
          for (int i = j + 1 + tid; i < n; i += TAU_NT) MM(i, j) = mul(r, MM(i, j));
        } else {
          for (int i = j + 1 + tid; i < n; i += TAU_NT) MM(i, j) = cdiv(MM(i, j), d);
        }
      } else if (tid == 0 && s_info == 0) {
        s_info = j + 1;
      }
      __syncthreads();
      const int m = n - 1 - j;  // ZGERU: A22 += A21 * (-A12)
      for (int e = tid; e < m * m; e += TAU_NT) {
        const int i = j + 1 + e % m, k = j + 1 + e / m;
        const cz y = MM(j, k);
        if (!nonzero(y)) continue;
        const cz t = {-y.re, -y.im};
        MM(i, k) = add(MM(i, k), mul(MM(i, j), t));
      }
    }
    __syncthreads();
    if (s_info) {  // invcomp stops here; the point's outputs are not written
      if (tid == 0 && at == 0 && J.info) J.info[p] = s_info;
      return;
    }
    // ZTRTI2, upper, non-unit: column j := -inv(U_jj) * T * U(0:j-1, j)
    for (int j = 0; j < n; j++) {
      if (tid == 0) MM(j, j) = recip(MM(j, j));
      __syncthreads();
      const cz ajj = {-MM(j, j).re, -MM(j, j).im};
      cz y{0.0, 0.0};
      if (tid < j) {  // ZTRMV's column sweep, as row tid sees it
        const int i = tid;
        y = MM(i, j);
        if (nonzero(y)) y = mul(y, MM(i, i));
        for (int jj = i + 1; jj < j; jj++) {
          const cz t = MM(jj, j);
          if (nonzero(t)) y = add(y, mul(t, MM(i, jj)));
        }
      }
      __syncthreads();
      if (tid < j) MM(tid, j) = mul(ajj, y);
      __syncthreads();
    }
    // ZGETRI: inv(A) from inv(U) and L, right to left
    for (int j = n - 1; j >= 0; j--) {
      for (int i = j + 1 + tid; i < n; i += TAU_NT) {
        work[i] = MM(i, j);
        MM(i, j) = cz{0.0, 0.0};
      }
      __syncthreads();
      for (int i = tid; i < n; i += TAU_NT) {
        cz y = MM(i, j);
        for (int jj = j + 1; jj < n; jj++) {
          const cz w = work[jj];
          if (!nonzero(w)) continue;
          y = add(y, mul(cz{-w.re, -w.im}, MM(i, jj)));
        }
        MM(i, j) = y;
      }
      __syncthreads();
    }
    for (int i = tid; i < n; i += TAU_NT)  // column interchanges, each row on its own
      for (int j = n - 2; j >= 0; j--) {
        const int jp = ipiv[j];
        if (jp != j) {
          const cz t = MM(i, j);
          MM(i, j) = MM(i, jp);
          MM(i, jp) = t;
        }
      }
    __syncthreads();
    if (at == 0) {
      if (J.norm[p])
        for (int e = tid; e < n * n; e += TAU_NT) {
          J.norm[p][2 * e] = M[e].re;
          J.norm[p][2 * e + 1] = M[e].im;
        }
      if (tid == 0 && J.info) J.info[p] = 0;
    }
  } else if (at == 0 && tid == 0 && J.info) {
    J.info[p] = 0;
  }

  // amplitudes of the atom's rows
  for (int e = tid; e < na * n; e += TAU_NT) {
    const int l = e / n, f = e % n, ii = row0 + l;
    cz a;
    if (basereel) {
      a = {0.0, 0.0};
      for (int k = 0; k < n; k++) {
        const cz nt = MM(f, k);
        const cz v = xval<CAL>(x, n, ii, k);
        if (CAL) a = add(a, mul(nt, v));
        else a = add(a, cz{nt.re * v.re, nt.im * v.re});
      }
    } else {
      a = xval<CAL>(x, n, ii, f);
    }
    amp[l * n + f] = a;
  }
  __syncthreads();
  double *tau = J.tau[p] + 2 * J.off[at];
  for (int e = tid; e < na * na; e += TAU_NT) {
    const int l1 = e / na, l2 = e % na;
    cz c{0.0, 0.0};
    for (int f = 0; f < n; f++) {
      const cz a1 = amp[l1 * n + f], a2 = amp[l2 * n + f];
      const cz ia1 = {0.0 * a1.re - 1.0 * a1.im, 0.0 * a1.im + 1.0 * a1.re};  // img * a1
      c = sub(c, mul(ia1, fdm_comp ? a2 : cz{a2.re, -a2.im}));
    }
    const cz old = accumulate ? cz{tau[2 * e], tau[2 * e + 1]} : cz{0.0, 0.0};
    tau[2 * e] = old.re + c.re;
    tau[2 * e + 1] = old.im + c.im;
  }
#undef MM
}

// Tau_RI (mat.f90:229-295, complex energy, FDM_comp, Test_reg = 0): one
// workgroup per (point, atom); entry (i_RI, Lm1, Lm2) of the atom's block is
// Tau_RI += sum_f Ampl(f, Lm1) Ampl_I(f, (i_RI-1) na + Lm2), each product
// added straight into it in f order as :286-288 does, with Ampl the regular
// solution's atom rows rounded to single precision (the reference keeps the
// regular solution as cmplx(smr, smi) of default kind, :211) and Ampl_I the
// irregular solution's (its first n_ato_RI * na rows of the atom's blocks).
__global__ __launch_bounds__(TAU_NT) void k_tau_ri(TauRIJob J, int n, int n_ato_ri, int n_ri, int accumulate) {
  const int p = blockIdx.x, at = blockIdx.y, na = J.na[at];
  const double *xr = J.xr[p], *xi = J.xi[p];
  double *t = J.tau[p] + 2 * J.off[at];
  for (int e = threadIdx.x; e < n_ri * na * na; e += TAU_NT) {
    const int i_ri = e / (na * na), l1 = (e / na) % na, l2 = e % na;
    const int lm = i_ri * na + l2;  // 0-based column of Ampl_I
    double tr = accumulate ? t[2 * e] : 0.0, ti = accumulate ? t[2 * e + 1] : 0.0;
    if (lm < n_ato_ri * na) {
      const size_t a = (size_t)(J.row_reg[at] + l1) * n, b = (size_t)(J.row_irr[at] + lm) * n;
      for (int f = 0; f < n; f++) {
        const double ar = (double)(float)xr[2 * (a + f)], ai = (double)(float)xr[2 * (a + f) + 1];
        const double br = xi[2 * (b + f)], bi = xi[2 * (b + f) + 1];
        tr = tr + (ar * br - ai * bi);
        ti = ti + (ar * bi + ai * br);
      }
    }
    t[2 * e] = tr;
    t[2 * e + 1] = ti;
  }
}

// The symmetrisation at the end of mat (mat.f90:596-641, one spin, no
// spin-orbit) of each atom's nlmsa^2 block, harmonics in (L, m) order: with
// complex harmonics (Ylm_comp) the reference visits (Lm1, Lm2) in order and
// sets it to c = 0.5 (T(Lm1, Lm2) + s T(P)) and its partner P = (L2 -m2,
// L1 -m1) to s c, s = (-1)^(m1+m2); a later visit of the partner finds the
// pair consistent and rewrites the same values, so every entry follows
// from the block before the sweep: the earlier of an entry and its partner
// in that order supplies c.  With real harmonics the reference keeps
// i Im(T).  One workgroup per (point, atom), the block staged in LDS.
__global__ __launch_bounds__(TAU_NT) void k_tau_sym(TauSymJob J, int ylm_comp) {
  extern __shared__ cz blk[];
  const int p = blockIdx.x, at = blockIdx.y, na = J.na[at];
  double *t = J.tau[p] + 2 * J.off[at];
  for (int e = threadIdx.x; e < na * na; e += TAU_NT) blk[e] = cz{t[2 * e], t[2 * e + 1]};
  __syncthreads();
  auto lm_of = [](int k, int *m) {  // 0-based (L, m)-ordered index -> m
    int L = (int)sqrt((double)k);
    while (L * L > k) L--;
    while ((L + 1) * (L + 1) <= k) L++;
    *m = k - L * L - L;
  };
  for (int e = threadIdx.x; e < na * na; e += TAU_NT) {
    const int a = e / na, b = e % na;
    cz v;
    if (ylm_comp) {
      int m1, m2;
      lm_of(a, &m1);
      lm_of(b, &m2);
      const int pa = b - 2 * m2, pb = a - 2 * m1, pe = pa * na + pb;
      const double s = ((m1 + m2) & 1) ? -1.0 : 1.0;
      if (e < pe) {
        v = cz{0.5 * (blk[e].re + s * blk[pe].re), 0.5 * (blk[e].im + s * blk[pe].im)};
      } else if (e > pe) {  // the partner came first: c = 0.5 (T(P) + s T(e)), this entry s c
        const cz c = {0.5 * (blk[pe].re + s * blk[e].re), 0.5 * (blk[pe].im + s * blk[e].im)};
        v = cz{s * c.re, s * c.im};
      } else {  // its own partner: c then s c written to the same place
        const cz c = {0.5 * (blk[e].re + s * blk[e].re), 0.5 * (blk[e].im + s * blk[e].im)};
        v = cz{s * c.re, s * c.im};
      }
    } else {
      v = cz{0.0, blk[e].im};
    }
    t[2 * e] = v.re;
    t[2 * e + 1] = v.im;
  }
}

}  // namespace

int tau_ri_launch(const TauRIJob &J, int npts, int natoms, int n, int n_ato_ri, int n_ri, bool accumulate,
                  hipStream_t s) {
  hipLaunchKernelGGL(k_tau_ri, dim3(npts, natoms), dim3(TAU_NT), 0, s, J, n, n_ato_ri, n_ri, (int)accumulate);
  FDM_HIP(hipGetLastError());
  return 0;
}

int tau_sym_launch(const TauSymJob &J, int npts, int natoms, int na_max, bool ylm_comp, hipStream_t s) {
  const size_t lds = sizeof(cz) * (size_t)na_max * na_max;
  if (lds > 64 * 1024)
    FDM_HIP(hipFuncSetAttribute((const void *)k_tau_sym, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  hipLaunchKernelGGL(k_tau_sym, dim3(npts, natoms), dim3(TAU_NT), lds, s, J, (int)ylm_comp);
  FDM_HIP(hipGetLastError());
  return 0;
}

size_t tau_lds_bytes(int n, int na_max) {
  return sizeof(cz) * ((size_t)n * n + n + (size_t)na_max * n) + sizeof(int) * n;
}

int tau_launch(const TauJob &J, int npts, int natoms, int na_max, int n, int nso, bool cal, bool basereel,
               bool fdm_comp, bool accumulate, hipStream_t s) {
  const size_t lds = tau_lds_bytes(n, na_max);
  const void *fn = cal ? (const void *)k_tau<true> : (const void *)k_tau<false>;
  if (lds > 64 * 1024) FDM_HIP(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  if (cal)
    hipLaunchKernelGGL(k_tau<true>, dim3(npts, natoms), dim3(TAU_NT), lds, s, J, n, nso, (int)basereel,
                       (int)fdm_comp, (int)accumulate, na_max);
  else
    hipLaunchKernelGGL(k_tau<false>, dim3(npts, natoms), dim3(TAU_NT), lds, s, J, n, nso, (int)basereel,
                       (int)fdm_comp, (int)accumulate, na_max);
  FDM_HIP(hipGetLastError());
  return 0;
}

}  // namespace fdmgpu
