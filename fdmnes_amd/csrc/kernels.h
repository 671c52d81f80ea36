// Device kernels of the block-skyline UL solver (gfx950).
//
// Storage ("block skyline"): the nligne x nligne profile matrix is cut into
// TB x TB tiles.  Block row I keeps the tiles J in [J0(I), J1(I)], the block
// envelope of its rows' [lb1, lb2] profiles (lb1 non-decreasing, J1 swept
// non-decreasing), contiguously from tbase(I).  A tile is TB*TB elements,
// row-major; complex elements are interleaved (re, im).  Because lb1 is
// monotone the envelope is closed under fill-in, so the UL factors overwrite
// the tiles in place: tile (K,K) holds U_D (upper incl. diagonal) and L_D
// (strict lower, unit diagonal implied) of the diagonal block; tiles (K,J<K)
// hold the unit-lower factor L; tiles (I<K,K) hold U.
// dinv(K) holds U_D^-1 and L_D^-1 of each diagonal block.
//
// Every kernel is grouped: one launch serves the same step of up to MAXG
// independent systems (energy points, clusters of different size).  A launch
// carries an Items list (system, block step, first workgroup); workgroup b
// serves the item g with off[g] <= b < off[g+1].
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fdmgpu {

constexpr int TB = 64;    // tile edge = panel width
constexpr int NT = 256;   // threads per workgroup
constexpr int RC = 16;    // right-hand sides per solve chunk
constexpr int MAXG = 16;  // systems per grouped launch

struct DevPlan {
  int n;   // real rows
  int NB;  // block rows (NB*TB >= n, padded rows are identity)
  const int *J0, *J1, *I0, *Jx;  // per block: first/last tile column, first block row
                                  // reaching column K, last block row whose J0 <= K
  const int64_t *tbase;           // per block row: first tile index
};

// One system of a group, as the kernels see it.
template <class R>
struct SysDev {
  DevPlan p;
  R *tiles;
  R *dinv;
  int *fail;
  R *X;     // right-hand sides / solution, (NB*TB) x nrhs, interleaved when complex
  R *part;  // solve partial sums
  int nrhs;
};

struct Items {
  int count;
  int sys[MAXG];
  int K[MAXG];
  int off[MAXG + 1];
};

__device__ __forceinline__ int find_item(const Items &it, int b) {
  int g = 0;
  while (g + 1 < it.count && b >= it.off[g + 1]) g++;
  return g;
}

template <class R, bool C> struct Elem;
template <class R> struct Elem<R, true> { static constexpr int NC = 2; };
template <class R> struct Elem<R, false> { static constexpr int NC = 1; };

template <class R, bool C>
__device__ __forceinline__ R *tile_ptr(const DevPlan &p, R *tiles, int I, int J) {
  return tiles + (p.tbase[I] + (J - p.J0[I])) * (int64_t)(TB * TB * Elem<R, C>::NC);
}

// ---------------------------------------------------------------- MFMA
template <class R> struct Mfma;
// Operands: lane l holds A[l%16][l/16] and B[l/16][l%16].  Output v of lane
// l is D[row(l/16, v)][l%16]; the row map differs between the FP64 and the
// FP32 instruction (checked on the device by tests/test_gpu_mfma_layout.py).
template <> struct Mfma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 run(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int v) { return lk + 4 * v; }
};
template <> struct Mfma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 run(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lk, int v) { return 4 * lk + v; }
};

enum GemmMode { G_UPDATE = 0, G_ROWTRSM = 1, G_COLTRSM = 2 };

// Stage one TB x TB tile (global, interleaved) into split LDS planes.
template <class R, bool C, int LD>
__device__ __forceinline__ void stage_tile(const R *__restrict__ g, R (*s)[TB][LD]) {
  const int tid = threadIdx.x;
  if constexpr (C) {
    typedef R r2 __attribute__((ext_vector_type(2)));
    const r2 *g2 = reinterpret_cast<const r2 *>(g);
#pragma unroll
    for (int q = 0; q < TB * TB / NT; q++) {
      int e = tid + q * NT;
      r2 v = g2[e];
      s[0][e / TB][e % TB] = v.x;
      s[1][e / TB][e % TB] = v.y;
    }
  } else {
#pragma unroll
    for (int q = 0; q < TB * TB / NT; q++) {
      int e = tid + q * NT;
      s[0][e / TB][e % TB] = g[e];
    }
  }
}

// C(TB x TB) <- f(A * B), one output tile per work item; a persistent grid
// strides over the work items of all grouped systems.
//   G_UPDATE : tile(I,J) -= tile(I,K) * tile(K,J)   per item nI * nJ workgroups
//   G_ROWTRSM: tile(K,J)  = Uinv(K)   * tile(K,J)   per item nJ          (L rows)
//   G_COLTRSM: tile(I,K)  = tile(I,K) * Linv(K)     per item nI          (U columns)
// Four waves; wave w owns the 32x32 quadrant (w/2, w%2), 2x2 MFMA blocks of
// 16x16; a complex product is four real MFMAs (no 3M recombination).
template <class R, bool C>
__global__ __launch_bounds__(NT) void k_tile_gemm(const SysDev<R> *__restrict__ sys, Items it, int mode) {
  constexpr int NC = Elem<R, C>::NC;
  constexpr int LD = TB + 1;
  __shared__ R As[NC][TB][LD];
  __shared__ R Bs[NC][TB][LD];
  const int total = it.off[it.count];
  for (int wb = blockIdx.x; wb < total; wb += gridDim.x) {
  if (wb != (int)blockIdx.x) __syncthreads();  // LDS reuse across work items
  const int g = find_item(it, wb);
  const int b = wb - it.off[g];
  const SysDev<R> &S = sys[it.sys[g]];
  const DevPlan &p = S.p;
  const int K = it.K[g];
  const int I0 = p.I0[K], J0 = p.J0[K];
  const int nJ = K - J0;
  const R *A;
  const R *B;
  R *Cg;
  if (mode == G_UPDATE) {
    const int I = I0 + b / nJ, J = J0 + b % nJ;
    A = tile_ptr<R, C>(p, S.tiles, I, K);
    B = tile_ptr<R, C>(p, S.tiles, K, J);
    Cg = tile_ptr<R, C>(p, S.tiles, I, J);
  } else if (mode == G_ROWTRSM) {
    A = S.dinv + (int64_t)K * 2 * TB * TB * NC;  // Uinv
    Cg = tile_ptr<R, C>(p, S.tiles, K, J0 + b);
    B = Cg;
  } else {
    Cg = tile_ptr<R, C>(p, S.tiles, I0 + b, K);
    A = Cg;
    B = S.dinv + ((int64_t)K * 2 + 1) * TB * TB * NC;  // Linv
  }
  stage_tile<R, C, LD>(A, As);
  stage_tile<R, C, LD>(B, Bs);
  __syncthreads();

  typedef typename Mfma<R>::v4 v4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r0 = (wave >> 1) * 32, c0 = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  v4 accr[2][2], acci[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int bb = 0; bb < 2; bb++) {
      accr[a][bb] = v4{0, 0, 0, 0};
      acci[a][bb] = v4{0, 0, 0, 0};
    }
#pragma unroll 4
  for (int k0 = 0; k0 < TB; k0 += 4) {
    R ar[2], ai[2], br[2], bi[2];
#pragma unroll
    for (int a = 0; a < 2; a++) {
      ar[a] = As[0][r0 + a * 16 + li][k0 + lk];
      if constexpr (C) ai[a] = As[1][r0 + a * 16 + li][k0 + lk];
    }
#pragma unroll
    for (int bb = 0; bb < 2; bb++) {
      br[bb] = Bs[0][k0 + lk][c0 + bb * 16 + li];
      if constexpr (C) bi[bb] = Bs[1][k0 + lk][c0 + bb * 16 + li];
    }
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int bb = 0; bb < 2; bb++) {
        accr[a][bb] = Mfma<R>::run(ar[a], br[bb], accr[a][bb]);
        if constexpr (C) {
          accr[a][bb] = Mfma<R>::run(-ai[a], bi[bb], accr[a][bb]);
          acci[a][bb] = Mfma<R>::run(ar[a], bi[bb], acci[a][bb]);
          acci[a][bb] = Mfma<R>::run(ai[a], br[bb], acci[a][bb]);
        }
      }
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int bb = 0; bb < 2; bb++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = r0 + a * 16 + Mfma<R>::row(lk, v), c = c0 + bb * 16 + li;
        const int64_t e = (int64_t)(r * TB + c) * NC;
        if (mode == G_UPDATE) {
          Cg[e] -= accr[a][bb][v];
          if constexpr (C) Cg[e + 1] -= acci[a][bb][v];
        } else {
          Cg[e] = accr[a][bb][v];
          if constexpr (C) Cg[e + 1] = acci[a][bb][v];
        }
      }
  }
}

// ---------------------------------------------------------------- trailing update
// Work items of the trailing-update kernel: per item an output rectangle of
// tiles [i0, i0 + n/nj) x [j0, j0 + nj) and a range of block steps
// [Kb, Kb + nk).  Tile (I,J) accumulates U(I,K) L(K,J) over the steps K of the
// range whose envelope reaches it (I0(K) <= I and J0(K) <= J; I0, J0 are
// non-decreasing in K so these steps are a prefix of the range).
constexpr int MAXT = 2 * MAXG;
struct TrailItems {
  int count;
  int sys[MAXT];
  int Kb[MAXT];
  int nk[MAXT];
  int i0[MAXT];
  int j0[MAXT];
  int nj[MAXT];
  int off[MAXT + 1];
};

__device__ __forceinline__ int find_trail(const TrailItems &it, int b) {
  int g = 0;
  while (g + 1 < it.count && b >= it.off[g + 1]) g++;
  return g;
}

constexpr int KC = 16;  // k-chunk of the LDS pipeline

// tile(I,J) -= sum_K tile(I,K) tile(K,J).  Persistent grid over the work
// items of all systems; per tile the K dimension (nk*TB) streams through a
// two-buffer LDS pipeline in chunks of KC, the next chunk's global loads in
// flight while the current chunk feeds the MFMAs.  Four waves, each a 32x32
// quadrant in 2x2 blocks of 16x16x4 MFMAs; a complex product is four real
// MFMAs.  LDS holds A chunks k-major (As[k][row]) so the MFMA A operand reads
// 16 consecutive rows per k.
template <class R, bool C>
__global__ __launch_bounds__(NT, 2) void k_trail(const SysDev<R> *__restrict__ sys, TrailItems it) {
  constexpr int NC = Elem<R, C>::NC;
  constexpr int LD = TB + 1;
  constexpr int PER = TB * KC / NT;  // elements per thread per chunk and operand (4)
  __shared__ R As[2][NC][KC][LD];
  __shared__ R Bs[2][NC][KC][LD];
  typedef typename Mfma<R>::v4 v4;
  typedef R r2 __attribute__((ext_vector_type(2)));
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int r0 = (wave >> 1) * 32, c0 = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  const int total = it.off[it.count];
  for (int wb = blockIdx.x; wb < total; wb += gridDim.x) {
    const int g = find_trail(it, wb);
    const int b = wb - it.off[g];
    const SysDev<R> &S = sys[it.sys[g]];
    const DevPlan &p = S.p;
    const int I = it.i0[g] + b / it.nj[g], J = it.j0[g] + b % it.nj[g];
    int m = 0;
    while (m < it.nk[g] && p.I0[it.Kb[g] + m] <= I && p.J0[it.Kb[g] + m] <= J) m++;
    const int nch = m * (TB / KC);
    v4 accr[2][2], acci[2][2];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int bb = 0; bb < 2; bb++) {
        accr[a][bb] = v4{0, 0, 0, 0};
        acci[a][bb] = v4{0, 0, 0, 0};
      }
    // chunk c: step K = Kb + c/4, columns kc..kc+15 of A = tile(I,K), rows of B = tile(K,J)
    r2 pa[PER], pb[PER];
    R qa[PER], qb[PER];
    auto fetch = [&](int c) {
      const int K = it.Kb[g] + c / (TB / KC), kc = (c % (TB / KC)) * KC;
      const R *A = tile_ptr<R, C>(p, S.tiles, I, K);
      const R *B = tile_ptr<R, C>(p, S.tiles, K, J);
#pragma unroll
      for (int q = 0; q < PER; q++) {
        const int e = tid + q * NT;
        const int ar = e / KC, ak = e % KC;  // A: 64 rows x 16 k
        const int bk = e / TB, bc = e % TB;  // B: 16 k x 64 cols
        if constexpr (C) {
          pa[q] = reinterpret_cast<const r2 *>(A)[ar * TB + kc + ak];
          pb[q] = reinterpret_cast<const r2 *>(B)[(kc + bk) * TB + bc];
        } else {
          qa[q] = A[ar * TB + kc + ak];
          qb[q] = B[(kc + bk) * TB + bc];
        }
      }
    };
    if (wb != (int)blockIdx.x) __syncthreads();  // previous item done with LDS
    if (nch > 0) fetch(0);
    for (int c = 0; c < nch; c++) {
      const int buf = c & 1;
#pragma unroll
      for (int q = 0; q < PER; q++) {
        const int e = tid + q * NT;
        const int ar = e / KC, ak = e % KC, bk = e / TB, bc = e % TB;
        if constexpr (C) {
          As[buf][0][ak][ar] = pa[q].x;
          As[buf][NC - 1][ak][ar] = pa[q].y;
          Bs[buf][0][bk][bc] = pb[q].x;
          Bs[buf][NC - 1][bk][bc] = pb[q].y;
        } else {
          As[buf][0][ak][ar] = qa[q];
          Bs[buf][0][bk][bc] = qb[q];
        }
      }
      __syncthreads();
      if (c + 1 < nch) fetch(c + 1);
#pragma unroll
      for (int k4 = 0; k4 < KC; k4 += 4) {
        R ar[2], ai[2], br[2], bi[2];
#pragma unroll
        for (int a = 0; a < 2; a++) {
          ar[a] = As[buf][0][k4 + lk][r0 + a * 16 + li];
          if constexpr (C) ai[a] = As[buf][NC - 1][k4 + lk][r0 + a * 16 + li];
        }
#pragma unroll
        for (int bb = 0; bb < 2; bb++) {
          br[bb] = Bs[buf][0][k4 + lk][c0 + bb * 16 + li];
          if constexpr (C) bi[bb] = Bs[buf][NC - 1][k4 + lk][c0 + bb * 16 + li];
        }
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
          for (int bb = 0; bb < 2; bb++) {
            accr[a][bb] = Mfma<R>::run(ar[a], br[bb], accr[a][bb]);
            if constexpr (C) {
              accr[a][bb] = Mfma<R>::run(-ai[a], bi[bb], accr[a][bb]);
              acci[a][bb] = Mfma<R>::run(ar[a], bi[bb], acci[a][bb]);
              acci[a][bb] = Mfma<R>::run(ai[a], br[bb], acci[a][bb]);
            }
          }
      }
    }
    if (nch > 0) {
      R *Cg = tile_ptr<R, C>(p, S.tiles, I, J);
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int bb = 0; bb < 2; bb++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int r = r0 + a * 16 + Mfma<R>::row(lk, v), cc = c0 + bb * 16 + li;
            const int64_t e = (int64_t)(r * TB + cc) * NC;
            Cg[e] -= accr[a][bb][v];
            if constexpr (C) Cg[e + 1] -= acci[a][bb][v];
          }
    }
  }
}

// ---------------------------------------------------------------- panel
// UL factorisation of the diagonal tile (K,K) without pivoting, from its last
// row to its first as mat_solve_gaussian.f90:196-302 does row by row, then the
// inverses U_D^-1 and L_D^-1 into dinv(K).  The pivot test is the reference's:
// |a|^2 < 1e-20 (complex) or |a| < 1e-10 (real); the largest failing row
// (the first in elimination order) is recorded 1-based in *fail.  One
// workgroup per item.
template <class R, bool C>
__global__ __launch_bounds__(NT) void k_panel(const SysDev<R> *__restrict__ sys, Items it) {
  constexpr int NC = Elem<R, C>::NC;
  constexpr int LD = TB + 1;
  __shared__ R D[NC][TB][LD];
  __shared__ R X[NC][TB][LD];
  const int tid = threadIdx.x;
  const SysDev<R> &S = sys[it.sys[blockIdx.x]];
  const DevPlan &p = S.p;
  const int K = it.K[blockIdx.x];
  R *T = tile_ptr<R, C>(p, S.tiles, K, K);
  stage_tile<R, C, LD>(T, D);
  for (int q = TB - 1; q >= 0; --q) {
    __syncthreads();
    const R dr = D[0][q][q];
    R di = 0;
    if constexpr (C) di = D[NC - 1][q][q];
    R vr, vi = 0;
    if constexpr (C) {
      const R den = dr * dr + di * di;
      if (tid == 0 && den < R(1e-20) && K * TB + q < p.n) atomicMax(S.fail, K * TB + q + 1);
      vr = dr / den;
      vi = -di / den;
    } else {
      if (tid == 0 && fabs(dr) < R(1e-10) && K * TB + q < p.n) atomicMax(S.fail, K * TB + q + 1);
      vr = R(1) / dr;
    }
    for (int k = tid; k < q; k += NT) {  // normalise the stored row, :274-280 / :296-300
      if constexpr (C) {
        const R xr = D[0][q][k], xi = D[NC - 1][q][k];
        D[0][q][k] = xr * vr - xi * vi;
        D[NC - 1][q][k] = xi * vr + xr * vi;
      } else {
        D[0][q][k] = D[0][q][k] * vr;
      }
    }
    __syncthreads();
    for (int idx = tid; idx < q * q; idx += NT) {  // eliminate column q from the rows above
      const int r = idx / q, k = idx % q;
      if constexpr (C) {
        const R fr = D[0][r][q], fi = D[NC - 1][r][q];
        const R lr = D[0][q][k], li = D[NC - 1][q][k];
        D[0][r][k] = D[0][r][k] - fr * lr + fi * li;
        D[NC - 1][r][k] = D[NC - 1][r][k] - fr * li - fi * lr;
      } else {
        D[0][r][k] = D[0][r][k] - D[0][r][q] * D[0][q][k];
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < TB * TB; e += NT) {
    T[(int64_t)e * NC] = D[0][e / TB][e % TB];
    if constexpr (C) T[(int64_t)e * NC + 1] = D[NC - 1][e / TB][e % TB];
  }
  // X = U_D^-1 (upper, non-unit): X U = I, columns j ascending.
  for (int e = tid; e < TB * TB; e += NT) {
    X[0][e / TB][e % TB] = 0;
    if constexpr (C) X[NC - 1][e / TB][e % TB] = 0;
  }
  __syncthreads();
  for (int j = 0; j < TB; j++) {
    const R ujr = D[0][j][j];
    R uji = 0;
    if constexpr (C) uji = D[NC - 1][j][j];
    R ivr, ivi = 0;
    if constexpr (C) {
      const R den = ujr * ujr + uji * uji;
      ivr = ujr / den;
      ivi = -uji / den;
    } else {
      ivr = R(1) / ujr;
    }
    if (tid < j) {
      const int i = tid;
      R sr = 0, si = 0;
      for (int k = i; k < j; k++) {
        const R xr = X[0][i][k], ur = D[0][k][j];
        if constexpr (C) {
          const R xi = X[NC - 1][i][k], ui = D[NC - 1][k][j];
          sr += xr * ur - xi * ui;
          si += xr * ui + xi * ur;
        } else {
          sr += xr * ur;
        }
      }
      if constexpr (C) {
        X[0][i][j] = -(sr * ivr - si * ivi);
        X[NC - 1][i][j] = -(sr * ivi + si * ivr);
      } else {
        X[0][i][j] = -sr * ivr;
      }
    } else if (tid == j) {
      X[0][j][j] = ivr;
      if constexpr (C) X[NC - 1][j][j] = ivi;
    }
    __syncthreads();
  }
  R *Ui = S.dinv + (int64_t)K * 2 * TB * TB * NC;
  for (int e = tid; e < TB * TB; e += NT) {
    Ui[(int64_t)e * NC] = X[0][e / TB][e % TB];
    if constexpr (C) Ui[(int64_t)e * NC + 1] = X[NC - 1][e / TB][e % TB];
  }
  __syncthreads();
  // X = L_D^-1 (unit lower): X L = I, columns j descending.
  for (int e = tid; e < TB * TB; e += NT) {
    const int r = e / TB, c = e % TB;
    X[0][r][c] = (r == c) ? R(1) : R(0);
    if constexpr (C) X[NC - 1][r][c] = 0;
  }
  __syncthreads();
  for (int j = TB - 2; j >= 0; j--) {
    const int i = j + 1 + tid;
    if (i < TB) {
      R sr = 0, si = 0;
      for (int k = j + 1; k <= i; k++) {
        const R xr = X[0][i][k], lr = D[0][k][j];
        if constexpr (C) {
          const R xi = X[NC - 1][i][k], li = D[NC - 1][k][j];
          sr += xr * lr - xi * li;
          si += xr * li + xi * lr;
        } else {
          sr += xr * lr;
        }
      }
      X[0][i][j] = -sr;
      if constexpr (C) X[NC - 1][i][j] = -si;
    }
    __syncthreads();
  }
  R *Li = S.dinv + ((int64_t)K * 2 + 1) * TB * TB * NC;
  for (int e = tid; e < TB * TB; e += NT) {
    Li[(int64_t)e * NC] = X[0][e / TB][e % TB];
    if constexpr (C) Li[(int64_t)e * NC + 1] = X[NC - 1][e / TB][e % TB];
  }
}

// ---------------------------------------------------------------- solves
// Block triangular solves x_K = Dinv'(K) (r_K - sum_{J in S(K)} T(K,J) x_J),
// in place on X.  Kinds:
//   S_LOWER  : L x = y,   K ascending,  S = [J0(K), K),   T = tile(K,J),   Dinv' = Linv
//   S_UPPER  : U y = b,   K descending, S = (K, J1(K)],   T = tile(K,J),   Dinv' = Uinv
//   S_LOWER_T: L^T z = s, K descending, S = (K, Jx(K)],   T = tile(J,K)^T, Dinv' = Linv^T
//   S_UPPER_T: U^T l = z, K ascending,  S = [I0(K), K),   T = tile(J,K)^T, Dinv' = Uinv^T
enum SolveKind { S_LOWER = 0, S_UPPER = 1, S_LOWER_T = 2, S_UPPER_T = 3 };
constexpr int TPW = 4;  // tiles per workgroup in the partial sums

__host__ __device__ inline void solve_range(int kind, int K, int J0, int J1, int I0, int Jx, int *jb, int *je) {
  switch (kind) {
    case S_LOWER: *jb = J0; *je = K; break;
    case S_UPPER: *jb = K + 1; *je = J1 + 1; break;
    case S_LOWER_T: *jb = K + 1; *je = Jx + 1; break;
    default: *jb = I0; *je = K; break;
  }
}

// Partial sums over TPW tiles per workgroup; grid (sum of nparts, nrhs chunks).
// Partials of item g, part q, chunk c land at part + ((c*maxparts + q)*TB*RC)*NC.
template <class R, bool C>
__global__ __launch_bounds__(NT) void k_solve_partial(const SysDev<R> *__restrict__ sys, Items it, int kind,
                                                      int maxparts) {
  constexpr int NC = Elem<R, C>::NC;
  constexpr int LD = TB + 1;
  __shared__ R Ts[NC][TB][LD];
  __shared__ R Xs[NC][TB][RC + 1];
  const int g = find_item(it, blockIdx.x);
  const int q = blockIdx.x - it.off[g];
  const SysDev<R> &S = sys[it.sys[g]];
  const DevPlan &p = S.p;
  const int K = it.K[g];
  const int nrhs = S.nrhs;
  int jb, je;
  solve_range(kind, K, p.J0[K], p.J1[K], p.I0[K], p.Jx[K], &jb, &je);
  const int t0 = jb + q * TPW, t1 = min(je, t0 + TPW);
  const int rc0 = blockIdx.y * RC;
  if (rc0 >= nrhs) return;
  const int tid = threadIdx.x, r = tid >> 2, gq = tid & 3;
  const bool trans = (kind == S_LOWER_T || kind == S_UPPER_T);
  R accr[4] = {0, 0, 0, 0}, acci[4] = {0, 0, 0, 0};
  for (int J = t0; J < t1; J++) {
    const R *T = trans ? tile_ptr<R, C>(p, S.tiles, J, K) : tile_ptr<R, C>(p, S.tiles, K, J);
    __syncthreads();
    stage_tile<R, C, LD>(T, Ts);
    for (int e = tid; e < TB * RC; e += NT) {
      const int rr = e / RC, cc = e % RC;
      const int64_t gi = ((int64_t)(J * TB + rr) * nrhs + rc0 + cc) * NC;
      const bool ok = rc0 + cc < nrhs;
      Xs[0][rr][cc] = ok ? S.X[gi] : R(0);
      if constexpr (C) Xs[NC - 1][rr][cc] = ok ? S.X[gi + 1] : R(0);
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < TB; k++) {
      const R tr = trans ? Ts[0][k][r] : Ts[0][r][k];
      R ti = 0;
      if constexpr (C) ti = trans ? Ts[NC - 1][k][r] : Ts[NC - 1][r][k];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const R xr = Xs[0][k][gq * 4 + u];
        if constexpr (C) {
          const R xi = Xs[NC - 1][k][gq * 4 + u];
          accr[u] += tr * xr - ti * xi;
          acci[u] += tr * xi + ti * xr;
        } else {
          accr[u] += tr * xr;
        }
      }
    }
  }
  R *P = S.part + ((int64_t)(blockIdx.y * maxparts + q) * TB * RC) * NC;
#pragma unroll
  for (int u = 0; u < 4; u++) {
    P[(int64_t)(r * RC + gq * 4 + u) * NC] = accr[u];
    if constexpr (C) P[(int64_t)(r * RC + gq * 4 + u) * NC + 1] = acci[u];
  }
}

// x_K = Dinv'(K) (r_K - sum of partials); grid (items, nrhs chunks).
template <class R, bool C>
__global__ __launch_bounds__(NT) void k_solve_finish(const SysDev<R> *__restrict__ sys, Items it, int kind,
                                                     int maxparts) {
  constexpr int NC = Elem<R, C>::NC;
  constexpr int LD = TB + 1;
  __shared__ R Ds[NC][TB][LD];
  __shared__ R Ts[NC][TB][RC + 1];
  const SysDev<R> &S = sys[it.sys[blockIdx.x]];
  const DevPlan &p = S.p;
  const int K = it.K[blockIdx.x];
  const int nrhs = S.nrhs;
  const int rc0 = blockIdx.y * RC;
  if (rc0 >= nrhs) return;
  int jb, je;
  solve_range(kind, K, p.J0[K], p.J1[K], p.I0[K], p.Jx[K], &jb, &je);
  const int nparts = (je - jb + TPW - 1) / TPW;
  const int tid = threadIdx.x, r = tid >> 2, gq = tid & 3;
  const bool upper = (kind == S_UPPER || kind == S_UPPER_T);
  const bool trans = (kind == S_LOWER_T || kind == S_UPPER_T);
  const R *Dg = S.dinv + ((int64_t)K * 2 + (upper ? 0 : 1)) * TB * TB * NC;
  stage_tile<R, C, LD>(Dg, Ds);
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int cc = gq * 4 + u;
    const int64_t gi = ((int64_t)(K * TB + r) * nrhs + rc0 + cc) * NC;
    const bool ok = rc0 + cc < nrhs;
    R sr = ok ? S.X[gi] : R(0), si = 0;
    if constexpr (C) si = ok ? S.X[gi + 1] : R(0);
    for (int pp = 0; pp < nparts; pp++) {
      const R *P = S.part + ((int64_t)(blockIdx.y * maxparts + pp) * TB * RC) * NC;
      sr -= P[(int64_t)(r * RC + cc) * NC];
      if constexpr (C) si -= P[(int64_t)(r * RC + cc) * NC + 1];
    }
    Ts[0][r][cc] = sr;
    if constexpr (C) Ts[NC - 1][r][cc] = si;
  }
  __syncthreads();
  R accr[4] = {0, 0, 0, 0}, acci[4] = {0, 0, 0, 0};
#pragma unroll 4
  for (int k = 0; k < TB; k++) {
    const R dr = trans ? Ds[0][k][r] : Ds[0][r][k];
    R di = 0;
    if constexpr (C) di = trans ? Ds[NC - 1][k][r] : Ds[NC - 1][r][k];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const R tr = Ts[0][k][gq * 4 + u];
      if constexpr (C) {
        const R ti = Ts[NC - 1][k][gq * 4 + u];
        accr[u] += dr * tr - di * ti;
        acci[u] += dr * ti + di * tr;
      } else {
        accr[u] += dr * tr;
      }
    }
  }
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int cc = gq * 4 + u;
    if (rc0 + cc < nrhs) {
      const int64_t gi = ((int64_t)(K * TB + r) * nrhs + rc0 + cc) * NC;
      S.X[gi] = accr[u];
      if constexpr (C) S.X[gi + 1] = acci[u];
    }
  }
}

// ---------------------------------------------------------------- assembly
// Zero a buffer (graph-safe replacement for a memset node).
template <class R>
__global__ void k_fill(R *__restrict__ a, size_t count, R v) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += (size_t)gridDim.x * blockDim.x)
    a[i] = v;
}

// Scatter compressed rows (1-based columns) into the tiles; padded rows get a
// unit diagonal.  Values are interleaved (re, im) when complex.  One thread
// per row.
template <class R, bool C>
__global__ void k_scatter_csr(DevPlan p, R *__restrict__ tiles, const int64_t *__restrict__ row_ptr,
                              const int *__restrict__ col, const double *__restrict__ vals) {
  constexpr int NC = Elem<R, C>::NC;
  const int row = blockIdx.x * blockDim.x + threadIdx.x;
  if (row >= p.NB * TB) return;
  const int I = row / TB, rr = row % TB;
  if (row >= p.n) {
    R *T = tile_ptr<R, C>(p, tiles, I, I);
    T[(int64_t)(rr * TB + rr) * NC] = R(1);
    return;
  }
  for (int64_t q = row_ptr[row]; q < row_ptr[row + 1]; q++) {
    const int c = col[q] - 1;
    R *T = tile_ptr<R, C>(p, tiles, I, c / TB);
    const int64_t e = (int64_t)(rr * TB + (c % TB)) * NC;
    T[e] = R(vals[q * NC]);
    if constexpr (C) T[e + 1] = R(vals[q * NC + 1]);
  }
}

// X (precision R, rows NB*TB) <- b (FP64, n rows), zero padding rows.
template <class R>
__global__ void k_load_rhs(R *__restrict__ X, const double *__restrict__ b, size_t nb, size_t total) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < total; i += (size_t)gridDim.x * blockDim.x)
    X[i] = (i < nb) ? R(b[i]) : R(0);
}

// x (FP64, n rows) <- X (precision R).
template <class R>
__global__ void k_store_x(double *__restrict__ x, const R *__restrict__ X, size_t nb) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += (size_t)gridDim.x * blockDim.x)
    x[i] = double(X[i]);
}

}  // namespace fdmgpu
