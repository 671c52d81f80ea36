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
  int *done;  // per block: solved flag of the persistent substitution
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
// quadrant in 2x2 blocks of 16x16x4 MFMAs.  LDS holds A chunks k-major (As[k][row]) so the MFMA A operand reads
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
  // FP32 complex products use three real MFMAs (Gauss: P1 = ar br,
  // P2 = ai bi, P3 = (ar + ai)(br + bi); re = P1 - P2, im = P3 - P1 - P2);
  // the FP64 path keeps four so its rounding matches an ordinary product.
  constexpr bool M3 = C && sizeof(R) == 4;
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
    constexpr int NX = M3 ? 2 : 1;  // P2 accumulators only for the 3M form
    v4 accr[2][2], acci[2][2], accx[NX][NX];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int bb = 0; bb < 2; bb++) {
        accr[a][bb] = v4{0, 0, 0, 0};
        acci[a][bb] = v4{0, 0, 0, 0};
      }
#pragma unroll
    for (int a = 0; a < NX; a++)
#pragma unroll
      for (int bb = 0; bb < NX; bb++) accx[a][bb] = v4{0, 0, 0, 0};
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
        if constexpr (M3) {
          R as[2], bs[2];
#pragma unroll
          for (int a = 0; a < 2; a++) as[a] = ar[a] + ai[a];
#pragma unroll
          for (int bb = 0; bb < 2; bb++) bs[bb] = br[bb] + bi[bb];
#pragma unroll
          for (int a = 0; a < 2; a++)
#pragma unroll
            for (int bb = 0; bb < 2; bb++) {
              accr[a][bb] = Mfma<R>::run(ar[a], br[bb], accr[a][bb]);   // P1
              accx[a % NX][bb % NX] = Mfma<R>::run(ai[a], bi[bb], accx[a % NX][bb % NX]);  // P2
              acci[a][bb] = Mfma<R>::run(as[a], bs[bb], acci[a][bb]);   // P3
            }
        } else {
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
    }
    if constexpr (M3) {
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int bb = 0; bb < 2; bb++) {
          acci[a][bb] = acci[a][bb] - accr[a][bb] - accx[a % NX][bb % NX];
          accr[a][bb] = accr[a][bb] - accx[a % NX][bb % NX];
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
//
// Elimination step q (one barrier): every row r < q subtracts
// (a_rq / a_qq) * row q from its columns k < q; thread (r = tid/4) owns row r
// and the columns k = tid%4 (mod 4).  The normalised row q (the stored L row,
// a_qk / a_qq) is kept in a register and written back during step q-1, when
// row q is no longer read.  The inverses are column-parallel substitutions,
// four lanes per column combining partial sums by lane shuffles.
template <class R>
__device__ __forceinline__ R lane_sum4(R v) {
  v += __shfl_xor(v, 1);
  v += __shfl_xor(v, 2);
  return v;
}

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
  __syncthreads();
  const int myr = tid >> 2, part = tid & 3;
  int pend = -1;
  R pr = 0, pi = 0;
  for (int q = TB - 1; q >= 0; --q) {
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
    if (pend >= 0) {  // normalised row q+1, element k = tid
      D[0][pend][tid] = pr;
      if constexpr (C) D[NC - 1][pend][tid] = pi;
    }
    pend = -1;
    if (tid < q) {  // normalise the stored row, :274-280 / :296-300
      const R xr = D[0][q][tid];
      if constexpr (C) {
        const R xi = D[NC - 1][q][tid];
        pr = xr * vr - xi * vi;
        pi = xi * vr + xr * vi;
      } else {
        pr = xr * vr;
      }
      pend = q;
    }
    if (myr < q) {  // eliminate column q from row myr
      R fr = D[0][myr][q], fi = 0;
      if constexpr (C) {
        fi = D[NC - 1][myr][q];
        const R t = fr * vr - fi * vi;
        fi = fr * vi + fi * vr;
        fr = t;
      } else {
        fr = fr * vr;
      }
      for (int k = part; k < q; k += 4) {
        const R lr = D[0][q][k];
        if constexpr (C) {
          const R li = D[NC - 1][q][k];
          D[0][myr][k] = D[0][myr][k] - fr * lr + fi * li;
          D[NC - 1][myr][k] = D[NC - 1][myr][k] - fr * li - fi * lr;
        } else {
          D[0][myr][k] = D[0][myr][k] - fr * lr;
        }
      }
    }
    __syncthreads();
  }
  for (int e = tid; e < TB * TB; e += NT) {
    T[(int64_t)e * NC] = D[0][e / TB][e % TB];
    if constexpr (C) T[(int64_t)e * NC + 1] = D[NC - 1][e / TB][e % TB];
  }
  const int j = tid >> 2;
  // X = U_D^-1: U X = I, column j from its last row up.
  for (int i = TB - 1; i >= 0; --i) {
    if (i <= j) {
      R sr = 0, si = 0;
      for (int k = i + 1 + part; k <= j; k += 4) {
        const R ur = D[0][i][k], xr = X[0][k][j];
        if constexpr (C) {
          const R ui = D[NC - 1][i][k], xi = X[NC - 1][k][j];
          sr += ur * xr - ui * xi;
          si += ur * xi + ui * xr;
        } else {
          sr += ur * xr;
        }
      }
      sr = lane_sum4(sr);
      if constexpr (C) si = lane_sum4(si);
      sr = (i == j ? R(1) : R(0)) - sr;
      si = -si;
      const R ur = D[0][i][i];
      if constexpr (C) {
        const R ui = D[NC - 1][i][i], den = ur * ur + ui * ui;
        const R ivr = ur / den, ivi = -ui / den;
        X[0][i][j] = sr * ivr - si * ivi;
        X[NC - 1][i][j] = sr * ivi + si * ivr;
      } else {
        X[0][i][j] = sr / ur;
      }
    } else if (part == 0) {
      X[0][i][j] = 0;
      if constexpr (C) X[NC - 1][i][j] = 0;
    }
  }
  __syncthreads();
  R *Ui = S.dinv + (int64_t)K * 2 * TB * TB * NC;
  for (int e = tid; e < TB * TB; e += NT) {
    Ui[(int64_t)e * NC] = X[0][e / TB][e % TB];
    if constexpr (C) Ui[(int64_t)e * NC + 1] = X[NC - 1][e / TB][e % TB];
  }
  __syncthreads();
  // X = L_D^-1 (unit lower): L X = I, column j from its diagonal down.
  for (int i = 0; i < TB; ++i) {
    if (i >= j) {
      R sr = 0, si = 0;
      for (int k = j + part; k < i; k += 4) {
        const R lr = D[0][i][k], xr = X[0][k][j];
        if constexpr (C) {
          const R li = D[NC - 1][i][k], xi = X[NC - 1][k][j];
          sr += lr * xr - li * xi;
          si += lr * xi + li * xr;
        } else {
          sr += lr * xr;
        }
      }
      sr = lane_sum4(sr);
      if constexpr (C) si = lane_sum4(si);
      X[0][i][j] = (i == j ? R(1) : R(0)) - sr;
      if constexpr (C) X[NC - 1][i][j] = -si;
    } else if (part == 0) {
      X[0][i][j] = 0;
      if constexpr (C) X[NC - 1][i][j] = 0;
    }
  }
  __syncthreads();
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

// Persistent block triangular solve: one launch runs every block step of
// every listed system.  Task t (in dependency order: block steps aligned at
// their end across systems, as the factorisation's launches are) is taken
// from a ticket counter, so a task only ever waits on tasks with smaller
// tickets, which running workgroups hold: progress needs no scheduling
// assumption.  A task accumulates T(K,J) x_J over J in S(K), waiting on each
// J's done flag (agent-scope acquire), then writes x_K and publishes its own
// flag (agent-scope release, so other XCDs' caches see x_K).  Waits are
// bounded: past SPIN_LIMIT polls the task records an error and goes on, so
// the kernel always terminates.
constexpr int SPIN_LIMIT = 1 << 22;

// Systems of one persistent solve: system sys[i] runs its NB[i] block steps
// at global steps off[i] .. off[i] + NB[i] - 1 (aligned at the end).  Tasks
// are numbered step by step, systems in list order within a step.
struct SolveSet {
  int count;
  int maxNB;
  int sys[MAXG];
  int off[MAXG];
  int nb[MAXG];
};

__device__ __forceinline__ int tasks_before(const SolveSet &ss, int step) {
  int t = 0;
  for (int i = 0; i < ss.count; i++) t += max(0, min(step, ss.off[i] + ss.nb[i]) - ss.off[i]);
  return t;
}

__device__ __forceinline__ void wait_flag(const int *flag, int *err) {
  int n = 0;
  while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) == 0) {
    if (__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) break;
    if (++n > SPIN_LIMIT) {
      atomicExch(err, 1);
      break;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

template <class R, bool C>
__global__ __launch_bounds__(NT) void k_solve_persist(const SysDev<R> *__restrict__ sys, SolveSet ss, int kind,
                                                      int *counter, int *err) {
  constexpr int NC = Elem<R, C>::NC;
  constexpr int LD = TB + 1;
  __shared__ R Ts[NC][TB][LD];
  __shared__ R Xs[NC][TB][RC + 1];
  __shared__ int tk;
  const int tid = threadIdx.x, r = tid >> 2, gq = tid & 3;
  const bool trans = (kind == S_LOWER_T || kind == S_UPPER_T);
  const bool upper = (kind == S_UPPER || kind == S_UPPER_T);
  const bool asc = (kind == S_LOWER || kind == S_UPPER_T);
  const int ntasks = tasks_before(ss, ss.maxNB);
  for (;;) {
    __syncthreads();
    if (tid == 0) tk = atomicAdd(counter, 1);
    __syncthreads();
    const int t = tk;
    if (t >= ntasks) break;
    // task -> (step, system): last step whose first task is <= t
    int lo = 0, hi = ss.maxNB - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (tasks_before(ss, mid) <= t) lo = mid;
      else hi = mid - 1;
    }
    int k = t - tasks_before(ss, lo), i = 0;
    for (; i < ss.count; i++) {
      if (lo >= ss.off[i] && lo < ss.off[i] + ss.nb[i]) {
        if (k == 0) break;
        k--;
      }
    }
    const int g = ss.sys[i];
    const int tg = lo - ss.off[i];
    const int K = asc ? tg : ss.nb[i] - 1 - tg;
    const SysDev<R> &S = sys[g];
    const DevPlan &p = S.p;
    const int nrhs = S.nrhs;
    int jb, je;
    solve_range(kind, K, p.J0[K], p.J1[K], p.I0[K], p.Jx[K], &jb, &je);
    for (int rc0 = 0; rc0 < nrhs; rc0 += RC) {
      R accr[4] = {0, 0, 0, 0}, acci[4] = {0, 0, 0, 0};
      // oldest dependencies first: the last one is the neighbour block
      for (int q = 0; q < je - jb; q++) {
        const int J = asc ? jb + q : je - 1 - q;
        if (tid == 0) wait_flag(&S.done[J], err);
        __syncthreads();
        const R *T = trans ? tile_ptr<R, C>(p, S.tiles, J, K) : tile_ptr<R, C>(p, S.tiles, K, J);
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
        __syncthreads();
      }
      // x_K = Dinv'(K) (r_K - acc)
      const R *Dg = S.dinv + ((int64_t)K * 2 + (upper ? 0 : 1)) * TB * TB * NC;
      stage_tile<R, C, LD>(Dg, Ts);
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int cc = gq * 4 + u;
        const int64_t gi = ((int64_t)(K * TB + r) * nrhs + rc0 + cc) * NC;
        const bool ok = rc0 + cc < nrhs;
        R sr = ok ? S.X[gi] : R(0), si = 0;
        if constexpr (C) si = ok ? S.X[gi + 1] : R(0);
        Xs[0][r][cc] = sr - accr[u];
        if constexpr (C) Xs[NC - 1][r][cc] = si - acci[u];
      }
      __syncthreads();
      R yr[4] = {0, 0, 0, 0}, yi[4] = {0, 0, 0, 0};
#pragma unroll 4
      for (int k = 0; k < TB; k++) {
        const R dr = trans ? Ts[0][k][r] : Ts[0][r][k];
        R di = 0;
        if constexpr (C) di = trans ? Ts[NC - 1][k][r] : Ts[NC - 1][r][k];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          const R tr = Xs[0][k][gq * 4 + u];
          if constexpr (C) {
            const R ti = Xs[NC - 1][k][gq * 4 + u];
            yr[u] += dr * tr - di * ti;
            yi[u] += dr * ti + di * tr;
          } else {
            yr[u] += dr * tr;
          }
        }
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int cc = gq * 4 + u;
        if (rc0 + cc < nrhs) {
          const int64_t gi = ((int64_t)(K * TB + r) * nrhs + rc0 + cc) * NC;
          S.X[gi] = yr[u];
          if constexpr (C) S.X[gi + 1] = yi[u];
        }
      }
      __syncthreads();
    }
    // publish x_K to every XCD, then the flag
    __threadfence();
    __syncthreads();
    if (tid == 0) __hip_atomic_store(&S.done[K], 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
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
