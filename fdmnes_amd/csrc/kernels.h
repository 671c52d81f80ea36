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
// dinv(K) holds DSLOTS tiles per diagonal block: U_D^-1, L_D^-1, and the
// neighbour products MU(K) = U_D^-1 U(K,K+1) and ML(K) = L_D^-1 L(K,K-1) the
// substitutions use (x_K = y_K - M x_neighbour).
//
// Every kernel is grouped: one launch serves the same step of up to MAXG
// independent systems (energy points, clusters of different size).  A launch
// carries an Items list (system, block step, first workgroup); workgroup b
// serves the item g with off[g] <= b < off[g+1].
#pragma once
#include <hip/hip_runtime.h>

#include <type_traits>
#include <stdint.h>

namespace fdmgpu {

constexpr int TB = 64;    // tile edge = panel width
constexpr int NT = 256;   // threads per workgroup
constexpr int RC = 16;    // right-hand sides per solve chunk
constexpr int MAXG = 32;  // systems per grouped launch

// Device pointers held in structs are generic (flat) to the compiler, and
// flat loads count in lgkmcnt as well as vmcnt, so a wait for LDS results
// would also wait for every global load in flight.  Kernels access them
// through global-address-space views (global_* instructions, vmcnt only).
template <class T>
using gptr = __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ gptr<T> G(T *p) {
  return (gptr<T>)p;
}
// A pointer the caller knows to be the same in every lane, moved to scalar
// registers: loads then take it as their scalar base plus a 32-bit lane
// offset instead of a 64-bit address formed in every lane.
template <class T>
__device__ __forceinline__ gptr<T> uniform(gptr<T> p) {
  const uint64_t v = (uint64_t)p;
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return (gptr<T>)(((uint64_t)hi << 32) | lo);
}
// Element at byte offset off (32-bit, unsigned) from a scalar base: the form
// global loads encode directly as base register pair + lane offset.
template <class T, class P>
__device__ __forceinline__ T load_at(gptr<const P> base, uint32_t off) {
  return *(gptr<const T>)((gptr<const char>)base + off);
}

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

constexpr int DSLOTS = 4;
enum DinvSlot { D_UINV = 0, D_LINV = 1, D_MU = 2, D_ML = 3 };

template <class R, bool C> struct Elem;
template <class R> struct Elem<R, true> { static constexpr int NC = 2; };
template <class R> struct Elem<R, false> { static constexpr int NC = 1; };

template <class R, bool C>
__device__ __forceinline__ gptr<R> tile_ptr(const DevPlan &p, R *tiles, int I, int J) {
  return G(tiles) + (G(p.tbase)[I] + (J - G(p.J0)[I])) * (int64_t)(TB * TB * Elem<R, C>::NC);
}

template <class R, bool C>
__device__ __forceinline__ gptr<R> dinv_tile(const SysDev<R> &S, int K, int slot) {
  return G(S.dinv) + ((int64_t)K * DSLOTS + slot) * (int64_t)(TB * TB * Elem<R, C>::NC);
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

enum GemmMode { G_UPDATE = 0, G_TRSM = 1, G_NEIGH = 2 };

// Stage one TB x TB tile (global, interleaved) into split LDS planes.
// Stage the ROWS x COLS block at (row0, col0) of one TB x TB tile (global,
// interleaved) into split LDS planes.
template <class R, bool C, int LD, int ROWS, int COLS, class P>
__device__ __forceinline__ void stage_block(P g, int row0, int col0, R (*s)[ROWS][LD]) {
  static_assert(ROWS * COLS % NT == 0, "whole rounds of the workgroup");
  const int tid = threadIdx.x;
#pragma unroll
  for (int q = 0; q < ROWS * COLS / NT; q++) {
    const int e = tid + q * NT, r = e / COLS, c = e % COLS;
    const int64_t o = (int64_t)(row0 + r) * TB + col0 + c;
    if constexpr (C) {
      typedef R r2 __attribute__((ext_vector_type(2)));
      const r2 v = ((gptr<const r2>)g)[o];
      s[0][r][c] = v.x;
      s[1][r][c] = v.y;
    } else {
      s[0][r][c] = g[o];
    }
  }
}

template <class R, bool C, int LD, class P>
__device__ __forceinline__ void stage_tile(P g, R (*s)[TB][LD]) {
  const int tid = threadIdx.x;
  if constexpr (C) {
    typedef R r2 __attribute__((ext_vector_type(2)));
    const gptr<const r2> g2 = (gptr<const r2>)g;
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
//   G_UPDATE: tile(I,J) -= tile(I,K) * tile(K,J)   per item nI * nJ workgroups
//   G_TRSM:   the products of block step K once the panel has formed U_D^-1
//             and L_D^-1, per item nJ + nI workgroups:
//     tile(K,J) = Uinv(K) * tile(K,J), J0(K) <= J < K      (L row)
//     tile(I,K) = tile(I,K) * Linv(K), I0(K) <= I < K      (U column)
//   G_NEIGH:  after the last block step, the substitutions' neighbour
//             products of every block K, per item 2 NB workgroups:
//     MU(K) = Uinv(K) * U(K,K+1), ML(K) = Linv(K) * L(K,K-1) (where the tiles exist)
// Four waves; wave w owns the 32x32 quadrant (w/2, w%2), 2x2 MFMA blocks of
// 16x16.  The k dimension is staged in two halves (half the LDS).
template <class R, bool C>
__global__ __launch_bounds__(NT) void k_tile_gemm(const SysDev<R> *__restrict__ sys, Items it, int mode) {
  constexpr int NC = Elem<R, C>::NC;
  constexpr int KH = TB / 2;  // k staged in two halves: A columns / B rows kh .. kh+31
  constexpr int LDA = KH + 1, LDB = TB + 1;
  __shared__ R As[NC][TB][LDA];
  __shared__ R Bs[NC][KH][LDB];
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
  gptr<const R> A;
  gptr<const R> B;
  gptr<R> Cg;
  if (mode == G_UPDATE) {
    const int I = I0 + b / nJ, J = J0 + b % nJ;
    A = tile_ptr<R, C>(p, S.tiles, I, K);
    B = tile_ptr<R, C>(p, S.tiles, K, J);
    Cg = tile_ptr<R, C>(p, S.tiles, I, J);
  } else if (mode == G_TRSM) {
    if (b < nJ) {
      Cg = tile_ptr<R, C>(p, S.tiles, K, J0 + b);
      A = dinv_tile<R, C>(S, K, D_UINV);
      B = Cg;
    } else {
      Cg = tile_ptr<R, C>(p, S.tiles, I0 + (b - nJ), K);
      A = Cg;
      B = dinv_tile<R, C>(S, K, D_LINV);
    }
  } else {
    const int Kn = b >> 1;
    if (b & 1) {
      if (Kn - 1 < p.J0[Kn]) continue;  // no tile (K, K-1): block row K starts at K
      A = dinv_tile<R, C>(S, Kn, D_LINV);
      B = tile_ptr<R, C>(p, S.tiles, Kn, Kn - 1);
      Cg = dinv_tile<R, C>(S, Kn, D_ML);
    } else {
      if (Kn + 1 > p.J1[Kn]) continue;  // no tile (K, K+1)
      A = dinv_tile<R, C>(S, Kn, D_UINV);
      B = tile_ptr<R, C>(p, S.tiles, Kn, Kn + 1);
      Cg = dinv_tile<R, C>(S, Kn, D_MU);
    }
  }
  typedef typename Mfma<R>::v4 v4;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r0 = (wave >> 1) * 32, c0 = (wave & 1) * 32;
  const int li = lane & 15, lk = lane >> 4;
  // FP32 complex products use three real MFMAs (P1 = ar br, P2 = ai bi,
  // P3 = (ar + ai)(br + bi); re = P1 - P2, im = P3 - P1 - P2); FP64 keeps four
  constexpr bool M3 = C && sizeof(R) == 4;
  v4 accr[2][2], acci[2][2], accx[M3 ? 2 : 1][M3 ? 2 : 1];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int bb = 0; bb < 2; bb++) {
      accr[a][bb] = v4{0, 0, 0, 0};
      acci[a][bb] = v4{0, 0, 0, 0};
      if constexpr (M3) accx[a][bb] = v4{0, 0, 0, 0};
    }
#pragma unroll 1
  for (int kh = 0; kh < TB; kh += KH) {
    if (kh) __syncthreads();  // every read of the first halves done
    stage_block<R, C, LDA, TB, KH>(A, 0, kh, As);
    stage_block<R, C, LDB, KH, TB>(B, kh, 0, Bs);
    __syncthreads();
#pragma unroll 4
    for (int k0 = 0; k0 < KH; k0 += 4) {
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
          if constexpr (M3) {
            accr[a][bb] = Mfma<R>::run(ar[a], br[bb], accr[a][bb]);
            accx[a][bb] = Mfma<R>::run(ai[a], bi[bb], accx[a][bb]);
            acci[a][bb] = Mfma<R>::run(ar[a] + ai[a], br[bb] + bi[bb], acci[a][bb]);
          } else {
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
        acci[a][bb] = acci[a][bb] - accr[a][bb] - accx[a][bb];
        accr[a][bb] = accr[a][bb] - accx[a][bb];
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
constexpr int MAXGK = 32;  // block steps per trailing-update group
struct TrailItems {
  int count;
  int sys[MAXT];
  int Kb[MAXT];
  int nk[MAXT];
  int i0[MAXT];
  int j0[MAXT];
  int nj[MAXT];
  int ni[MAXT];
  int off[MAXT + 1];   // tile offsets (one tile per workgroup)
  int moff[MAXT + 1];  // offsets of 1 x 2 tile blocks (two tiles per workgroup)
};

__device__ __forceinline__ int find_trail(const TrailItems &it, int b) {
  int g = 0;
  while (g + 1 < it.count && b >= it.off[g + 1]) g++;
  return g;
}

constexpr int KC = 16;  // k-chunk of the LDS pipeline

// tile(I,J) -= sum_K tile(I,K) tile(K,J) over the work items of all
// systems, either by a persistent grid striding over the output blocks or
// (grid at least the block count) by one workgroup per output block; the
// blocks a workgroup takes in turn get the next block's first chunks loaded
// during the current one's last.  A workgroup updates TJ consecutive tiles of a block
// row (a 64 x 64*TJ output: TJ = 1 or 2); the K dimension (m steps of 64)
// streams through a two-buffer LDS pipeline in chunks of KC, the next chunks'
// global loads in flight while the current chunk feeds the MFMAs.  Four waves,
// each 32 rows x 32*TJ columns in 2 x 2*TJ blocks of 16x16x4 MFMAs.  LDS holds
// A chunks k-major (As[k][row]) so the MFMA A operand reads 16 consecutive
// rows per k.  With TJ = 2 every A chunk feeds twice the MFMA work, the
// barriers per multiply-add halve, and the bytes loaded per multiply-add drop
// by a quarter; at an item's last odd column the second tile is absent and its
// B columns load as zero.  The steps of the range reaching a tile are a prefix
// (I0, J0 non-decreasing in K); the workgroup runs the longest prefix of its
// tiles, and B columns of a tile (K, J) that does not exist (J0(K) > J) load
// as zero, so every tile receives exactly its own steps.
//
// Work order: dispatch places workgroup w on XCD w % NXCD, and each XCD has
// its own L2.  The work list is cut into NXCD contiguous ranges, one per XCD
// (the grid is a multiple of NXCD), and inside an item the output blocks run
// in bands of GM block rows, column by column, so the workgroups resident on
// one XCD at a time cover a compact patch and share its U row tiles and L
// column tiles in their L2 instead of each fetching them again.
constexpr int NXCD = 8;
constexpr int GM = 8;
template <class R, bool C, int TJ>
__global__ __launch_bounds__(NT, (sizeof(R) == 4 && TJ == 1) ? 3 : 2) void k_trail(const SysDev<R> *__restrict__ sys, TrailItems it) {
  static_assert(TJ == 1 || TJ == 2, "one or two tiles per workgroup");
  constexpr int NC = Elem<R, C>::NC;
  constexpr int W = TB * TJ;  // output columns per workgroup
  constexpr int NBJ = 2 * TJ;  // 16-column MFMA blocks per wave
  // FP32: A rows of 80 words, so the four k rows an MFMA operand read touches
  // (lanes 16*lk .. 16*lk+15 read row k4+lk) fall in four disjoint groups of
  // 16 banks; A is stored with its row index XOR-ed by sw(k) (a multiple of 4
  // below 16), so the staging writes, 16 k values x 4 rows per wave, also hit
  // 64 distinct banks.  B: rows of 80 words (TJ = 1), or of 128 words with
  // the column XOR-ed by swb(k) = 16*(k%4) (TJ = 2, which keeps three
  // workgroups' LDS within a CU).  FP64 keeps rows of 65 elements.
  constexpr bool F32 = sizeof(R) == 4;
  constexpr int LDA = F32 ? TB + 16 : TB + 1;
  constexpr int LDB = F32 ? (TJ == 1 ? TB + 16 : W) : W + 1;
  constexpr int PERA = TB * KC / NT;  // A elements per thread per chunk (4)
  constexpr int PERB = W * KC / NT;   // B elements per thread per chunk (4 * TJ)
  auto sw = [](int k) { return F32 ? ((k & 15) >> 2) << 2 : 0; };
  auto swb = [](int k) { return (F32 && TJ == 2) ? (k & 3) << 4 : 0; };
  constexpr int64_t TE = (int64_t)TB * TB * NC;  // elements per tile
  __shared__ R As[2][NC][KC][LDA];
  __shared__ R Bs[2][NC][KC][LDB];
  // per step t of the current item (K = Kb + t): I0(K), J0(K), tbase(K),
  // loaded once per item; the tiles of the item then need no global metadata
  // beyond tbase(I), J0(I) of their row
  __shared__ int kI0[MAXGK], kJ0[MAXGK];
  __shared__ int64_t ktb[MAXGK];
  typedef typename Mfma<R>::v4 v4;
  typedef R r2 __attribute__((ext_vector_type(2)));
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int r0 = (wave >> 1) * 32, c0 = (wave & 1) * 32 * TJ;
  const int li = lane & 15, lk = lane >> 4;
  // FP32 complex products use three real MFMAs (Gauss: P1 = ar br,
  // P2 = ai bi, P3 = (ar + ai)(br + bi); re = P1 - P2, im = P3 - P1 - P2);
  // the FP64 path keeps four so its rounding matches an ordinary product.
  constexpr bool M3 = C && sizeof(R) == 4;
  const int total = TJ == 1 ? it.off[it.count] : it.moff[it.count];
  const int per = (total + NXCD - 1) / NXCD, xcd = blockIdx.x % NXCD, lanes = gridDim.x / NXCD;
  // output block named by work index wb: item g, block row I, first tile
  // column J, tiles present in the item; rowA = position of tile (I, Kb)
  struct Blk {
    int g, I, J, cols_in;
    int64_t rowA;
  };
  auto decode = [&](int wb) {
    Blk k;
    int g = 0;
    if constexpr (TJ == 1) g = find_trail(it, wb);
    else
      while (g + 1 < it.count && wb >= it.moff[g + 1]) g++;
    const int b = wb - (TJ == 1 ? it.off[g] : it.moff[g]);
    const int nj = it.nj[g], ni = it.ni[g], mj = (nj + TJ - 1) / TJ;
    const int band = b / (GM * mj), hr = min(GM, ni - band * GM), q = b - band * GM * mj;
    k.g = g;
    k.I = it.i0[g] + band * GM + q % hr;
    k.J = it.j0[g] + TJ * (q / hr);
    k.cols_in = min(TJ, it.j0[g] + nj - k.J);
    const DevPlan &pk = sys[it.sys[g]].p;
    k.rowA = G(pk.tbase)[k.I] + (it.Kb[g] - G(pk.J0)[k.I]);
    return k;
  };
  // DEPTH register sets keep that many chunks' loads in flight while the
  // MFMAs consume the LDS copy of the current one (two for FP32 with one tile
  // per workgroup, whose register budget allows it); they live across blocks
  // because the next block's first chunks load during this block's epilogue
  constexpr int DEPTH = (sizeof(R) == 4 && TJ == 1) ? 2 : 1;
  r2 pa[DEPTH][PERA], pb[DEPTH][PERB];
  R qa[DEPTH][PERA], qb[DEPTH][PERB];
  int jw = blockIdx.x / NXCD;
  if (jw >= per || xcd * per + jw >= total) return;
  Blk cur = decode(xcd * per + jw);
  int gcur = -1;
  bool pre = false;  // the first chunks of cur are already in the prefetch registers
  for (bool first = true;; first = false) {
    const int g = cur.g, I = cur.I, Jc = cur.J, cin = cur.cols_in;
    const SysDev<R> &S = sys[it.sys[g]];
    const DevPlan &p = S.p;
    const int Kb = it.Kb[g], nk = it.nk[g];
    if (!first) __syncthreads();  // previous block done with LDS
    if (g != gcur) {
      if (tid < nk) {
        kI0[tid] = G(p.I0)[Kb + tid];
        kJ0[tid] = G(p.J0)[Kb + tid];
        ktb[tid] = G(p.tbase)[Kb + tid];
      }
      __syncthreads();
      gcur = g;
    }
    // steps of the range reaching the last tile of a block: a prefix, counted by a ballot
    auto steps = [&](int Ib, int Jb, int cb) {
      const bool reach = lane < nk && kI0[lane] <= Ib && kJ0[lane] <= Jb + cb - 1;
      const uint64_t rb = __ballot(reach);
      return ~rb == 0 ? 64 : __builtin_ctzll(~rb);
    };
    const int m = steps(I, Jc, cin);
    const int nch = m * (TB / KC);
    const gptr<const R> tiles = G(S.tiles);
    const gptr<const R> Ab = tiles + cur.rowA * TE;
    // the next block of this workgroup; when it belongs to the same item (its
    // step metadata already in LDS) its first chunks load during this block's
    // last chunks
    const int jn = jw + lanes;
    const bool has_next = jn < per && xcd * per + jn < total;
    Blk nxt{};
    if (has_next) nxt = decode(xcd * per + jn);
    const bool chain = has_next && nxt.g == g && steps(nxt.I, nxt.J, nxt.cols_in) > 0;
    const gptr<const R> An = tiles + nxt.rowA * TE;
    constexpr int NX = M3 ? NBJ : 1;  // P2 accumulators only for the 3M form
    v4 accr[2][NBJ], acci[2][NBJ], accx[M3 ? 2 : 1][NX];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int bb = 0; bb < NBJ; bb++) {
        accr[a][bb] = v4{0, 0, 0, 0};
        acci[a][bb] = v4{0, 0, 0, 0};
      }
#pragma unroll
    for (int a = 0; a < (M3 ? 2 : 1); a++)
#pragma unroll
      for (int bb = 0; bb < NX; bb++) accx[a][bb] = v4{0, 0, 0, 0};
    // chunk c: step K = Kb + c/4, columns kc..kc+15 of A = tile(I,K), rows of
    // B = tile(K,J..J+TJ-1), loaded into the register sets declared above.
    // src() forms the addresses (reading the step metadata in LDS), load()
    // issues the global loads, so the loads can follow a barrier with no LDS
    // wait in between.
    struct Src {
      gptr<const R> A, B[TJ];
      uint32_t kc;  // unsigned lane offsets: loads take them as 32-bit offsets to the scalar base
      bool ok[TJ];
    };
    // chunk c of the block whose tile (I, Kb) is at A0, output tiles J .. J + cols_in - 1
    auto src = [&](gptr<const R> A0, int J, int cols_in, int c) {
      Src x;
      const int st = c / (TB / KC);
      x.kc = (uint32_t)(c % (TB / KC)) * KC;
      x.A = uniform(A0 + st * TE);
#pragma unroll
      for (int tb = 0; tb < TJ; tb++) {
        x.ok[tb] = TJ == 1 || (tb < cols_in && kJ0[st] <= J + tb);
        x.B[tb] = uniform(tiles + (ktb[st] + (J + tb - kJ0[st])) * TE);
      }
      return x;
    };
    auto load = [&](const Src &x, auto set) {
      constexpr int s = decltype(set)::value;
#pragma unroll
      for (int q = 0; q < PERA; q++) {
        const int e = tid + q * NT;
        const uint32_t ar = e / KC, ak = e % KC;  // A: 64 rows x 16 k
        const uint32_t off = (ar * TB + x.kc + ak) * (uint32_t)(NC * sizeof(R));
        if constexpr (C) pa[s][q] = load_at<r2>(x.A, off);
        else qa[s][q] = load_at<R>(x.A, off);
      }
#pragma unroll
      for (int q = 0; q < PERB; q++) {
        const int e = tid + q * NT;
        const uint32_t bk = e / W, bc = e % W, tb = bc / TB;  // B: 16 k x W cols, tile J + tb
        const uint32_t off = ((x.kc + bk) * TB + bc % TB) * (uint32_t)(NC * sizeof(R));
        if constexpr (C) pb[s][q] = x.ok[tb] ? load_at<r2>(x.B[tb], off) : r2{0, 0};
        else qb[s][q] = x.ok[tb] ? load_at<R>(x.B[tb], off) : R(0);
      }
    };
    // the chunk loaded DEPTH chunks ahead: this block's chunk, else the next
    // block's first chunks (chain), else this block's last chunk again
    // (unused).  Every path issues the same loads, so the wait before a
    // chunk's LDS staging counts only the loads issued before it and leaves
    // the later set in flight.
    auto ahead = [&](int c) {
      const int cn = c + DEPTH;
      const bool own = cn < nch, nb = !own && chain;
      return src(own || !nb ? Ab : An, own || !nb ? Jc : nxt.J, own || !nb ? cin : nxt.cols_in,
                 own ? cn : (nb ? cn - nch : nch - 1));
    };
    auto step = [&](int c, auto set) {
      constexpr int s = decltype(set)::value;
      const int buf = c & 1;
      const Src nx = ahead(c);
#pragma unroll
      for (int q = 0; q < PERA; q++) {
        const int e = tid + q * NT;
        const int ar = e / KC, ak = e % KC;
        if constexpr (C) {
          As[buf][0][ak][ar ^ sw(ak)] = pa[s][q].x;
          As[buf][NC - 1][ak][ar ^ sw(ak)] = pa[s][q].y;
        } else {
          As[buf][0][ak][ar ^ sw(ak)] = qa[s][q];
        }
      }
#pragma unroll
      for (int q = 0; q < PERB; q++) {
        const int e = tid + q * NT;
        const int bk = e / W, bc = e % W;
        if constexpr (C) {
          Bs[buf][0][bk][bc ^ swb(bk)] = pb[s][q].x;
          Bs[buf][NC - 1][bk][bc ^ swb(bk)] = pb[s][q].y;
        } else {
          Bs[buf][0][bk][bc ^ swb(bk)] = qb[s][q];
        }
      }
      __syncthreads();
      load(nx, set);
      __builtin_amdgcn_s_setprio(1);
#pragma unroll
      for (int k4 = 0; k4 < KC; k4 += 4) {
        const int k = k4 + lk;
        R ar[2], ai[2], br[NBJ], bi[NBJ];
#pragma unroll
        for (int a = 0; a < 2; a++) {
          ar[a] = As[buf][0][k][r0 + a * 16 + (li ^ sw(k))];
          if constexpr (C) ai[a] = As[buf][NC - 1][k][r0 + a * 16 + (li ^ sw(k))];
        }
#pragma unroll
        for (int bb = 0; bb < NBJ; bb++) {
          br[bb] = Bs[buf][0][k][(c0 + bb * 16 + li) ^ swb(k)];
          if constexpr (C) bi[bb] = Bs[buf][NC - 1][k][(c0 + bb * 16 + li) ^ swb(k)];
        }
        if constexpr (M3) {
          R as[2], bs[NBJ];
#pragma unroll
          for (int a = 0; a < 2; a++) as[a] = ar[a] + ai[a];
#pragma unroll
          for (int bb = 0; bb < NBJ; bb++) bs[bb] = br[bb] + bi[bb];
#pragma unroll
          for (int a = 0; a < 2; a++)
#pragma unroll
            for (int bb = 0; bb < NBJ; bb++) {
              accr[a][bb] = Mfma<R>::run(ar[a], br[bb], accr[a][bb]);                       // P1
              accx[a % 2][bb % NX] = Mfma<R>::run(ai[a], bi[bb], accx[a % 2][bb % NX]);  // P2
              acci[a][bb] = Mfma<R>::run(as[a], bs[bb], acci[a][bb]);                       // P3
            }
        } else {
#pragma unroll
          for (int a = 0; a < 2; a++)
#pragma unroll
            for (int bb = 0; bb < NBJ; bb++) {
              accr[a][bb] = Mfma<R>::run(ar[a], br[bb], accr[a][bb]);
              if constexpr (C) {
                accr[a][bb] = Mfma<R>::run(-ai[a], bi[bb], accr[a][bb]);
                acci[a][bb] = Mfma<R>::run(ar[a], bi[bb], acci[a][bb]);
                acci[a][bb] = Mfma<R>::run(ai[a], br[bb], acci[a][bb]);
              }
            }
        }
      }
      __builtin_amdgcn_s_setprio(0);
    };
    using S0 = std::integral_constant<int, 0>;
    using S1 = std::integral_constant<int, DEPTH - 1>;
    static_assert((TB / KC) % 2 == 0, "chunks per step even: the loop below takes them in pairs");
    if (!pre && nch > 0) {
      load(src(Ab, Jc, cin, 0), S0{});
      // keeps the two sets' loads in order, as the loop issues them: the
      // waits at the loop head are counted over both entries to the loop
      __builtin_amdgcn_sched_barrier(0);
      if constexpr (DEPTH == 2) load(src(Ab, Jc, cin, 1), S1{});
    }
    for (int c = 0; c < nch; c += DEPTH) {
      step(c, S0{});
      if constexpr (DEPTH == 2) step(c + 1, S1{});
    }
    pre = nch > 0 && chain;
    if constexpr (M3) {
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int bb = 0; bb < NBJ; bb++) {
          acci[a][bb] = acci[a][bb] - accr[a][bb] - accx[a % 2][bb % NX];
          accr[a][bb] = accr[a][bb] - accx[a % 2][bb % NX];
        }
    }
    // this wave's columns c0 .. c0 + 32*TJ - 1 lie in tile Jc + c0/64
    const int tb = c0 / TB;
    if (nch > 0 && tb < cin) {
      // all loads of the read-modify-write first, then the stores
      const gptr<R> Cg = tile_ptr<R, C>(p, S.tiles, I, Jc + tb);
      const int cl0 = c0 % TB;
      R cr[2][NBJ][4], ci[2][NBJ][4];
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int bb = 0; bb < NBJ; bb++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int r = r0 + a * 16 + Mfma<R>::row(lk, v), cc = cl0 + bb * 16 + li;
            if constexpr (C) {
              const r2 t = ((gptr<r2>)Cg)[r * TB + cc];
              cr[a][bb][v] = t.x;
              ci[a][bb][v] = t.y;
            } else {
              cr[a][bb][v] = Cg[r * TB + cc];
            }
          }
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int bb = 0; bb < NBJ; bb++)
#pragma unroll
          for (int v = 0; v < 4; v++) {
            const int r = r0 + a * 16 + Mfma<R>::row(lk, v), cc = cl0 + bb * 16 + li;
            if constexpr (C)
              ((gptr<r2>)Cg)[r * TB + cc] = r2{cr[a][bb][v] - accr[a][bb][v], ci[a][bb][v] - acci[a][bb][v]};
            else
              Cg[r * TB + cc] = cr[a][bb][v] - accr[a][bb][v];
          }
    }
    if (!has_next) break;
    cur = nxt;
    jw = jn;
  }
}

// ---------------------------------------------------------------- panel
// UL factorisation of the diagonal tile (K,K) without pivoting, from its last
// row to its first as mat_solve_gaussian.f90:196-302 does row by row, then the
// inverses U_D^-1 and L_D^-1 into dinv(K).  The pivot test is the
// reference's: |a|^2 < 1e-20 (complex) or |a| < 1e-10 (real); the largest
// failing row (the first in elimination order) is recorded 1-based in *fail.
// One workgroup of PNT threads per item.
//
// Elimination: PL lanes per row; lane p of row r keeps the elements
// (r, p + PL t) in registers (the step loop stays rolled, so the element of
// the current step is picked by selects over t).  Step q: the lanes of row q
// take the pivot from their row (lane shuffle), normalise the row's lower part
// (the stored L row, a_qk / a_qq) and publish it to LDS with zeros from column
// q on; after one barrier every row r < q subtracts a_rq * row q, a_rq coming
// from its own row's lanes.  Every element receives the reference's updates in
// its order.  Inverses: the two halves of the workgroup form U_D^-1 and L_D^-1
// at the same time, IL lanes per column j holding x_k (k = p + IL t); entries
// not yet computed are zero, so each substitution step is a full 64-term dot
// product with the published factor row, reduced across the IL lanes.
constexpr int PNT = 1024;  // panel workgroup
constexpr int PL = 16;     // lanes per row in the elimination
constexpr int IL = 8;      // lanes per column in the inverses

template <int CTRL>
__device__ __forceinline__ float dpp(float v) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double v) {
  const long long b = __builtin_bit_cast(long long, v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)(b & 0xffffffffLL), CTRL, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xf, 0xf, false);
  return __builtin_bit_cast(double, (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

// Sum over each group of eight lanes (every lane gets its group's sum):
// within quads by DPP quad permutations, then across the two quads of the
// group by the half-row mirror (lane i and lane 7 - i).
template <class R>
__device__ __forceinline__ R lane_sum8(R v) {
  v += dpp<0xB1>(v);   // lanes 1,0,3,2
  v += dpp<0x4E>(v);   // lanes 2,3,0,1
  v += dpp<0x141>(v);  // half-row mirror
  return v;
}

// X = U_D^-1 (UPPER: from the last row up, Vd[i] = 1 / U(i,i)) or L_D^-1
// (the unit-lower part, from the first row down) of the factored tile in P,
// written to X; thread t of the IL * TB taking part: column t / IL.
template <class R, bool C, int LD, bool UPPER>
__device__ void tri_inverse(const R (*P)[TB][LD], const R (*Vd)[TB], gptr<R> X, int t) {
  constexpr int NC = Elem<R, C>::NC;
  constexpr int IQ = TB / IL;
  const int j = t / IL, part = t % IL;
  R xr[IQ], xi[IQ];
#pragma unroll
  for (int u = 0; u < IQ; u++) xr[u] = xi[u] = 0;
#pragma unroll 1
  for (int s = 0; s < TB; ++s) {
    const int i = UPPER ? TB - 1 - s : s;
    R sr = 0, si = 0;
#pragma unroll
    for (int u = 0; u < IQ; u++) {
      const R fr = P[0][i][part + IL * u];
      if constexpr (C) {
        const R fi = P[NC - 1][i][part + IL * u];
        sr += fr * xr[u] - fi * xi[u];
        si += fr * xi[u] + fi * xr[u];
      } else {
        sr += fr * xr[u];
      }
    }
    sr = lane_sum8(sr);
    if constexpr (C) si = lane_sum8(si);
    const R br = (i == j ? R(1) : R(0)) - sr, bi = -si;
    R nr = br, ni = bi;
    if constexpr (UPPER) {
      if constexpr (C) {
        const R vr = Vd[0][i], vi = Vd[NC - 1][i];
        nr = br * vr - bi * vi;
        ni = br * vi + bi * vr;
      } else {
        nr = br * Vd[0][i];
      }
    }
    const int ti = part == i % IL ? i / IL : -1;
#pragma unroll
    for (int u = 0; u < IQ; u++) {
      xr[u] = u == ti ? nr : xr[u];
      xi[u] = u == ti ? ni : xi[u];
    }
  }
#pragma unroll
  for (int u = 0; u < IQ; u++) {
    const int64_t e = (int64_t)((part + IL * u) * TB + j) * NC;
    X[e] = xr[u];
    if constexpr (C) X[e + 1] = xi[u];
  }
}

template <class R, bool C>
__global__ __launch_bounds__(PNT) void k_panel(const SysDev<R> *__restrict__ sys, Items it) {
  static_assert(PNT == TB * PL && PNT == 2 * TB * IL, "one lane group per row, two inverses of IL lanes per column");
  constexpr int NC = Elem<R, C>::NC;
  constexpr int LD = TB + 1;
  constexpr int NQ = TB / PL;
  __shared__ R P[NC][TB][LD];  // published rows, then the factored tile
  __shared__ R Vd[NC][TB];     // 1 / U_D(i,i)
  const int tid = threadIdx.x;
  const int myr = tid / PL, part = tid % PL;
  const SysDev<R> &S = sys[it.sys[blockIdx.x]];
  const DevPlan &p = S.p;
  const int K = it.K[blockIdx.x];
  const gptr<R> T = tile_ptr<R, C>(p, S.tiles, K, K);
  R ar[NQ], ai[NQ];
#pragma unroll
  for (int t = 0; t < NQ; t++) {
    const int64_t e = (int64_t)(myr * TB + part + PL * t) * NC;
    ar[t] = T[e];
    ai[t] = C ? T[e + 1] : R(0);
  }
#pragma unroll 1
  for (int q = TB - 1; q >= 0; --q) {
    const int tq = q / PL;
    R mr = 0, mi = 0;
#pragma unroll
    for (int t = 0; t < NQ; t++) {
      mr = t == tq ? ar[t] : mr;
      mi = t == tq ? ai[t] : mi;
    }
    const R fr = __shfl(mr, q % PL, PL);  // a_{myr,q}; the pivot in row q's lanes
    const R fi = C ? __shfl(mi, q % PL, PL) : R(0);
    if (myr == q) {
      R vr, vi = 0;
      if constexpr (C) {
        const R den = fr * fr + fi * fi;
        if (part == 0 && den < R(1e-20) && K * TB + q < p.n) atomicMax(S.fail, K * TB + q + 1);
        vr = fr / den;
        vi = -fi / den;
      } else {
        if (part == 0 && fabs(fr) < R(1e-10) && K * TB + q < p.n) atomicMax(S.fail, K * TB + q + 1);
        vr = R(1) / fr;
      }
#pragma unroll
      for (int t = 0; t < NQ; t++) {
        const int k = part + PL * t;
        if (k < q) {  // normalise the stored row, :274-280 / :296-300
          const R lr = ar[t] * vr - ai[t] * vi, li = ai[t] * vr + ar[t] * vi;
          ar[t] = lr;
          ai[t] = li;
          P[0][q][k] = lr;
          if constexpr (C) P[NC - 1][q][k] = li;
        } else {
          P[0][q][k] = 0;
          if constexpr (C) P[NC - 1][q][k] = 0;
        }
      }
      if (part == 0) {
        Vd[0][q] = vr;
        if constexpr (C) Vd[NC - 1][q] = vi;
      }
    }
    __syncthreads();
    if (myr < q) {  // eliminate column q from row myr
#pragma unroll
      for (int t = 0; t < NQ; t++) {
        const int k = part + PL * t;
        const R lr = P[0][q][k];
        if constexpr (C) {
          const R li = P[NC - 1][q][k];
          ar[t] = ar[t] - fr * lr + fi * li;
          ai[t] = ai[t] - fr * li - fi * lr;
        } else {
          ar[t] = ar[t] - fr * lr;
        }
      }
    }
  }
  __syncthreads();  // every read of the published rows done
#pragma unroll
  for (int t = 0; t < NQ; t++) {
    const int k = part + PL * t;
    const int64_t e = (int64_t)(myr * TB + k) * NC;
    T[e] = ar[t];
    if constexpr (C) T[e + 1] = ai[t];
    P[0][myr][k] = ar[t];
    if constexpr (C) P[NC - 1][myr][k] = ai[t];
  }
  __syncthreads();
  if (tid < TB * IL)
    tri_inverse<R, C, LD, true>(P, Vd, dinv_tile<R, C>(S, K, D_UINV), tid);
  else
    tri_inverse<R, C, LD, false>(P, Vd, dinv_tile<R, C>(S, K, D_LINV), tid - TB * IL);
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
    const gptr<const R> T = trans ? tile_ptr<R, C>(p, S.tiles, J, K) : tile_ptr<R, C>(p, S.tiles, K, J);
    __syncthreads();
    stage_tile<R, C, LD>(T, Ts);
    for (int e = tid; e < TB * RC; e += NT) {
      const int rr = e / RC, cc = e % RC;
      const int64_t gi = ((int64_t)(J * TB + rr) * nrhs + rc0 + cc) * NC;
      const bool ok = rc0 + cc < nrhs;
      Xs[0][rr][cc] = ok ? G(S.X)[gi] : R(0);
      if constexpr (C) Xs[NC - 1][rr][cc] = ok ? G(S.X)[gi + 1] : R(0);
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
  const gptr<R> P = G(S.part) + ((int64_t)(blockIdx.y * maxparts + q) * TB * RC) * NC;
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
  const gptr<const R> Dg = dinv_tile<R, C>(S, K, upper ? D_UINV : D_LINV);
  stage_tile<R, C, LD>(Dg, Ds);
#pragma unroll
  for (int u = 0; u < 4; u++) {
    const int cc = gq * 4 + u;
    const int64_t gi = ((int64_t)(K * TB + r) * nrhs + rc0 + cc) * NC;
    const bool ok = rc0 + cc < nrhs;
    R sr = ok ? G(S.X)[gi] : R(0), si = 0;
    if constexpr (C) si = ok ? G(S.X)[gi + 1] : R(0);
    for (int pp = 0; pp < nparts; pp++) {
      const gptr<const R> P = G(S.part) + ((int64_t)(blockIdx.y * maxparts + pp) * TB * RC) * NC;
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
      G(S.X)[gi] = accr[u];
      if constexpr (C) G(S.X)[gi + 1] = acci[u];
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
// s_waitcnt immediate: vmcnt(0), expcnt and lgkmcnt at their maxima (no wait)
constexpr int WAIT_VM0 = 0x0F70;

// Publishes done[K] = 1 after the workgroup's stores of x_K: one thread,
// after a barrier at which every wave has waited for its own stores.  The
// agent-scope release fence writes the XCD's L2 back to memory (buffer_wbl2);
// the explicit vmcnt(0) makes that write-back complete before the flag store
// leaves.  An atomic release store alone compiles to buffer_wbl2 followed at
// once by the flag store when no store is pending (the wait the memory
// model places after the write-back is dropped), so a consumer on another
// XCD could see the flag while x_K's lines are still in this XCD's L2 and
// read the block's old contents from memory.
__device__ __forceinline__ void publish_flag(int *flag) {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __builtin_amdgcn_s_waitcnt(WAIT_VM0);
  __hip_atomic_store(flag, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

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

// Dependency q of a task is block base + dir * q.  Called by the 64 lanes of
// wave 0: polls the flags of dependencies q .. q+63 (those below n) with
// relaxed loads until dependency q is complete, then takes one agent-scope
// acquire fence, so the x blocks those flags publish are visible to the
// workgroup after its next barrier, and returns how many consecutive
// dependencies from q on are complete.  Polls are bounded: past SPIN_LIMIT
// the error flag is set, and once it is set the wait returns 0.
__device__ __forceinline__ int ready_run(const int *done, int base, int dir, int q, int n, int *err) {
  const int lane = threadIdx.x & 63;
  for (int spins = 0;; spins++) {
    const bool ok = q + lane < n &&
                    __hip_atomic_load(&done[base + dir * (q + lane)], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    const uint64_t ready = __ballot(ok);
    if (ready & 1) {
      // the acquire invalidates this CU's L1 and this XCD's L2; the wait
      // lets the invalidation complete before the caller's barrier releases
      // the other waves to load the published x blocks
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
      __builtin_amdgcn_s_waitcnt(WAIT_VM0);
      return ~ready == 0 ? 64 : __builtin_ctzll(~ready);
    }
    if (__builtin_amdgcn_readfirstlane(__hip_atomic_load(err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)) != 0)
      return 0;
    if (spins > SPIN_LIMIT) {
      if (lane == 0) atomicExch(err, 1);
      return 0;
    }
    __builtin_amdgcn_s_sleep(2);
  }
}

// acc += T x_J for thread (r, gq): row r of the staged tile (transposed when
// trans) times columns gq*4 .. gq*4+3 of the staged x block.
template <class R, bool C, int LD>
__device__ __forceinline__ void tile_mac(const R (*Ts)[TB][LD], const R (*Xs)[TB][RC + 1], bool trans, int r,
                                         int gq, R *accr, R *acci) {
#pragma unroll 4
  for (int k = 0; k < TB; k++) {
    const R tr = trans ? Ts[0][k][r] : Ts[0][r][k];
    R ti = 0;
    if constexpr (C) ti = trans ? Ts[1][k][r] : Ts[1][r][k];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const R xr = Xs[0][k][gq * 4 + u];
      if constexpr (C) {
        const R xi = Xs[1][k][gq * 4 + u];
        accr[u] += tr * xr - ti * xi;
        acci[u] += tr * xi + ti * xr;
      } else {
        accr[u] += tr * xr;
      }
    }
  }
}

// acc(64 x RC) += op(T)(64 x 64) X(64 x RC) on MFMA 16x16x4 (RC = 16): wave w
// computes rows 16w .. 16w+15; lane l accumulates rows 16w + Mfma<R>::row(l/16, v),
// column l%16.  op(T) = T, or T^T when trans.  Complex: four real MFMAs.
template <class R, bool C, int LD>
__device__ __forceinline__ void tile_mfma(const R (*Ts)[TB][LD], const R (*Xs)[TB][RC + 1], bool trans,
                                          typename Mfma<R>::v4 &accr, typename Mfma<R>::v4 &acci) {
  static_assert(RC == 16, "one 16-column MFMA block per wave");
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int row = (threadIdx.x >> 6) * 16 + li;
#pragma unroll 4
  for (int k4 = 0; k4 < TB; k4 += 4) {
    const int k = k4 + lk;
    const R ar = trans ? Ts[0][k][row] : Ts[0][row][k];
    const R br = Xs[0][k][li];
    accr = Mfma<R>::run(ar, br, accr);
    if constexpr (C) {
      const R ai = trans ? Ts[1][k][row] : Ts[1][row][k];
      const R bi = Xs[1][k][li];
      accr = Mfma<R>::run(-ai, bi, accr);
      acci = Mfma<R>::run(ar, bi, acci);
      acci = Mfma<R>::run(ai, br, acci);
    }
  }
}

// Stages rows J*TB .. J*TB+63, columns rc0 .. rc0+RC-1 of X (zero past nrhs).
template <class R, bool C>
__device__ __forceinline__ void stage_x(gptr<const R> X, int J, int rc0, int nrhs, R (*Xs)[TB][RC + 1]) {
  constexpr int NC = Elem<R, C>::NC;
  for (int e = threadIdx.x; e < TB * RC; e += NT) {
    const int rr = e / RC, cc = e % RC;
    const int64_t gi = ((int64_t)(J * TB + rr) * nrhs + rc0 + cc) * NC;
    const bool ok = rc0 + cc < nrhs;
    Xs[0][rr][cc] = ok ? X[gi] : R(0);
    if constexpr (C) Xs[NC - 1][rr][cc] = ok ? X[gi + 1] : R(0);
  }
}

// A tile and an x block (rows J*TB.., columns rc0..rc0+RC-1) held in registers
// between their global load and their LDS store, so the next pair's loads are
// in flight while the current pair is used.
template <class R, bool C>
struct TileRegs {
  typedef R r2 __attribute__((ext_vector_type(2)));
  static constexpr int PER = TB * TB / NT;
  r2 v2[C ? PER : 1];
  R v1[C ? 1 : PER];
  __device__ __forceinline__ void load(gptr<const R> g) {
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = threadIdx.x + q * NT;
      if constexpr (C) v2[q] = ((gptr<const r2>)g)[e];
      else v1[q] = g[e];
    }
  }
  template <int LD>
  __device__ __forceinline__ void store(R (*s)[TB][LD]) const {
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = threadIdx.x + q * NT;
      if constexpr (C) {
        s[0][e / TB][e % TB] = v2[q].x;
        s[1][e / TB][e % TB] = v2[q].y;
      } else {
        s[0][e / TB][e % TB] = v1[q];
      }
    }
  }
};

template <class R, bool C>
struct XRegs {
  static constexpr int NC = Elem<R, C>::NC;
  static constexpr int PER = TB * RC / NT;
  R re[PER], im[PER];
  // every lane loads (columns past nrhs read column rc0 and are zeroed), so
  // the loads sit on no branch
  __device__ __forceinline__ void load(gptr<const R> X, int J, int rc0, int nrhs) {
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = threadIdx.x + q * NT, rr = e / RC, cc = e % RC;
      const bool ok = rc0 + cc < nrhs;
      const int64_t gi = ((int64_t)(J * TB + rr) * nrhs + rc0 + (ok ? cc : 0)) * NC;
      const R a = X[gi];
      const R b = C ? X[gi + NC - 1] : R(0);
      re[q] = ok ? a : R(0);
      im[q] = ok ? b : R(0);
    }
  }
  __device__ __forceinline__ void store(R (*Xs)[TB][RC + 1]) const {
#pragma unroll
    for (int q = 0; q < PER; q++) {
      const int e = threadIdx.x + q * NT, rr = e / RC, cc = e % RC;
      Xs[0][rr][cc] = re[q];
      if constexpr (C) Xs[NC - 1][rr][cc] = im[q];
    }
  }
};

// Task K of a persistent solve accumulates its dependencies oldest first,
// each run of complete ones after a single acquire, so it follows the wave
// of completed block steps.  The last dependency is the neighbour nb = K -/+ 1.
// For the forward substitutions everything else is done before waiting on it:
// y_K = D^-1 (r_K - sum_old T(K,J) x_J) and the staging of the neighbour
// product M(K) = D^-1 T(K,nb) (MU / ML, formed during the factorisation), so
// the chain from one block step to the next carries only x_nb, one 64x64
// tile product x_K = y_K - M(K) x_nb and the publication.  The transposed
// substitutions stage T(nb,K)^T and D^-1 before the wait and apply both after.
// Depth of the persistent solve's register ring: FP32 complex stays at 256
// VGPRs with two sets (two workgroups per CU, as its LDS allows); FP64 runs
// one workgroup per CU (its LDS) and holds three sets in 449 VGPRs
template <class R>
constexpr int solve_pd() {
  return sizeof(R) == 4 ? 2 : 3;
}
template <class R, bool C>
__global__ __launch_bounds__(NT, sizeof(R) == 4 ? 2 : 1) void k_solve_persist(const SysDev<R> *__restrict__ sys, SolveSet ss, int kind,
                                                      int *counter, int *err) {
  constexpr int NC = Elem<R, C>::NC;
  constexpr int LD = TB + 1;
  __shared__ R Ts[NC][TB][LD];
  __shared__ R Ds[NC][TB][LD];
  __shared__ R Xs[NC][TB][RC + 1];
  __shared__ int tk, nrun;
  // tile index (tbase + column - J0 of its block row) of each dependency of
  // the current run: formed once per run by wave 0, so the ring's tile
  // addresses need no global load (a load's wait counts every load in flight
  // before it, the ring's tiles included)
  __shared__ int64_t toff[64];
  const int tid = threadIdx.x;
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
    // neighbour nb (the last dependency to complete) and the older range [ob, oe)
    const bool has_nb = je > jb;
    const int nb = asc ? je - 1 : jb;
    const int ob = has_nb ? (asc ? jb : jb + 1) : 0, oe = has_nb ? (asc ? je - 1 : je) : 0;
    auto tile = [&](int J) { return trans ? tile_ptr<R, C>(p, S.tiles, J, K) : tile_ptr<R, C>(p, S.tiles, K, J); };
    const gptr<const R> Dg = dinv_tile<R, C>(S, K, upper ? D_UINV : D_LINV);
    // thread (wave w, lane l) holds rows rows(v) = 16w + Mfma<R>::row(l/16, v),
    // v = 0..3, of right-hand side column rc0 + l%16 (the MFMA output layout)
    typedef typename Mfma<R>::v4 v4;
    const int lane = tid & 63, li = lane & 15, lk = lane >> 4, w16 = (tid >> 6) * 16;
    auto xidx = [&](int v, int rc0) {
      return ((int64_t)(K * TB + w16 + Mfma<R>::row(lk, v)) * nrhs + rc0 + li) * NC;
    };
    for (int rc0 = 0; rc0 < nrhs; rc0 += RC) {
      v4 accr = {0, 0, 0, 0}, acci = {0, 0, 0, 0};
      const bool colok = rc0 + li < nrhs;
      const int nold = oe - ob, base = asc ? ob : oe - 1, dir = asc ? 1 : -1;
      const gptr<const R> tiles0 = uniform(G((const R *)S.tiles));
      constexpr int64_t TE = (int64_t)TB * TB * NC;
      for (int q = 0; q < nold;) {
        if (tid < 64) {
          int run = ready_run(S.done, base, dir, q, nold, err);
          if (run == 0) run = min(64, nold - q);  // error: finish without waiting
          if (tid < run) {
            const int J = base + dir * (q + tid);
            toff[tid] = trans ? G(p.tbase)[J] + (K - G(p.J0)[J]) : G(p.tbase)[K] + (J - G(p.J0)[K]);
          }
          if (tid == 0) nrun = run;
        }
        __syncthreads();
        const int run = nrun;
        constexpr int PD = solve_pd<R>();
        // the run's tiles and x blocks load into a ring of PD register
        // sets, PD - 1 pairs ahead of the pair being multiplied, so that
        // several tiles per workgroup are in flight (one alone leaves the
        // stream bound by memory latency); every load is of a dependency
        // whose flag this run has already seen
        // Loads are issued on every path (past the end of the run they
        // repeat its last pair, which the L2 holds), and the ring's body has
        // no branch around them, so the compiler's wait before staging a
        // pair counts only the older loads and the younger sets stay in
        // flight; a conditional load would make it wait for all of them.
        TileRegs<R, C> tr[PD];
        XRegs<R, C> xr[PD];
#pragma unroll
        for (int s = 0; s < PD; s++) {
          const int v = min(s, run - 1);
          tr[s].load(tiles0 + toff[v] * TE);
          xr[s].load(G(S.X), base + dir * (q + v), rc0, nrhs);
        }
        const int full = run - run % PD;
        for (int u0 = 0; u0 < full; u0 += PD) {
#pragma unroll
          for (int s = 0; s < PD; s++) {
            tr[s].store(Ts);
            xr[s].store(Xs);
            __syncthreads();
            const int v = min(u0 + s + PD, run - 1);
            tr[s].load(tiles0 + toff[v] * TE);
            xr[s].load(G(S.X), base + dir * (q + v), rc0, nrhs);
            tile_mfma<R, C, LD>(Ts, Xs, trans, accr, acci);
            __syncthreads();
          }
        }
#pragma unroll
        for (int s = 0; s < PD; s++) {
          if (full + s < run) {
            tr[s].store(Ts);
            xr[s].store(Xs);
            __syncthreads();
            tile_mfma<R, C, LD>(Ts, Xs, trans, accr, acci);
            __syncthreads();
          }
        }
        q += run;
      }
      // everything independent of the neighbour
      if (has_nb) stage_tile<R, C, LD>(trans ? tile(nb) : dinv_tile<R, C>(S, K, upper ? D_MU : D_ML), Ts);
      stage_tile<R, C, LD>(Dg, Ds);
      R sr[4] = {0, 0, 0, 0}, si[4] = {0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < 4; v++) {
        if (colok) {
          sr[v] = G(S.X)[xidx(v, rc0)];
          if constexpr (C) si[v] = G(S.X)[xidx(v, rc0) + 1];
        }
      }
      v4 yr = {0, 0, 0, 0}, yi = {0, 0, 0, 0};
      if (!trans) {
        // y_K = D^-1 (r_K - acc), then x_K = y_K - M(K) x_nb
#pragma unroll
        for (int v = 0; v < 4; v++) {
          Xs[0][w16 + Mfma<R>::row(lk, v)][li] = sr[v] - accr[v];
          if constexpr (C) Xs[NC - 1][w16 + Mfma<R>::row(lk, v)][li] = si[v] - acci[v];
        }
        __syncthreads();
        tile_mfma<R, C, LD>(Ds, Xs, false, yr, yi);
        if (has_nb) {
          if (tid < 64) (void)ready_run(S.done, nb, 1, 0, 1, err);
          __syncthreads();
          stage_x<R, C>(G(S.X), nb, rc0, nrhs, Xs);
          __syncthreads();
          v4 zr = {0, 0, 0, 0}, zi = {0, 0, 0, 0};
          tile_mfma<R, C, LD>(Ts, Xs, false, zr, zi);
          yr -= zr;
          yi -= zi;
        }
      } else {
        if (has_nb) {
          if (tid < 64) (void)ready_run(S.done, nb, 1, 0, 1, err);
          __syncthreads();
          stage_x<R, C>(G(S.X), nb, rc0, nrhs, Xs);
          __syncthreads();
          tile_mfma<R, C, LD>(Ts, Xs, trans, accr, acci);
        }
        __syncthreads();
        // x_K = Dinv'(K) (r_K - acc)
#pragma unroll
        for (int v = 0; v < 4; v++) {
          Xs[0][w16 + Mfma<R>::row(lk, v)][li] = sr[v] - accr[v];
          if constexpr (C) Xs[NC - 1][w16 + Mfma<R>::row(lk, v)][li] = si[v] - acci[v];
        }
        __syncthreads();
        tile_mfma<R, C, LD>(Ds, Xs, trans, yr, yi);
      }
#pragma unroll
      for (int v = 0; v < 4; v++) {
        if (colok) {
          G(S.X)[xidx(v, rc0)] = yr[v];
          if constexpr (C) G(S.X)[xidx(v, rc0) + 1] = yi[v];
        }
      }
      __syncthreads();
    }
    // publish x_K to every XCD, then the flag: each wave waits for its own
    // stores to complete, and one write-back of the L2 after the barrier
    // covers the whole workgroup
    __builtin_amdgcn_s_waitcnt(WAIT_VM0);
    __syncthreads();
    if (tid == 0) publish_flag(&S.done[K]);
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
// unit diagonal.  Values are interleaved (re, im) when complex.  One wave per
// row, its lanes striding over the row's entries (the outer-sphere rows hold
// thousands), the row's block metadata loaded once.
constexpr int SCATTER_ROWS = 4;  // rows (waves) per workgroup
template <class R, bool C>
__global__ __launch_bounds__(64 * SCATTER_ROWS) void k_scatter_csr(DevPlan p, R *__restrict__ tiles,
                                                                   const int64_t *__restrict__ row_ptr,
                                                                   const int *__restrict__ col,
                                                                   const double *__restrict__ vals) {
  constexpr int NC = Elem<R, C>::NC;
  constexpr int64_t TE = (int64_t)TB * TB * NC;
  const int lane = threadIdx.x & 63;
  const int row = blockIdx.x * SCATTER_ROWS + (threadIdx.x >> 6);
  if (row >= p.NB * TB) return;
  const int I = row / TB, rr = row % TB;
  if (row >= p.n) {
    if (lane == 0) tile_ptr<R, C>(p, tiles, I, I)[(int64_t)(rr * TB + rr) * NC] = R(1);
    return;
  }
  const gptr<R> T0 = G(tiles) + G(p.tbase)[I] * TE;  // tile (I, J0(I))
  const int j0 = G(p.J0)[I];
  const int64_t q1 = G(row_ptr)[row + 1];
  for (int64_t q = G(row_ptr)[row] + lane; q < q1; q += 64) {
    const int c = G(col)[q] - 1;
    const int64_t e = (c / TB - j0) * TE + (int64_t)(rr * TB + (c % TB)) * NC;
    T0[e] = R(G(vals)[q * NC]);
    if constexpr (C) T0[e + 1] = R(G(vals)[q * NC + 1]);
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
