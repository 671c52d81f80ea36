// Block-skyline UL factorisation and block triangular solves (host side).
#include "solver.h"

#include <algorithm>
#include <cstring>

namespace fdmgpu {

static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
const char *get_error() { return g_err.c_str(); }

int Plan::build(int n_, const int *lb1, const int *lb2) {
  if (n_ <= 0) {
    set_error("nligne must be positive");
    return -1;
  }
  for (int i = 0; i < n_; i++) {
    const int ii = i + 1;
    if (lb1[i] < 1 || lb1[i] > ii || lb2[i] < ii || lb2[i] > n_) {
      set_error("profile bounds out of range at row " + std::to_string(ii));
      return -1;
    }
    if (i > 0 && lb1[i] < lb1[i - 1]) {
      set_error("lb1 must be non-decreasing (newind's sweep, mat.f90:1882); row " + std::to_string(ii));
      return -1;
    }
  }
  release();
  n = n_;
  NB = (n + TB - 1) / TB;
  J0.assign(NB, 0);
  J1.assign(NB, 0);
  I0.assign(NB, 0);
  Jx.assign(NB, 0);
  tbase.assign(NB, 0);
  for (int I = 0; I < NB; I++) {
    const int r0 = I * TB, r1 = std::min(n, (I + 1) * TB);
    int j0 = I, j1 = I;
    if (r0 < n) j0 = std::min(j0, (lb1[r0] - 1) / TB);
    for (int r = r0; r < r1; r++) j1 = std::max(j1, (lb2[r] - 1) / TB);
    J0[I] = j0;
    J1[I] = j1;
  }
  for (int I = 1; I < NB; I++) J1[I] = std::max(J1[I], J1[I - 1]);
  int64_t acc = 0;
  for (int I = 0; I < NB; I++) {
    tbase[I] = acc;
    acc += J1[I] - J0[I] + 1;
  }
  ntiles = acc;
  int I = 0;
  for (int K = 0; K < NB; K++) {  // first block row reaching column K (J1 non-decreasing)
    while (J1[I] < K) I++;
    I0[K] = I;
  }
  int J = NB - 1;
  for (int K = NB - 1; K >= 0; K--) {  // last block row whose envelope starts at or before K
    while (J0[J] > K) J--;
    Jx[K] = J;
  }
  maxlen = 0;
  for (int kind = 0; kind < 4; kind++)
    for (int K = 0; K < NB; K++) {
      int jb, je;
      solve_range(kind, K, J0[K], J1[K], I0[K], Jx[K], &jb, &je);
      maxlen = std::max(maxlen, je - jb);
    }
  return 0;
}

int Plan::upload() {
  FDM_HIP(hipMalloc(&dJ0, sizeof(int) * NB));
  FDM_HIP(hipMalloc(&dJ1, sizeof(int) * NB));
  FDM_HIP(hipMalloc(&dI0, sizeof(int) * NB));
  FDM_HIP(hipMalloc(&dJx, sizeof(int) * NB));
  FDM_HIP(hipMalloc(&dtbase, sizeof(int64_t) * NB));
  FDM_HIP(hipMemcpy(dJ0, J0.data(), sizeof(int) * NB, hipMemcpyHostToDevice));
  FDM_HIP(hipMemcpy(dJ1, J1.data(), sizeof(int) * NB, hipMemcpyHostToDevice));
  FDM_HIP(hipMemcpy(dI0, I0.data(), sizeof(int) * NB, hipMemcpyHostToDevice));
  FDM_HIP(hipMemcpy(dJx, Jx.data(), sizeof(int) * NB, hipMemcpyHostToDevice));
  FDM_HIP(hipMemcpy(dtbase, tbase.data(), sizeof(int64_t) * NB, hipMemcpyHostToDevice));
  return 0;
}

void Plan::release() {
  if (dJ0) (void)hipFree(dJ0);
  if (dJ1) (void)hipFree(dJ1);
  if (dI0) (void)hipFree(dI0);
  if (dJx) (void)hipFree(dJx);
  if (dtbase) (void)hipFree(dtbase);
  dJ0 = dJ1 = dI0 = dJx = nullptr;
  dtbase = nullptr;
}

// Reference elimination work on the element profile: row ii is updated by
// every row j in (ii, lb2(ii)] over that row's stored part [lb1(j), j).
double Plan::factor_macs(int n, const int *lb1, const int *lb2) {
  double s = 0;
  for (int i = 0; i < n; i++) {
    const int ii = i + 1;
    for (int j = ii + 1; j <= lb2[i]; j++) s += (double)(j - lb1[j - 1]);
  }
  return s;
}

double Plan::solve_macs(int n, const int *lb1, const int *lb2) {
  double s = 0;
  for (int i = 0; i < n; i++) s += (double)(lb2[i] - (i + 1)) + (double)((i + 1) - lb1[i]);
  return s;
}

double Plan::tile_other_macs() const {
  double s = 0;
  for (int K = 0; K < NB; K++) s += 0.5 * (double)((K - I0[K]) + (K - J0[K])) + 2.0 / 3.0;
  return s * (double)TB * TB * TB;
}

double Plan::tile_update_macs() const {
  double s = 0;
  for (int K = 0; K < NB; K++) s += (double)(K - I0[K]) * (double)(K - J0[K]);
  return s * (double)TB * TB * TB;
}

template <class R, bool C>
int Group<R, C>::add(Plan *plan, int nrhs, R *X) {
  if ((int)m.size() >= MAXG) {
    set_error("a group holds at most " + std::to_string(MAXG) + " systems");
    return -1;
  }
  Member mb;
  mb.plan = plan;
  mb.nrhs = nrhs;
  const size_t te = (size_t)TB * TB * NC;
  FDM_HIP(hipMalloc(&mb.tiles, sizeof(R) * te * (size_t)plan->ntiles));
  FDM_HIP(hipMalloc(&mb.dinv, sizeof(R) * te * DSLOTS * (size_t)plan->NB));
  FDM_HIP(hipMalloc(&mb.fail, sizeof(int)));
  FDM_HIP(hipMalloc(&mb.done, sizeof(int) * plan->NB));
  if (X) {
    mb.X = X;
    owns_x.push_back(0);
  } else {
    FDM_HIP(hipMalloc(&mb.X, sizeof(R) * (size_t)plan->NB * TB * nrhs * NC));
    owns_x.push_back(1);
  }
  m.push_back(mb);
  active.push_back(1);
  return 0;
}

template <class R, bool C>
int Group<R, C>::finalize(hipStream_t s) {
  stream = s;
  maxparts = 1;
  for (auto &mb : m) maxparts = std::max(maxparts, (mb.plan->maxlen + TPW - 1) / TPW);
  for (auto &mb : m) {
    const int nch = (mb.nrhs + RC - 1) / RC;
    if (mb.part) (void)hipFree(mb.part);
    FDM_HIP(hipMalloc(&mb.part, sizeof(R) * (size_t)maxparts * nch * TB * RC * NC));
  }
  std::vector<SysDev<R>> h(m.size());
  for (size_t g = 0; g < m.size(); g++)
    h[g] = SysDev<R>{m[g].plan->dev(), m[g].tiles, m[g].dinv, m[g].fail, m[g].X, m[g].part, m[g].nrhs, m[g].done};
  if (d_sys) (void)hipFree(d_sys);
  if (!ctl) {
    FDM_HIP(hipMalloc(&ctl, sizeof(int) * 2));
    FDM_HIP(hipMemset(ctl, 0, sizeof(int) * 2));
  }
  if (!side) FDM_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
  if (!ev_fact) FDM_HIP(hipEventCreateWithFlags(&ev_fact, hipEventDisableTiming));
  if (!ev_rest) FDM_HIP(hipEventCreateWithFlags(&ev_rest, hipEventDisableTiming));
  if (!ev_step) FDM_HIP(hipEventCreateWithFlags(&ev_step, hipEventDisableTiming));
  if (!ev_left) FDM_HIP(hipEventCreateWithFlags(&ev_left, hipEventDisableTiming));
  FDM_HIP(hipMalloc(&d_sys, sizeof(SysDev<R>) * h.size()));
  FDM_HIP(hipMemcpy(d_sys, h.data(), sizeof(SysDev<R>) * h.size(), hipMemcpyHostToDevice));
  return 0;
}

template <class R, bool C>
void Group<R, C>::release() {
  for (size_t g = 0; g < m.size(); g++) {
    auto &mb = m[g];
    if (mb.tiles) (void)hipFree(mb.tiles);
    if (mb.dinv) (void)hipFree(mb.dinv);
    if (mb.part) (void)hipFree(mb.part);
    if (mb.fail) (void)hipFree(mb.fail);
    if (mb.done) (void)hipFree(mb.done);
    if (mb.X && owns_x[g]) (void)hipFree(mb.X);
  }
  m.clear();
  active.clear();
  owns_x.clear();
  if (d_sys) (void)hipFree(d_sys);
  d_sys = nullptr;
  if (ctl) (void)hipFree(ctl);
  ctl = nullptr;
  for (auto e : ev) (void)hipEventDestroy(e);
  ev.clear();
  if (ev_fact) (void)hipEventDestroy(ev_fact);
  if (ev_rest) (void)hipEventDestroy(ev_rest);
  if (ev_step) (void)hipEventDestroy(ev_step);
  if (ev_left) (void)hipEventDestroy(ev_left);
  ev_step = ev_left = nullptr;
  if (side) (void)hipStreamDestroy(side);
  ev_fact = ev_rest = nullptr;
  side = nullptr;
}

template <class R, bool C>
size_t Group<R, C>::factor_bytes() const {
  size_t s = 0;
  for (auto &mb : m) s += sizeof(R) * (size_t)TB * TB * NC * ((size_t)mb.plan->ntiles + DSLOTS * (size_t)mb.plan->NB);
  return s;
}

template <class R, bool C>
int Group<R, C>::zero(int g) {
  auto &mb = m[g];
  const size_t cnt = (size_t)TB * TB * NC * (size_t)mb.plan->ntiles;
  hipLaunchKernelGGL((k_fill<R>), dim3(2048), dim3(256), 0, stream, mb.tiles, cnt, R(0));
  hipLaunchKernelGGL((k_fill<int>), dim3(1), dim3(64), 0, stream, mb.fail, (size_t)1, 0);
  FDM_HIP(hipGetLastError());
  return 0;
}

template <class R, bool C>
int Group<R, C>::scatter_csr(int g, const int64_t *d_row_ptr, const int *d_col, const double *d_vals) {
  auto &mb = m[g];
  const int rows = mb.plan->NB * TB;
  hipLaunchKernelGGL((k_scatter_csr<R, C>), dim3((rows + SCATTER_ROWS - 1) / SCATTER_ROWS), dim3(64 * SCATTER_ROWS), 0,
                     stream, mb.plan->dev(), mb.tiles, d_row_ptr, d_col, d_vals);
  FDM_HIP(hipGetLastError());
  return 0;
}

namespace {
inline void push(Items &it, int g, int K, int nwg) {
  if (it.count == 0) it.off[0] = 0;
  it.sys[it.count] = g;
  it.K[it.count] = K;
  it.off[it.count + 1] = it.off[it.count] + nwg;
  it.count++;
}
}  // namespace

namespace {
inline void push_trail(TrailItems &it, int g, int Kb, int nk, int i0, int ni, int j0, int nj) {
  if (ni <= 0 || nj <= 0 || nk <= 0) return;
  if (it.count == 0) it.off[0] = it.moff[0] = 0;
  const int c = it.count;
  it.sys[c] = g;
  it.Kb[c] = Kb;
  it.nk[c] = nk;
  it.i0[c] = i0;
  it.j0[c] = j0;
  it.nj[c] = nj;
  it.ni[c] = ni;
  it.off[c + 1] = it.off[c] + ni * nj;
  it.moff[c + 1] = it.moff[c] + ni * ((nj + 1) / 2);
  it.count++;
}
}  // namespace

// Grid of k_trail: at most persist_wg workgroups, a multiple of NXCD.
static int xcd_grid(int items, int persist_wg) {
  const int g = std::max(std::min(items, persist_wg), 1);
  return (g + NXCD - 1) / NXCD * NXCD;
}

// Blocked UL from the last block row up.  With gk > 0 the block steps are
// taken in outer groups of gk.  Inside a group the update is left-looking:
// just before step K factorises its diagonal tile, block row K (columns
// J0(K)..K) and block column K (rows I0(K)..K-1) receive the contributions
// of the group's earlier steps K+1..Khi in one launch ("cross"); step K then
// forms its L row and U column.  The rest of the trailing matrix (rows and
// columns below the group) receives the whole group's rank-(gk*TB) update in
// one pass at the group's last step.  Without groups (gk = 0) every step
// updates its whole trailing block right-looking.
//
// Lookahead: the next group's chain (its left-looking launches, panels and
// triangular products) reads and writes only the next group's block rows and
// columns and tiles outside the current group's rectangle.  So at a group's
// last step the rectangle is split: its part in the next group's block rows
// and columns is updated on `stream` before the next chain, the rest on
// `side` beside it (started after that part); the next group's rectangle
// waits for the rest.
template <class R, bool C>
int Group<R, C>::factor() {
  int maxNB = 0;
  for (size_t g = 0; g < m.size(); g++)
    if (active[g]) maxNB = std::max(maxNB, m[g].plan->NB);
  if (time_updates) {
    for (auto e : ev) (void)hipEventDestroy(e);
    ev.assign(6 * (size_t)maxNB, nullptr);
    for (auto &e : ev) FDM_HIP(hipEventCreate(&e));
  }
  int nev = 0;
  // A start event queued behind unfinished kernels would be stamped before
  // they complete; drain the launch's stream first so the pair brackets one
  // launch (work on the other stream may still run beside it).
  auto timed = [&](hipStream_t st, auto launch) -> int {
    if (time_updates) {
      FDM_HIP(hipStreamSynchronize(st));
      FDM_HIP(hipEventRecord(ev[nev], st));
    }
    launch();
    if (time_updates) {
      FDM_HIP(hipEventRecord(ev[nev + 1], st));
      nev += 2;
    }
    return 0;
  };
  auto launch_trail = [&](hipStream_t st, const TrailItems &it, int wg) {
    if (trail_tile == 128)
      hipLaunchKernelGGL((k_trail<R, C, 2>), dim3(xcd_grid(it.moff[it.count], wg)), dim3(NT), 0, st,
                         (const SysDev<R> *)d_sys, it);
    else
      hipLaunchKernelGGL((k_trail<R, C, 1>), dim3(xcd_grid(it.off[it.count], wg)), dim3(NT), 0, st,
                         (const SysDev<R> *)d_sys, it);
  };
  const int rest_wg = lookahead > 1 ? lookahead : persist_wg;
  bool rest_pending = false;  // a remainder on `side` not yet waited for by `stream`
  bool left_pending = false;  // an overlapped left-looking update on `side` not yet joined
  for (int s = 0; s < maxNB; s++) {
    Items ip{}, itx{}, itu{};
    TrailItems cross{}, trail{}, rest{}, left_next{};
    for (int g = 0; g < (int)m.size(); g++) {
      if (!active[g]) continue;
      const Plan &pl = *m[g].plan;
      const int t = s - (maxNB - pl.NB);
      if (t < 0) continue;
      const int K = pl.NB - 1 - t;
      const int nJ = K - pl.J0[K], nI = K - pl.I0[K];
      push(ip, g, K, 1);
      if (nJ + nI > 0) push(itx, g, K, nJ + nI);  // L row and U column in one launch
      if (gk <= 0) {
        if (nI > 0 && nJ > 0) push(itu, g, K, nI * nJ);
        continue;
      }
      const int Khi = pl.NB - 1 - gk * (t / gk), Klo = std::max(Khi - gk + 1, 0);
      if (K < Khi) {
        // block row K, columns J0(K)..K (diagonal tile included), steps
        // K+1..Khi; with the overlap only step K+1 here, the others having
        // run on `side` beside step K+1's panel
        const int nk = overlap ? 1 : Khi - K;
        push_trail(cross, g, K + 1, nk, K, 1, pl.J0[K], nJ + 1);
        // block column K, rows I0(K)..K-1
        push_trail(cross, g, K + 1, nk, pl.I0[K], nI, K, 1);
      }
      if (overlap && K < Khi && K > Klo) {
        // block row and column K-1 from steps K+1..Khi: they read only those
        // steps' finished products and write tiles that step K's panel,
        // products and single-step update neither read nor write
        const int K2 = K - 1, nJ2 = K2 - pl.J0[K2], nI2 = K2 - pl.I0[K2];
        push_trail(left_next, g, K + 1, Khi - K, K2, 1, pl.J0[K2], nJ2 + 1);
        push_trail(left_next, g, K + 1, Khi - K, pl.I0[K2], nI2, K2, 1);
      }
      if (K == Klo) {
        const int nk = Khi - Klo + 1, i0 = pl.I0[Klo], j0 = pl.J0[Klo];
        if (lookahead && Klo > 0) {
          // the next group's steps are Klo2..Klo-1
          const int Klo2 = std::max(Klo - gk, 0), a = std::max(i0, Klo2), b = std::max(j0, Klo2);
          push_trail(trail, g, Klo, nk, a, Klo - a, j0, Klo - j0);      // its block rows
          push_trail(trail, g, Klo, nk, i0, a - i0, b, Klo - b);        // its block columns above them
          push_trail(rest, g, Klo, nk, i0, a - i0, j0, b - j0);         // the rest
        } else {
          push_trail(trail, g, Klo, nk, i0, Klo - i0, j0, Klo - j0);
        }
      }
    }
    if (left_next.count) FDM_HIP(hipEventRecord(ev_step, stream));  // step K+1 done: its products are final
    if (left_pending) {  // the earlier steps' part of this step's update
      FDM_HIP(hipStreamWaitEvent(stream, ev_left, 0));
      left_pending = false;
    }
    if (cross.count) timed(stream, [&]() { launch_trail(stream, cross, persist_wg); });
    if (left_next.count) {
      FDM_HIP(hipStreamWaitEvent(side, ev_step, 0));
      timed(side, [&]() { launch_trail(side, left_next, persist_wg); });
      FDM_HIP(hipEventRecord(ev_left, side));
      left_pending = true;
    }
    hipLaunchKernelGGL((k_panel<R, C>), dim3(ip.count), dim3(PNT), 0, stream, (const SysDev<R> *)d_sys, ip);
    if (itx.count)
      hipLaunchKernelGGL((k_tile_gemm<R, C>), dim3(std::min(itx.off[itx.count], persist_wg)), dim3(NT), 0, stream,
                         (const SysDev<R> *)d_sys, itx, (int)G_TRSM);
    if (time_updates) {
      const double tb = (double)TB * TB * NC * sizeof(R);
      for (int c = 0; c < itu.count; c++) {
        const Plan &pl = *m[itu.sys[c]].plan;
        const int K = itu.K[c], nI = K - pl.I0[K], nJ = K - pl.J0[K];
        update_bytes += tb * (2.0 * nI * nJ + nI + nJ);
      }
      for (const TrailItems *ti : {&cross, &trail, &rest})
        for (int c = 0; c < ti->count; c++)
          update_bytes += tb * (2.0 * ti->ni[c] * ti->nj[c] + (double)(ti->ni[c] + ti->nj[c]) * ti->nk[c]);
    }
    if (itu.count)
      timed(stream, [&]() {
        hipLaunchKernelGGL((k_tile_gemm<R, C>), dim3(std::min(itu.off[itu.count], persist_wg)), dim3(NT), 0, stream,
                           (const SysDev<R> *)d_sys, itu, (int)G_UPDATE);
      });
    if (trail.count) {
      // the rectangle (or its part in the next group's rows and columns)
      // overlaps the remainders of earlier groups
      if (rest_pending) {
        FDM_HIP(hipStreamWaitEvent(stream, ev_rest, 0));
        rest_pending = false;
      }
      timed(stream, [&]() { launch_trail(stream, trail, persist_wg); });
    }
    if (rest.count) {
      // the remainder starts once the next group's rows and columns are
      // updated, so it does not hold back the next group's first panel
      FDM_HIP(hipEventRecord(ev_fact, stream));
      FDM_HIP(hipStreamWaitEvent(side, ev_fact, 0));
      timed(side, [&]() { launch_trail(side, rest, rest_wg); });
      FDM_HIP(hipEventRecord(ev_rest, side));
      rest_pending = true;
    }
  }
  if (rest_pending) FDM_HIP(hipStreamWaitEvent(stream, ev_rest, 0));  // join before the substitutions
  if (left_pending) FDM_HIP(hipStreamWaitEvent(stream, ev_left, 0));
  // the substitutions' neighbour products MU(K), ML(K) of every block, off
  // the per-step chain: their tiles are final once steps K and K+1 are done
  Items nb{};
  for (int g = 0; g < (int)m.size(); g++)
    if (active[g]) push(nb, g, 0, 2 * m[g].plan->NB);
  if (nb.count)
    hipLaunchKernelGGL((k_tile_gemm<R, C>), dim3(std::min(nb.off[nb.count], persist_wg)), dim3(NT), 0, stream,
                       (const SysDev<R> *)d_sys, nb, (int)G_NEIGH);
  FDM_HIP(hipGetLastError());
  if (time_updates) {
    FDM_HIP(hipStreamSynchronize(side));
    FDM_HIP(hipStreamSynchronize(stream));
    for (int e = 0; e < nev; e += 2) {
      float ms = 0;
      FDM_HIP(hipEventElapsedTime(&ms, ev[e], ev[e + 1]));
      update_ms += ms;
      update_launches++;
    }
  }
  return 0;
}

template <class R, bool C>
int Group<R, C>::solve(int kind) {
  int maxNB = 0, nch = 1;
  for (size_t g = 0; g < m.size(); g++) {
    if (!active[g]) continue;
    maxNB = std::max(maxNB, m[g].plan->NB);
    nch = std::max(nch, (m[g].nrhs + RC - 1) / RC);
  }
  const bool asc = (kind == S_LOWER || kind == S_UPPER_T);
  if (persist_solve) {
    SolveSet ss{};
    ss.maxNB = maxNB;
    int ntask = 0;
    for (int g = 0; g < (int)m.size(); g++) {
      if (!active[g]) continue;
      ss.sys[ss.count] = g;
      ss.nb[ss.count] = m[g].plan->NB;
      ss.off[ss.count] = maxNB - m[g].plan->NB;
      ss.count++;
      ntask += m[g].plan->NB;
      hipLaunchKernelGGL((k_fill<int>), dim3((m[g].plan->NB + 255) / 256), dim3(256), 0, stream, m[g].done,
                         (size_t)m[g].plan->NB, 0);
    }
    hipLaunchKernelGGL((k_fill<int>), dim3(1), dim3(64), 0, stream, ctl, (size_t)1, 0);
    hipLaunchKernelGGL((k_solve_persist<R, C>), dim3(std::min(ntask, solve_wg)), dim3(NT), 0, stream,
                       (const SysDev<R> *)d_sys, ss, kind, ctl, ctl + 1);
    FDM_HIP(hipGetLastError());
    return 0;
  }
  for (int s = 0; s < maxNB; s++) {
    Items ipart{}, ifin{};
    for (int g = 0; g < (int)m.size(); g++) {
      if (!active[g]) continue;
      const Plan &pl = *m[g].plan;
      const int t = s - (maxNB - pl.NB);
      if (t < 0) continue;
      const int K = asc ? t : pl.NB - 1 - t;
      int jb, je;
      solve_range(kind, K, pl.J0[K], pl.J1[K], pl.I0[K], pl.Jx[K], &jb, &je);
      const int np = (je - jb + TPW - 1) / TPW;
      if (np > 0) push(ipart, g, K, np);
      push(ifin, g, K, 1);
    }
    if (ipart.count)
      hipLaunchKernelGGL((k_solve_partial<R, C>), dim3(ipart.off[ipart.count], nch), dim3(NT), 0, stream,
                         (const SysDev<R> *)d_sys, ipart, kind, maxparts);
    hipLaunchKernelGGL((k_solve_finish<R, C>), dim3(ifin.count, nch), dim3(NT), 0, stream, (const SysDev<R> *)d_sys,
                       ifin, kind, maxparts);
  }
  FDM_HIP(hipGetLastError());
  return 0;
}

template <class R, bool C>
int Group<R, C>::read_fail(std::vector<int> &rows) {
  rows.assign(m.size(), 0);
  int h[2] = {0, 0};
  for (size_t g = 0; g < m.size(); g++)
    FDM_HIP(hipMemcpyAsync(&rows[g], m[g].fail, sizeof(int), hipMemcpyDeviceToHost, stream));
  if (ctl) FDM_HIP(hipMemcpyAsync(h, ctl, sizeof(int) * 2, hipMemcpyDeviceToHost, stream));
  FDM_HIP(hipStreamSynchronize(stream));
  if (h[1]) {
    set_error("persistent substitution: a dependency wait timed out");
    (void)hipMemsetAsync(ctl + 1, 0, sizeof(int), stream);
    return -2;
  }
  return 0;
}

template <class R, bool C>
int Group<R, C>::check_solve() {
  int h = 0;
  if (!ctl) return 0;
  FDM_HIP(hipMemcpyAsync(&h, ctl + 1, sizeof(int), hipMemcpyDeviceToHost, stream));
  FDM_HIP(hipStreamSynchronize(stream));
  if (h) {
    set_error("persistent substitution: a dependency wait timed out");
    (void)hipMemsetAsync(ctl + 1, 0, sizeof(int), stream);
    return -2;
  }
  return 0;
}

template struct Group<double, true>;
template struct Group<double, false>;
template struct Group<float, true>;
template struct Group<float, false>;

}  // namespace fdmgpu
