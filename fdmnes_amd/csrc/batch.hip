// Device-resident solver: grouped factorisation of several energy points or
// clusters, hipGraph-captured, FP64 or FP32 + FP64 refinement, adjoint.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "batch.h"

namespace fdmgpu {

// sigma = -Im sum_i s_i x_i over n*nrhs entries, in two deterministic
// stages: workgroup b of SIG_PARTS sums a fixed contiguous range into
// part[b] (fixed-order tree), then one workgroup adds the parts in order.
__device__ __forceinline__ double block_sum256(double v, double *sh) {
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  return sh[0];
}

template <bool C>
__global__ __launch_bounds__(256) void k_sigma_part(const double *__restrict__ s, const double *__restrict__ x,
                                                    size_t cnt, double *__restrict__ part) {
  __shared__ double sh[256];
  const size_t chunk = (cnt + SIG_PARTS - 1) / SIG_PARTS;
  const size_t i0 = blockIdx.x * chunk, i1 = min(cnt, i0 + chunk);
  double acc = 0;
  for (size_t i = i0 + threadIdx.x; i < i1; i += 256) {
    if constexpr (C)
      acc += s[2 * i] * x[2 * i + 1] + s[2 * i + 1] * x[2 * i];
    else
      acc += 0.0 * s[i] * x[i];
  }
  const double t = block_sum256(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ __launch_bounds__(256) void k_sigma_final(const double *__restrict__ part, double *__restrict__ out) {
  __shared__ double sh[256];
  const double t = block_sum256(part[threadIdx.x], sh);
  if (threadIdx.x == 0) *out = -t;
}

// grad_i = Im sum_L lam[i][L] x[i][L]  (d sigma / d V_i for sigma = -Im s^T x,
// A x = s, A^T lam = s, dA/dV_i = e_i e_i^T).
template <bool C>
__global__ void k_adjoint_grad(const double *__restrict__ lam, const double *__restrict__ x, int n, int nrhs,
                               double *__restrict__ grad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0;
  for (int l = 0; l < nrhs; l++) {
    const size_t e = (size_t)i * nrhs + l;
    if constexpr (C) acc += lam[2 * e] * x[2 * e + 1] + lam[2 * e + 1] * x[2 * e];
  }
  grad[i] = acc;
}

template <class R, bool C>
Solver<R, C>::~Solver() {
  fb.clear();
  for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto &kv : sweeps_g) (void)hipGraphExecDestroy(kv.second);
  if (hnorm) (void)hipHostFree(hnorm);
  if (hsig) (void)hipHostFree(hsig);
  if (sig_buf) (void)hipFree(sig_buf);
  grp.release();
  for (auto &s : slots) {
    if (s.row_ptr) (void)hipFree(s.row_ptr);
    if (s.long_rows) (void)hipFree(s.long_rows);
    if (s.col) (void)hipFree(s.col);
    if (s.vals) (void)hipFree(s.vals);
    if (s.B) (void)hipFree(s.B);
    if (s.Xd) (void)hipFree(s.Xd);
    if (s.Rr) (void)hipFree(s.Rr);
    if (s.scr) (void)hipFree(s.scr);
  }
  if (own_stream && stream) (void)hipStreamDestroy(stream);
}

template <class R, bool C>
int Solver<R, C>::init(int nrhs_, hipStream_t s) {
  nrhs = nrhs_;
  if (s) {
    stream = s;
  } else {
    FDM_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    own_stream = true;
  }
  return 0;
}

template <class R, bool C>
int Solver<R, C>::add_system(int n, const int *lb1, const int *lb2, const int64_t *row_ptr, const int *col,
                             int *slot) {
  if (finalized) {
    set_error("add_system after the first solve");
    return -1;
  }
  if ((int)slots.size() >= MAXG) {
    set_error("a solver holds at most " + std::to_string(MAXG) + " systems");
    return -1;
  }
  if (row_ptr[0] != 0) {
    set_error("row_ptr[0] must be 0");
    return -1;
  }
  for (int i = 0; i < n; i++)
    for (int64_t q = row_ptr[i]; q < row_ptr[i + 1]; q++)
      if (col[q] < lb1[i] || col[q] > lb2[i]) {
        set_error("column " + std::to_string(col[q]) + " of row " + std::to_string(i + 1) +
                  " lies outside its profile [lb1, lb2]");
        return -1;
      }
  Slot sl;
  sl.plan.reset(new Plan());
  if (sl.plan->build(n, lb1, lb2)) return -1;
  if (sl.plan->upload()) return -2;
  sl.n = n;
  sl.nnz = row_ptr[n];
  sl.factor_macs = Plan::factor_macs(n, lb1, lb2);
  sl.solve_macs = Plan::solve_macs(n, lb1, lb2);
  const size_t xe = (size_t)sl.plan->NB * TB * nrhs * NC;
  FDM_HIP(hipMalloc(&sl.row_ptr, sizeof(int64_t) * (n + 1)));
  FDM_HIP(hipMalloc(&sl.col, sizeof(int) * std::max<int64_t>(sl.nnz, 1)));
  FDM_HIP(hipMalloc(&sl.vals, sizeof(double) * std::max<int64_t>(sl.nnz, 1) * NC));
  FDM_HIP(hipMalloc(&sl.B, sizeof(double) * xe));
  FDM_HIP(hipMemcpy(sl.row_ptr, row_ptr, sizeof(int64_t) * (n + 1), hipMemcpyHostToDevice));
  FDM_HIP(hipMemcpy(sl.col, col, sizeof(int) * sl.nnz, hipMemcpyHostToDevice));
  FDM_HIP(hipMemset(sl.B, 0, sizeof(double) * xe));
  std::vector<int> lr;
  for (int i = 0; i < n; i++)
    if (row_ptr[i + 1] - row_ptr[i] > LONG_ROW) lr.push_back(i);
  sl.nlong = (int)lr.size();
  if (sl.nlong) {
    FDM_HIP(hipMalloc(&sl.long_rows, sizeof(int) * lr.size()));
    FDM_HIP(hipMemcpy(sl.long_rows, lr.data(), sizeof(int) * lr.size(), hipMemcpyHostToDevice));
  }
  // refinement buffers (FP32 factors always refine; FP64 ones on request)
  FDM_HIP(hipMalloc(&sl.Xd, sizeof(double) * xe));
  FDM_HIP(hipMalloc(&sl.Rr, sizeof(double) * xe));
  FDM_HIP(hipMemset(sl.Rr, 0, sizeof(double) * xe));
  FDM_HIP(hipMalloc(&sl.scr, sizeof(double) * NORM_GRID * std::max(nrhs, 1)));
  if (grp.add(sl.plan.get(), nrhs)) return -2;
  *slot = (int)slots.size();
  slots.push_back(std::move(sl));
  return 0;
}

template <class R, bool C>
int Solver<R, C>::finalize() {
  if (finalized) return 0;
  if (grp.finalize(stream)) return -2;
  finalized = true;
  return 0;
}

template <class R, bool C>
int Solver<R, C>::set_active(int count, const int *ids) {
  std::fill(grp.active.begin(), grp.active.end(), 0);
  for (int k = 0; k < count; k++) {
    if (ids[k] < 0 || ids[k] >= (int)slots.size()) {
      set_error("bad slot id");
      return -1;
    }
    if (grp.active[ids[k]]) {
      set_error("a slot appears twice in one call");
      return -1;
    }
    grp.active[ids[k]] = 1;
  }
  return 0;
}

// Zero + scatter every active slot, load its RHS (FP64 path), factorise, and
// for FP64 also run both substitutions.  Captured once per active mask.
template <class R, bool C>
int Solver<R, C>::factor_phase(uint32_t mask) {
  auto record = [&]() -> int {
    for (size_t g = 0; g < slots.size(); g++) {
      if (!grp.active[g]) continue;
      if (grp.zero((int)g) || grp.scatter_csr((int)g, slots[g].row_ptr, slots[g].col, slots[g].vals)) return -2;
      if (std::is_same<R, double>::value) {
        const size_t tot = (size_t)slots[g].plan->NB * TB * nrhs * NC;
        hipLaunchKernelGGL((k_load_rhs<R>), dim3(1024), dim3(256), 0, stream, grp.m[g].X, (const double *)slots[g].B,
                           tot, tot);
      }
    }
    if (grp.factor()) return -2;
    if (std::is_same<R, double>::value)
      if (grp.solve(S_UPPER) || grp.solve(S_LOWER)) return -2;
    return 0;
  };
  if (!use_graphs || grp.time_updates) return record();
  auto it = graphs.find(mask);
  if (it == graphs.end()) {
    hipGraph_t graph;
    FDM_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int rc = record();
    hipError_t e = hipStreamEndCapture(stream, &graph);
    if (rc) return rc;
    if (e != hipSuccess) {
      set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      return -2;
    }
    hipGraphExec_t exec;
    FDM_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    (void)hipGraphDestroy(graph);
    it = graphs.emplace(mask, exec).first;
  }
  FDM_HIP(hipGraphLaunch(it->second, stream));
  return 0;
}

// One refinement sweep over the active slots: d = A32^-1 r (FP32 factors),
// x += d, r = b - A x (FP64), squared residual norms per right-hand side.
// Captured once per active mask.
template <class R, bool C>
int Solver<R, C>::sweep_phase(uint32_t mask) {
  auto record = [&]() -> int {
    for (size_t g = 0; g < slots.size(); g++) {
      if (!grp.active[g]) continue;
      const size_t tot = (size_t)slots[g].plan->NB * TB * nrhs * NC;
      if (convert<R>(slots[g].Rr, grp.m[g].X, tot, stream)) return -2;
    }
    if (grp.solve(S_UPPER) || grp.solve(S_LOWER)) return -2;
    for (size_t g = 0; g < slots.size(); g++) {
      if (!grp.active[g]) continue;
      Slot &sl = slots[g];
      const size_t tot = (size_t)sl.plan->NB * TB * nrhs * NC;
      DevCsr a{sl.n, sl.nnz, sl.row_ptr, sl.col, sl.vals, sl.long_rows, sl.nlong};
      if (accumulate<R>(sl.Xd, grp.m[g].X, tot, stream) || residual<C>(a, sl.B, sl.Xd, sl.Rr, nrhs, stream) ||
          launch_norms2<C>(sl.Rr, sl.n, nrhs, sl.scr, stream))
        return -2;
    }
    return 0;
  };
  if (!use_graphs || grp.time_updates) return record();
  auto it = sweeps_g.find(mask);
  if (it == sweeps_g.end()) {
    hipGraph_t graph;
    FDM_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int rc = record();
    hipError_t e = hipStreamEndCapture(stream, &graph);
    if (rc) return rc;
    if (e != hipSuccess) {
      set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      return -2;
    }
    hipGraphExec_t exec;
    FDM_HIP(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    (void)hipGraphDestroy(graph);
    it = sweeps_g.emplace(mask, exec).first;
  }
  FDM_HIP(hipGraphLaunch(it->second, stream));
  return 0;
}

// FP64 factors for one point whose FP32 factorisation or refinement failed:
// the direct solve, then (as the FP64 solver) at most max_sweeps corrections
// with the same factors while max_l ||r_l||/||b_l|| > tol; *rel gets the
// final residual.
template <class R, bool C>
int Solver<R, C>::fallback_fp64(int g, int *fail_row, int max_sweeps, double tol, double *rel) {
  Slot &sl = slots[g];
  if (fb.size() < slots.size()) fb.resize(slots.size());
  if (!fb[g]) {
    std::unique_ptr<Group<double, C>> f(new Group<double, C>());
    f->gk = grp.gk;
    f->persist_wg = grp.persist_wg;
    f->trail_tile = grp.trail_tile;
    f->persist_solve = grp.persist_solve;
    f->solve_wg = grp.solve_wg;
    if (f->add(sl.plan.get(), nrhs) || f->finalize(stream)) return -2;
    fb[g] = std::move(f);
  }
  Group<double, C> &f = *fb[g];
  const size_t tot = (size_t)sl.plan->NB * TB * nrhs * NC;
  if (f.zero(0) || f.scatter_csr(0, sl.row_ptr, sl.col, sl.vals)) return -2;
  hipLaunchKernelGGL((k_load_rhs<double>), dim3(1024), dim3(256), 0, stream, f.m[0].X, (const double *)sl.B, tot, tot);
  if (f.factor() || f.solve(S_UPPER) || f.solve(S_LOWER)) return -2;
  std::vector<int> fr;
  if (f.read_fail(fr)) return -2;
  *fail_row = fr[0];
  *rel = NAN;
  FDM_HIP(hipMemcpyAsync(sl.Xd, f.m[0].X, sizeof(double) * tot, hipMemcpyDeviceToDevice, stream));
  if (fr[0]) {
    FDM_HIP(hipStreamSynchronize(stream));
    return 0;
  }
  const DevCsr a{sl.n, sl.nnz, sl.row_ptr, sl.col, sl.vals, sl.long_rows, sl.nlong};
  const int scr = NORM_GRID * std::max(nrhs, 1);
  std::vector<double> bn(nrhs), rn(nrhs);
  if (col_norms2<C>(sl.B, sl.n, nrhs, bn.data(), sl.scr, scr, stream)) return -2;
  auto max_rel = [&]() {
    double mx = 0;
    for (int l = 0; l < nrhs; l++) mx = std::max(mx, bn[l] > 0 ? std::sqrt(rn[l] / bn[l]) : std::sqrt(rn[l]));
    return mx;
  };
  for (int it = 0;; it++) {
    if (residual<C>(a, sl.B, sl.Xd, sl.Rr, nrhs, stream) || col_norms2<C>(sl.Rr, sl.n, nrhs, rn.data(), sl.scr, scr,
                                                                          stream))
      return -2;
    *rel = max_rel();
    if (*rel <= tol || *rel != *rel || it >= max_sweeps) break;
    if (convert<double>(sl.Rr, f.m[0].X, tot, stream) || f.solve(S_UPPER) || f.solve(S_LOWER) ||
        accumulate<double>(sl.Xd, f.m[0].X, tot, stream))
      return -2;
  }
  FDM_HIP(hipStreamSynchronize(stream));
  return 0;
}

template <class R, bool C>
int Solver<R, C>::snap_factors(int count, const int *ids) {
  FDM_HIP(hipStreamSynchronize(stream));
  snap_factor.resize(slots.size());
  snap_d.assign(slots.size(), {});
  snap_x.assign(slots.size(), {});
  const size_t te = (size_t)TB * TB * NC * sizeof(R);
  for (int k = 0; k < count; k++) {
    const int g = ids[k];
    const Plan &pl = *slots[g].plan;
    const size_t bt = te * (size_t)pl.ntiles, bd = te * DSLOTS * (size_t)pl.NB;
    snap_factor[g].resize(bt + bd);
    FDM_HIP(hipMemcpy(snap_factor[g].data(), grp.m[g].tiles, bt, hipMemcpyDeviceToHost));
    FDM_HIP(hipMemcpy(snap_factor[g].data() + bt, grp.m[g].dinv, bd, hipMemcpyDeviceToHost));
  }
  return 0;
}

template <class R, bool C>
int Solver<R, C>::snap_sweep(const std::vector<int> &slot_ids) {
  FDM_HIP(hipStreamSynchronize(stream));
  for (int g : slot_ids) {
    const size_t tot = (size_t)slots[g].plan->NB * TB * nrhs * NC;
    std::vector<char> d(tot * sizeof(R)), x(tot * sizeof(double));
    FDM_HIP(hipMemcpy(d.data(), grp.m[g].X, d.size(), hipMemcpyDeviceToHost));
    FDM_HIP(hipMemcpy(x.data(), slots[g].Xd, x.size(), hipMemcpyDeviceToHost));
    snap_d[g].push_back(std::move(d));
    snap_x[g].push_back(std::move(x));
  }
  return 0;
}

template <class R, bool C>
int Solver<R, C>::solve(int count, const int *ids, const double *const *vals, const double *const *b,
                        double *const *x, int max_sweeps, double tol, int *fail_rows, int *sweeps, double *rel) {
  if (finalize()) return -2;
  if (set_active(count, ids)) return -1;
  uint32_t mask = 0;
  for (int k = 0; k < count; k++) mask |= 1u << ids[k];
  // stage the inputs into the slots' fixed buffers (the graph reads those)
  for (int k = 0; k < count; k++) {
    Slot &sl = slots[ids[k]];
    FDM_HIP(hipMemcpyAsync(sl.vals, vals[k], sizeof(double) * sl.nnz * NC, hipMemcpyDeviceToDevice, stream));
    FDM_HIP(hipMemcpyAsync(sl.B, b[k], sizeof(double) * (size_t)sl.n * nrhs * NC, hipMemcpyDeviceToDevice, stream));
  }
  if (factor_phase(mask)) return -2;
  std::vector<int> fr;
  if (grp.read_fail(fr)) return -2;
  if (snapshots && snap_factors(count, ids)) return -2;
  last_sweeps.assign(count, 0);
  last_rel.assign(count, 0.0);
  constexpr bool F64 = std::is_same<R, double>::value;
  if (F64 && max_sweeps <= 0) {  // the direct FP64 solve only
    for (int k = 0; k < count; k++) {
      const int g = ids[k];
      fail_rows[k] = fr[g];
      if (sweeps) sweeps[k] = 0;
      if (rel) rel[k] = 0;
      hipLaunchKernelGGL((k_store_x<R>), dim3(1024), dim3(256), 0, stream, x[k], (const R *)grp.m[g].X,
                         (size_t)slots[g].n * nrhs * NC);
    }
    FDM_HIP(hipGetLastError());
    FDM_HIP(hipStreamSynchronize(stream));
    return 0;
  }
  // Refinement in FP64 over the members whose factorisation held.  FP32
  // factors: x = 0, r = b, and every sweep (the first is the direct solve)
  // solves with the FP32 factors.  FP64 factors: the direct solve already
  // ran with the factorisation; x = that solution, r = b - A x, and the
  // sweeps are corrections with the FP64 factors.
  std::vector<int> act;
  for (int k = 0; k < count; k++)
    if (fr[ids[k]] == 0) act.push_back(k);
  if (!hnorm) FDM_HIP(hipHostMalloc(&hnorm, sizeof(double) * 2 * MAXG * NORM_GRID * std::max(nrhs, 1)));
  const size_t np = (size_t)NORM_GRID * nrhs;
  double *hres = hnorm + (size_t)MAXG * np;  // second half: residual norms of the direct FP64 solve
  std::vector<std::vector<double>> bn(count, std::vector<double>(nrhs)), rn(count, std::vector<double>(nrhs));
  std::vector<double> r(count, INFINITY);
  std::vector<int> sw(count, 0);
  auto max_rel = [&](int k) {
    double mx = 0;
    for (int l = 0; l < nrhs; l++)
      mx = std::max(mx, bn[k][l] > 0 ? std::sqrt(rn[k][l] / bn[k][l]) : std::sqrt(rn[k][l]));
    return mx;
  };
  for (int k : act) {
    Slot &sl = slots[ids[k]];
    const size_t tot = (size_t)sl.plan->NB * TB * nrhs * NC;
    if (launch_norms2<C>(sl.B, sl.n, nrhs, sl.scr, stream)) return -2;
    FDM_HIP(hipMemcpyAsync(hnorm + k * np, sl.scr, sizeof(double) * np, hipMemcpyDeviceToHost, stream));
    if (F64) {
      FDM_HIP(hipMemcpyAsync(sl.Xd, grp.m[ids[k]].X, sizeof(double) * tot, hipMemcpyDeviceToDevice, stream));
      DevCsr a{sl.n, sl.nnz, sl.row_ptr, sl.col, sl.vals, sl.long_rows, sl.nlong};
      if (residual<C>(a, sl.B, sl.Xd, sl.Rr, nrhs, stream) || launch_norms2<C>(sl.Rr, sl.n, nrhs, sl.scr, stream))
        return -2;
      FDM_HIP(hipMemcpyAsync(hres + k * np, sl.scr, sizeof(double) * np, hipMemcpyDeviceToHost, stream));
    } else {
      FDM_HIP(hipMemsetAsync(sl.Xd, 0, sizeof(double) * tot, stream));
      FDM_HIP(hipMemcpyAsync(sl.Rr, sl.B, sizeof(double) * tot, hipMemcpyDeviceToDevice, stream));
    }
  }
  FDM_HIP(hipStreamSynchronize(stream));
  for (int k : act) sum_norm_partials(hnorm + k * np, nrhs, bn[k].data());
  if (F64) {
    std::vector<int> keep;
    for (int k : act) {
      sum_norm_partials(hres + k * np, nrhs, rn[k].data());
      r[k] = max_rel(k);
      if (!(r[k] <= tol) && r[k] == r[k]) keep.push_back(k);
    }
    act.swap(keep);
  }
  for (int it = F64 ? 1 : 0; !act.empty(); it++) {
    std::vector<int> ids_act;
    uint32_t m = 0;
    for (int k : act) {
      ids_act.push_back(ids[k]);
      m |= 1u << ids[k];
    }
    if (set_active((int)ids_act.size(), ids_act.data())) return -1;
    if (sweep_phase(m)) return -2;
    for (int k : act)
      FDM_HIP(hipMemcpyAsync(hnorm + k * np, slots[ids[k]].scr, sizeof(double) * np, hipMemcpyDeviceToHost, stream));
    FDM_HIP(hipStreamSynchronize(stream));
    if (snapshots && snap_sweep(ids_act)) return -2;
    for (int k : act) {
      sum_norm_partials(hnorm + k * np, nrhs, rn[k].data());
      r[k] = max_rel(k);
      sw[k] = it;
    }
    std::vector<int> keep;
    for (int k : act)
      if (!(r[k] <= tol) && r[k] == r[k] && it < max_sweeps) keep.push_back(k);
    act.swap(keep);
  }
  if (F64) {  // no fallback: the refined FP64 solution, with its last residual
    for (int k = 0; k < count; k++) {
      const int g = ids[k];
      fail_rows[k] = fr[g];
      if (fr[g] == 0)
        FDM_HIP(hipMemcpyAsync(x[k], slots[g].Xd, sizeof(double) * (size_t)slots[g].n * nrhs * NC,
                               hipMemcpyDeviceToDevice, stream));
      if (sweeps) sweeps[k] = sw[k];
      if (rel) rel[k] = fr[g] == 0 ? r[k] : NAN;
    }
    FDM_HIP(hipStreamSynchronize(stream));
    if (grp.check_solve()) return -2;
    last_sweeps = sw;
    last_rel = r;
    return 0;
  }
  // members that did not converge (or whose FP32 pivots failed) fall back to FP64
  for (int k = 0; k < count; k++) {
    const int g = ids[k];
    int frow = 0;
    if (fr[g] != 0 || !(r[k] <= tol)) {
      if (fallback_fp64(g, &frow, max_sweeps, tol, &r[k])) return -2;
      sw[k] = -(std::max(sw[k], 1));
    }
    fail_rows[k] = frow;
    if (frow == 0)
      FDM_HIP(hipMemcpyAsync(x[k], slots[g].Xd, sizeof(double) * (size_t)slots[g].n * nrhs * NC,
                             hipMemcpyDeviceToDevice, stream));
    if (sweeps) sweeps[k] = sw[k];
    if (rel) rel[k] = r[k];
  }
  FDM_HIP(hipStreamSynchronize(stream));
  if (grp.check_solve()) return -2;
  last_sweeps = sw;
  last_rel = r;
  return 0;
}

template <class R, bool C>
int Solver<R, C>::adjoint(int count, const int *ids, const double *const *src, const double *const *x,
                          double *const *grad) {
  if (!std::is_same<R, double>::value || !C) {
    set_error("the adjoint needs FP64 complex factors");
    return -1;
  }
  if (set_active(count, ids)) return -1;
  for (int k = 0; k < count; k++) {
    const int g = ids[k];
    const size_t nb = (size_t)slots[g].n * nrhs * NC, tot = (size_t)slots[g].plan->NB * TB * nrhs * NC;
    hipLaunchKernelGGL((k_load_rhs<R>), dim3(1024), dim3(256), 0, stream, grp.m[g].X, src[k], nb, tot);
  }
  // A^T = L^T U^T: solve L^T z = s, then U^T lam = z, with the stored factors.
  if (grp.solve(S_LOWER_T) || grp.solve(S_UPPER_T)) return -2;
  for (int k = 0; k < count; k++) {
    const int g = ids[k];
    hipLaunchKernelGGL((k_adjoint_grad<C>), dim3((slots[g].n + 255) / 256), dim3(256), 0, stream,
                       (const double *)grp.m[g].X, x[k], slots[g].n, nrhs, grad[k]);
  }
  FDM_HIP(hipGetLastError());
  FDM_HIP(hipStreamSynchronize(stream));
  if (grp.check_solve()) return -2;
  return 0;
}

template <class R, bool C>
int Solver<R, C>::sigma_into(int count, const int *ids, const double *const *src, const double *const *x,
                             double *part, double *d_out) {
  if (count < 1 || count > MAXG) {
    set_error("sigma: 1 to " + std::to_string(MAXG) + " points per call");
    return -1;
  }
  for (int k = 0; k < count; k++)
    if (ids[k] < 0 || ids[k] >= (int)slots.size() || !src[k] || !x[k]) {
      set_error("sigma: bad slot id or null buffer for point " + std::to_string(k));
      return -1;
    }
  for (int k = 0; k < count; k++) {
    const int g = ids[k];
    hipLaunchKernelGGL((k_sigma_part<C>), dim3(SIG_PARTS), dim3(256), 0, stream, src[k], x[k],
                       (size_t)slots[g].n * nrhs, part + (size_t)k * SIG_PARTS);
    hipLaunchKernelGGL(k_sigma_final, dim3(1), dim3(256), 0, stream, part + (size_t)k * SIG_PARTS, d_out + k);
  }
  FDM_HIP(hipGetLastError());
  return 0;
}

template <class R, bool C>
int Solver<R, C>::sigma(int count, const int *ids, const double *const *src, const double *const *x, double *out) {
  // partials and results of up to MAXG points, allocated on first use
  if (!sig_buf) FDM_HIP(hipMalloc(&sig_buf, sizeof(double) * MAXG * (SIG_PARTS + 1)));
  if (!hsig) FDM_HIP(hipHostMalloc(&hsig, sizeof(double) * MAXG));
  double *d = sig_buf + MAXG * SIG_PARTS;
  const int rc = sigma_into(count, ids, src, x, sig_buf, d);
  if (rc) return rc;
  hipError_t e = hipMemcpyAsync(hsig, d, sizeof(double) * count, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e == hipSuccess)
    for (int k = 0; k < count; k++) out[k] = hsig[k];
  if (e != hipSuccess) {
    set_error(std::string("sigma: ") + hipGetErrorString(e));
    return -2;
  }
  return 0;
}

template <class R, bool C>
void Solver<R, C>::drop_graphs() {
  for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
  for (auto &kv : sweeps_g) (void)hipGraphExecDestroy(kv.second);
  graphs.clear();
  sweeps_g.clear();
  fb.clear();  // rebuilt with the new schedule on their next use
}

template struct Solver<double, true>;
template struct Solver<double, false>;
template struct Solver<float, true>;
template struct Solver<float, false>;

}  // namespace fdmgpu
