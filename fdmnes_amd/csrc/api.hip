// C ABI of libfdmgpu.so (declared in include/fdmgpu.h).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "../../include/fdmgpu.h"
#include "assemble.h"
#include "batch.h"
#include "sphere.h"
#include "tensor.h"
#include "tau.h"

using namespace fdmgpu;

namespace {

struct Lib {
  bool ready = false;
  int device = -1;
  hipStream_t stream = nullptr;
};
Lib g_lib;

// Scratch of fdmgpu_solver_sigma_noalloc: the partial sums of up to MAXG
// points, SIG_PARTS doubles each (64 KiB).  Each device holds its own copy
// until the process ends; host code uses only the device address that
// hipGetSymbolAddress gives for it (per device, taken in fdmgpu_init).
constexpr int SCRATCH_DEVICES = 64;
static __device__ __align__(256) double scratch[MAXG * SIG_PARTS];
double *g_scratch[SCRATCH_DEVICES] = {};

int ensure_init() {
  if (g_lib.ready) {
    FDM_HIP(hipSetDevice(g_lib.device));
    return 0;
  }
  return fdmgpu_init(0);
}

// Correction sweeps of the FP64 drop-in: the blocked factorisation applies
// explicit inverses of its 64 x 64 diagonal tiles, whose rounding on an
// ill-conditioned undamped system (real arithmetic) exceeds the row-by-row
// elimination's; refinement with the same factors restores the accuracy.
constexpr int DROPIN_FP64_SWEEPS = 2;
constexpr double DROPIN_FP64_TOL = 1e-12;

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// Device buffer that frees itself.
template <class T>
struct DBuf {
  T *p = nullptr;
  size_t n = 0;
  int alloc(size_t count) {
    release();
    n = count;
    FDM_HIP(hipMalloc(&p, sizeof(T) * (count ? count : 1)));
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
  }
  ~DBuf() { release(); }
};

int check_csr(int n, const int *lb1, const int *lb2, const int64_t *row_ptr, const int *col) {
  if (row_ptr[0] != 0) {
    set_error("row_ptr[0] must be 0");
    return FDMGPU_EINVAL;
  }
  for (int i = 0; i < n; i++) {
    if (row_ptr[i + 1] < row_ptr[i]) {
      set_error("row_ptr must be non-decreasing");
      return FDMGPU_EINVAL;
    }
    for (int64_t q = row_ptr[i]; q < row_ptr[i + 1]; q++)
      if (col[q] < lb1[i] || col[q] > lb2[i]) {
        set_error("column " + std::to_string(col[q]) + " of row " + std::to_string(i + 1) +
                  " lies outside its profile [lb1, lb2]");
        return FDMGPU_EINVAL;
      }
  }
  return 0;
}

// The drop-in keeps its one-slot device solver between calls: FDMNES calls
// mat_solve once per energy point with the same structure, so the block plan,
// the factor storage, the pattern on the device, the staging buffers and the
// captured launch graphs are reused while the precision, the arithmetic, the
// sizes, the profile and the pattern stay the same (compared exactly).  Any
// change rebuilds it; fdmgpu_finalize and a device change release it.
template <class R, bool C>
struct DropinCache {
  std::unique_ptr<Solver<R, C>> sv;
  int n = 0, nrhs = 0, slot = -1;
  std::vector<int> lb1, lb2, col;
  std::vector<int64_t> row_ptr;
  DBuf<double> dv, db, dx;
  void reset() {
    sv.reset();
    dv.release();
    db.release();
    dx.release();
    n = nrhs = 0;
    slot = -1;
    lb1.clear();
    lb2.clear();
    col.clear();
    row_ptr.clear();
  }
  bool matches(int n_, int nrhs_, const int *l1, const int *l2, const int64_t *rp, const int *cl) const {
    if (!sv || n_ != n || nrhs_ != nrhs || rp[n_] != row_ptr[n]) return false;
    return std::equal(l1, l1 + n_, lb1.begin()) && std::equal(l2, l2 + n_, lb2.begin()) &&
           std::equal(rp, rp + n_ + 1, row_ptr.begin()) && std::equal(cl, cl + rp[n_], col.begin());
  }
};
// heap objects: never destroyed at exit, when the HIP runtime may be gone
template <class R, bool C>
DropinCache<R, C> &dropin_cache() {
  static DropinCache<R, C> *c = new DropinCache<R, C>();
  return *c;
}
void release_dropin_caches() {
  dropin_cache<double, true>().reset();
  dropin_cache<double, false>().reset();
  dropin_cache<float, true>().reset();
  dropin_cache<float, false>().reset();
}

// Host-pointer drop-in on top of the cached one-slot device solver.
template <class R, bool C>
int host_solve(int n, int nrhs, const int *lb1, const int *lb2, const int64_t *row_ptr, const int *col,
               const double *a_re, const double *a_im, double *smr, double *smi, int max_sweeps, double tol,
               int *sweeps, double *final_rel, int *fail_row, double *t_fill, double *t_tria) {
  constexpr int NC = C ? 2 : 1;
  const double t0 = now();
  auto &c = dropin_cache<R, C>();
  const int64_t nnz = row_ptr[n];
  if (!c.matches(n, nrhs, lb1, lb2, row_ptr, col)) {
    c.reset();
    int rc = check_csr(n, lb1, lb2, row_ptr, col);
    if (rc) return rc;
    c.sv.reset(new (std::nothrow) Solver<R, C>());
    if (!c.sv) return FDMGPU_ENOMEM;
    if (c.sv->init(nrhs, g_lib.stream)) {
      c.reset();
      return FDMGPU_EHIP;
    }
    rc = c.sv->add_system(n, lb1, lb2, row_ptr, col, &c.slot);
    if (rc) {
      c.reset();
      return rc == -1 ? FDMGPU_EINVAL : FDMGPU_EHIP;
    }
    if (c.dv.alloc((size_t)std::max<int64_t>(nnz, 1) * NC) || c.db.alloc((size_t)n * nrhs * NC) ||
        c.dx.alloc((size_t)n * nrhs * NC)) {
      c.reset();
      return FDMGPU_ENOMEM;
    }
    c.n = n;
    c.nrhs = nrhs;
    c.lb1.assign(lb1, lb1 + n);
    c.lb2.assign(lb2, lb2 + n);
    c.row_ptr.assign(row_ptr, row_ptr + n + 1);
    c.col.assign(col, col + nnz);
  }
  Solver<R, C> &sv = *c.sv;
  std::vector<double> hv((size_t)std::max<int64_t>(nnz, 1) * NC), hb((size_t)n * nrhs * NC);
  for (int64_t q = 0; q < nnz; q++) {
    hv[q * NC] = a_re[q];
    if (C) hv[q * NC + 1] = a_im[q];
  }
  for (int64_t i = 0; i < (int64_t)n * nrhs; i++) {
    hb[i * NC] = smr[i];
    if (C) hb[i * NC + 1] = smi[i];
  }
  FDM_HIP(hipMemcpy(c.dv.p, hv.data(), sizeof(double) * hv.size(), hipMemcpyHostToDevice));
  FDM_HIP(hipMemcpy(c.db.p, hb.data(), sizeof(double) * hb.size(), hipMemcpyHostToDevice));
  const double t1 = now();
  const double *vp = c.dv.p, *bp = c.db.p;
  double *xp = c.dx.p;
  int fr = 0, sw = 0;
  double rel = 0;
  const int slot = c.slot;
  if (sv.solve(1, &slot, &vp, &bp, &xp, max_sweeps, tol, &fr, &sw, &rel)) {
    c.reset();  // a failed launch may leave the solver in any state
    return FDMGPU_EHIP;
  }
  const double t2 = now();
  if (t_fill) *t_fill += t1 - t0;
  if (t_tria) *t_tria += t2 - t1;
  if (fail_row) *fail_row = fr;
  if (fr) return FDMGPU_PIVOT;
  FDM_HIP(hipMemcpy(hb.data(), c.dx.p, sizeof(double) * hb.size(), hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < (int64_t)n * nrhs; i++) {
    smr[i] = hb[i * NC];
    if (C) smi[i] = hb[i * NC + 1];
  }
  if (sweeps) *sweeps = sw;
  if (final_rel) *final_rel = rel;
  return FDMGPU_OK;
}

}  // namespace

extern "C" {

int fdmgpu_init(int device) {
  if (g_lib.ready && g_lib.device == device) {
    FDM_HIP(hipSetDevice(device));
    return 0;
  }
  int count = 0;
  FDM_HIP(hipGetDeviceCount(&count));
  if (device < 0 || device >= count) {
    set_error("no HIP device " + std::to_string(device));
    return FDMGPU_EINVAL;
  }
  if (g_lib.ready) {  // another device: the cached drop-in solvers belong to the old one
    FDM_HIP(hipSetDevice(g_lib.device));
    release_dropin_caches();
  }
  FDM_HIP(hipSetDevice(device));
  if (device < SCRATCH_DEVICES && !g_scratch[device]) {
    double *p = nullptr;
    const hipError_t e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("fdmgpu_init: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return FDMGPU_EHIP;
    }
    g_scratch[device] = p;
  }
  if (g_lib.stream) (void)hipStreamDestroy(g_lib.stream);
  FDM_HIP(hipStreamCreateWithFlags(&g_lib.stream, hipStreamNonBlocking));
  g_lib.device = device;
  g_lib.ready = true;
  return 0;
}

int fdmgpu_finalize(void) {
  if (g_lib.ready) {
    (void)hipSetDevice(g_lib.device);
    release_dropin_caches();
  }
  if (g_lib.stream) (void)hipStreamDestroy(g_lib.stream);
  g_lib = Lib();
  return 0;
}

const char *fdmgpu_last_error(void) { return get_error(); }

void fdmgpu_solver_name(char out[5]) { memcpy(out, "HIP  ", 5); }

int fdmgpu_mat_solve(int nligne, int nlmso, int cal_comp, const int *lb1, const int *lb2, const int64_t *row_ptr,
                     const int *col, const double *a_re, const double *a_im, double *smr, double *smi,
                     int *fail_row, double *t_fill, double *t_tria) {
  if (nligne <= 0 || nlmso <= 0 || !lb1 || !lb2 || !row_ptr || !col || !a_re || !smr ||
      (cal_comp && (!a_im || !smi))) {
    set_error("fdmgpu_mat_solve: invalid argument");
    return FDMGPU_EINVAL;
  }
  int rc = ensure_init();
  if (rc) return rc;
  // the direct solve, then at most DROPIN_FP64_SWEEPS corrections with the
  // same FP64 factors while max_l ||r_l|| / ||b_l|| > DROPIN_FP64_TOL
  return cal_comp ? host_solve<double, true>(nligne, nlmso, lb1, lb2, row_ptr, col, a_re, a_im, smr, smi,
                                             DROPIN_FP64_SWEEPS, DROPIN_FP64_TOL, nullptr, nullptr, fail_row, t_fill,
                                             t_tria)
                  : host_solve<double, false>(nligne, nlmso, lb1, lb2, row_ptr, col, a_re, a_im, smr, smi,
                                              DROPIN_FP64_SWEEPS, DROPIN_FP64_TOL, nullptr, nullptr, fail_row, t_fill,
                                              t_tria);
}

int fdmgpu_mat_solve_mixed(int nligne, int nlmso, int cal_comp, const int *lb1, const int *lb2,
                           const int64_t *row_ptr, const int *col, const double *a_re, const double *a_im,
                           double *smr, double *smi, int max_sweeps, double tol, int *sweeps, double *final_rel,
                           int *fail_row) {
  if (nligne <= 0 || nlmso <= 0 || nlmso > 256 || !lb1 || !lb2 || !row_ptr || !col || !a_re || !smr ||
      (cal_comp && (!a_im || !smi)) || max_sweeps < 0) {
    set_error("fdmgpu_mat_solve_mixed: invalid argument");
    return FDMGPU_EINVAL;
  }
  int rc = ensure_init();
  if (rc) return rc;
  return cal_comp ? host_solve<float, true>(nligne, nlmso, lb1, lb2, row_ptr, col, a_re, a_im, smr, smi, max_sweeps,
                                            tol, sweeps, final_rel, fail_row, nullptr, nullptr)
                  : host_solve<float, false>(nligne, nlmso, lb1, lb2, row_ptr, col, a_re, a_im, smr, smi,
                                             max_sweeps, tol, sweeps, final_rel, fail_row, nullptr, nullptr);
}

// ---------------------------------------------------------------- device solver
struct fdmgpu_solver {
  int prec = 0, cal = 0;
  void *impl = nullptr;
};

#define FDM_DISPATCH(S, CALL)                                                          \
  ((S)->prec == FDMGPU_FP64                                                            \
       ? ((S)->cal ? ((Solver<double, true> *)(S)->impl)->CALL                         \
                   : ((Solver<double, false> *)(S)->impl)->CALL)                       \
       : ((S)->cal ? ((Solver<float, true> *)(S)->impl)->CALL                          \
                   : ((Solver<float, false> *)(S)->impl)->CALL))

static int map_rc(int rc) { return rc == 0 ? 0 : (rc == -1 ? FDMGPU_EINVAL : FDMGPU_EHIP); }

int fdmgpu_solver_create(int precision, int cal_comp, int nrhs, void *hip_stream, fdmgpu_solver **out) {
  if (!out || nrhs <= 0 || nrhs > 256 || (precision != FDMGPU_FP64 && precision != FDMGPU_FP32)) {
    set_error("fdmgpu_solver_create: invalid argument");
    return FDMGPU_EINVAL;
  }
  int rc = ensure_init();
  if (rc) return rc;
  fdmgpu_solver *s = new fdmgpu_solver();
  s->prec = precision;
  s->cal = cal_comp ? 1 : 0;
  hipStream_t st = hip_stream ? (hipStream_t)hip_stream : g_lib.stream;
  if (precision == FDMGPU_FP64) {
    if (s->cal) s->impl = new Solver<double, true>();
    else s->impl = new Solver<double, false>();
  } else {
    if (s->cal) s->impl = new Solver<float, true>();
    else s->impl = new Solver<float, false>();
  }
  rc = FDM_DISPATCH(s, init(nrhs, st));
  if (rc) {
    fdmgpu_solver_destroy(s);
    return map_rc(rc);
  }
  *out = s;
  return 0;
}

void fdmgpu_solver_destroy(fdmgpu_solver *s) {
  if (!s) return;
  if (s->impl) {
    if (s->prec == FDMGPU_FP64) {
      if (s->cal) delete (Solver<double, true> *)s->impl;
      else delete (Solver<double, false> *)s->impl;
    } else {
      if (s->cal) delete (Solver<float, true> *)s->impl;
      else delete (Solver<float, false> *)s->impl;
    }
  }
  delete s;
}

int fdmgpu_solver_add_system(fdmgpu_solver *s, int nligne, const int *lb1, const int *lb2, const int64_t *row_ptr,
                             const int *col, int *slot) {
  if (!s || nligne <= 0 || !lb1 || !lb2 || !row_ptr || !col || !slot) {
    set_error("fdmgpu_solver_add_system: invalid argument");
    return FDMGPU_EINVAL;
  }
  return map_rc(FDM_DISPATCH(s, add_system(nligne, lb1, lb2, row_ptr, col, slot)));
}

int fdmgpu_solver_solve(fdmgpu_solver *s, int count, const int *slots, const double *const *d_vals,
                        const double *const *d_b, double *const *d_x, int max_sweeps, double tol, int *fail_rows,
                        int *sweeps, double *rel) {
  if (!s || count <= 0 || count > MAXG || !slots || !d_vals || !d_b || !d_x || !fail_rows) {
    set_error("fdmgpu_solver_solve: invalid argument");
    return FDMGPU_EINVAL;
  }
  int rc = map_rc(FDM_DISPATCH(s, solve(count, slots, d_vals, d_b, d_x, max_sweeps, tol, fail_rows, sweeps, rel)));
  if (rc) return rc;
  for (int k = 0; k < count; k++)
    if (fail_rows[k]) return FDMGPU_PIVOT;
  return 0;
}

int fdmgpu_solver_adjoint(fdmgpu_solver *s, int count, const int *slots, const double *const *d_src,
                          const double *const *d_x, double *const *d_grad) {
  if (!s || count <= 0 || count > MAXG || !slots || !d_src || !d_x || !d_grad) {
    set_error("fdmgpu_solver_adjoint: invalid argument");
    return FDMGPU_EINVAL;
  }
  return map_rc(FDM_DISPATCH(s, adjoint(count, slots, d_src, d_x, d_grad)));
}

int fdmgpu_solver_sigma(fdmgpu_solver *s, int count, const int *slots, const double *const *d_src,
                        const double *const *d_x, double *sigma) {
  if (!s || count <= 0 || !slots || !d_src || !d_x || !sigma) {
    set_error("fdmgpu_solver_sigma: invalid argument");
    return FDMGPU_EINVAL;
  }
  return map_rc(FDM_DISPATCH(s, sigma(count, slots, d_src, d_x, sigma)));
}

int fdmgpu_solver_sigma_noalloc(fdmgpu_solver *s, int count, const int *slots, const double *const *d_src,
                                const double *const *d_x, double *d_sigma) {
  if (!s || count <= 0 || !slots || !d_src || !d_x || !d_sigma) {
    set_error("fdmgpu_solver_sigma_noalloc: invalid argument");
    return FDMGPU_EINVAL;
  }
  if (count > MAXG) {
    set_error("fdmgpu_solver_sigma_noalloc: at most " + std::to_string(MAXG) +
              " points per call (the scratch holds their partial sums)");
    return FDMGPU_EINVAL;
  }
  // the solver exists, so the library was initialised on this device
  if (!g_lib.ready || g_lib.device >= SCRATCH_DEVICES || !g_scratch[g_lib.device]) {
    set_error("fdmgpu_solver_sigma_noalloc: library not initialised on a device with a scratch address");
    return FDMGPU_ESTATE;
  }
  FDM_HIP(hipSetDevice(g_lib.device));
  return map_rc(FDM_DISPATCH(s, sigma_into(count, slots, d_src, d_x, g_scratch[g_lib.device], d_sigma)));
}

int fdmgpu_solver_options(fdmgpu_solver *s, int use_graphs, int time_updates, int persist_wg) {
  if (!s) return FDMGPU_EINVAL;
  auto apply = [&](auto *p) {
    // captured graphs hold the old grid sizes: rebuild them on a change
    if ((use_graphs != 0) != p->use_graphs || (persist_wg > 0 && persist_wg != p->grp.persist_wg)) p->drop_graphs();
    p->use_graphs = use_graphs != 0;
    p->grp.time_updates = time_updates != 0;
    if (persist_wg > 0) p->grp.persist_wg = persist_wg;
  };
  if (s->prec == FDMGPU_FP64) {
    if (s->cal) apply((Solver<double, true> *)s->impl);
    else apply((Solver<double, false> *)s->impl);
  } else {
    if (s->cal) apply((Solver<float, true> *)s->impl);
    else apply((Solver<float, false> *)s->impl);
  }
  return 0;
}

int fdmgpu_solver_overlap(fdmgpu_solver *s, int on) {
  if (!s || on < 0 || on > 1) {
    set_error("fdmgpu_solver_overlap: invalid argument");
    return FDMGPU_EINVAL;
  }
  auto apply = [&](auto *p) {
    p->grp.overlap = on;
    p->drop_graphs();
  };
  if (s->prec == FDMGPU_FP64) {
    if (s->cal) apply((Solver<double, true> *)s->impl);
    else apply((Solver<double, false> *)s->impl);
  } else {
    if (s->cal) apply((Solver<float, true> *)s->impl);
    else apply((Solver<float, false> *)s->impl);
  }
  return 0;
}

int fdmgpu_solver_tune(fdmgpu_solver *s, int gk, int trail_wg, int persist_solve, int solve_wg, int trail_tile,
                       int lookahead) {
  if (!s || gk > MAXGK || (trail_tile > 0 && trail_tile != 64 && trail_tile != 128)) {
    set_error("fdmgpu_solver_tune: gk must be <= 32 and trail_tile 64 or 128");
    return FDMGPU_EINVAL;
  }
  auto apply = [&](auto *p) {
    if (gk >= 0) p->grp.gk = gk;
    if (trail_tile > 0) p->grp.trail_tile = trail_tile;
    if (trail_wg > 0) p->grp.persist_wg = trail_wg;
    if (persist_solve >= 0) p->grp.persist_solve = persist_solve != 0;
    if (solve_wg > 0) p->grp.solve_wg = solve_wg;
    if (lookahead >= 0) p->grp.lookahead = lookahead;
    p->drop_graphs();
  };
  if (s->prec == FDMGPU_FP64) {
    if (s->cal) apply((Solver<double, true> *)s->impl);
    else apply((Solver<double, false> *)s->impl);
  } else {
    if (s->cal) apply((Solver<float, true> *)s->impl);
    else apply((Solver<float, false> *)s->impl);
  }
  return 0;
}

int fdmgpu_solver_stats(const fdmgpu_solver *s, double *out) {
  if (!s || !out) return FDMGPU_EINVAL;
  auto fill = [&](auto *p) {
    out[0] = p->grp.update_ms;
    out[1] = (double)p->grp.update_launches;
    double tm = 0, fm = 0, sm = 0, om = 0;
    for (auto &sl : p->slots) {
      tm += sl.plan->tile_update_macs();
      om += sl.plan->tile_other_macs();
      fm += sl.factor_macs;
      sm += sl.solve_macs;
    }
    out[2] = tm;
    out[3] = fm;
    out[4] = (double)p->grp.factor_bytes();
    out[5] = sm;
    out[6] = om;
    out[7] = p->grp.update_bytes;
  };
  if (s->prec == FDMGPU_FP64) {
    if (s->cal) fill((Solver<double, true> *)s->impl);
    else fill((Solver<double, false> *)s->impl);
  } else {
    if (s->cal) fill((Solver<float, true> *)s->impl);
    else fill((Solver<float, false> *)s->impl);
  }
  return 0;
}

int fdmgpu_solver_snapshots(fdmgpu_solver *s, int on) {
  if (!s || on < 0 || on > 1) {
    set_error("fdmgpu_solver_snapshots: invalid argument");
    return FDMGPU_EINVAL;
  }
  auto apply = [&](auto *p) { p->snapshots = on != 0; };
  if (s->prec == FDMGPU_FP64) {
    if (s->cal) apply((Solver<double, true> *)s->impl);
    else apply((Solver<double, false> *)s->impl);
  } else {
    if (s->cal) apply((Solver<float, true> *)s->impl);
    else apply((Solver<float, false> *)s->impl);
  }
  return 0;
}

int fdmgpu_solver_snapshot(const fdmgpu_solver *s, int slot, int what, int sweep, void *out, int64_t capacity,
                           int64_t *bytes) {
  if (!s || !bytes || what < 0 || what > 3 || slot < 0) {
    set_error("fdmgpu_solver_snapshot: invalid argument");
    return FDMGPU_EINVAL;
  }
  auto get = [&](auto *p) -> int {
    if (slot >= (int)p->snap_factor.size() || p->snap_factor[slot].empty()) {
      set_error("fdmgpu_solver_snapshot: no snapshot of this slot (snapshots off, or the slot was not solved)");
      return FDMGPU_ESTATE;
    }
    const int nsw = (int)p->snap_d[slot].size();
    if (what == 3) {
      *bytes = nsw;
      return 0;
    }
    if (what > 0 && (sweep < 0 || sweep >= nsw)) {
      set_error("fdmgpu_solver_snapshot: sweep out of range");
      return FDMGPU_EINVAL;
    }
    const std::vector<char> &v = what == 0 ? p->snap_factor[slot] : (what == 1 ? p->snap_d[slot][sweep]
                                                                                 : p->snap_x[slot][sweep]);
    *bytes = (int64_t)v.size();
    if (out) {
      if (capacity < (int64_t)v.size()) {
        set_error("fdmgpu_solver_snapshot: buffer too small");
        return FDMGPU_EINVAL;
      }
      memcpy(out, v.data(), v.size());
    }
    return 0;
  };
  if (s->prec == FDMGPU_FP64) return s->cal ? get((Solver<double, true> *)s->impl) : get((Solver<double, false> *)s->impl);
  return s->cal ? get((Solver<float, true> *)s->impl) : get((Solver<float, false> *)s->impl);
}

struct fdmgpu_assembler {
  Assembler a;
};

int fdmgpu_assembler_create_atoms(const fdmgpu_assembly_tables *t, const fdmgpu_atom_tables *at, int max_points,
                                  void *hip_stream, fdmgpu_assembler **out) {
  if (!t || !out) {
    set_error("fdmgpu_assembler_create: invalid argument");
    return FDMGPU_EINVAL;
  }
  int rc = ensure_init();
  if (rc) return rc;
  std::unique_ptr<fdmgpu_assembler> h(new (std::nothrow) fdmgpu_assembler());
  if (!h) return FDMGPU_ENOMEM;
  rc = h->a.init(t, at, max_points, hip_stream ? (hipStream_t)hip_stream : g_lib.stream);
  if (rc) return rc;
  *out = h.release();
  return 0;
}

int fdmgpu_assembler_create(const fdmgpu_assembly_tables *t, int max_points, void *hip_stream,
                            fdmgpu_assembler **out) {
  return fdmgpu_assembler_create_atoms(t, nullptr, max_points, hip_stream, out);
}

void fdmgpu_assembler_destroy(fdmgpu_assembler *a) { delete a; }

int fdmgpu_assembler_set_relativistic(fdmgpu_assembler *a, const fdmgpu_rel_tables *t) {
  if (!a || !t) {
    set_error("fdmgpu_assembler_set_relativistic: invalid argument");
    return FDMGPU_EINVAL;
  }
  FDM_HIP(hipSetDevice(g_lib.device));
  return a->a.init_rel(t);
}

int fdmgpu_assembler_set_complex(fdmgpu_assembler *a, const double *base_im, const double *ow_im,
                                 const double *ylm_im, const double *dself, const double *dself_im) {
  if (!a) {
    set_error("fdmgpu_assembler_set_complex: invalid argument");
    return FDMGPU_EINVAL;
  }
  FDM_HIP(hipSetDevice(g_lib.device));
  return a->a.init_complex(base_im, ow_im, ylm_im, dself, dself_im);
}

int fdmgpu_assembler_run_atoms(fdmgpu_assembler *a, int count, const double *Enervide, const double *Eimag,
                               const double *Ecinetic_out, int E_comp, int basereel, int update_only,
                               const double *const *d_V, const double *const *d_phi, int phi_cplx,
                               double *const *d_vals, double *const *d_rhs) {
  return fdmgpu_assembler_run_rel(a, count, Enervide, Eimag, Ecinetic_out, E_comp, basereel, update_only, d_V,
                                  nullptr, d_phi, phi_cplx, d_vals, d_rhs);
}

static int assembler_run(fdmgpu_assembler *a, int count, const double *Enervide, const double *Eimag,
                         const double *Ecinetic_out, int E_comp, int basereel, int update_only,
                         const double *const *d_V, const double *const *d_G, const double *const *d_phi,
                         const double *const *d_phi2, int phi_cplx, double *const *d_vals, double *const *d_rhs);

int fdmgpu_assembler_run_rel(fdmgpu_assembler *a, int count, const double *Enervide, const double *Eimag,
                             const double *Ecinetic_out, int E_comp, int basereel, int update_only,
                             const double *const *d_V, const double *const *d_G, const double *const *d_phi,
                             int phi_cplx, double *const *d_vals, double *const *d_rhs) {
  if (a && a->a.nri == 2) {
    set_error("fdmgpu_assembler_run: the irregular solution's assembler runs by fdmgpu_assembler_run_ri");
    return FDMGPU_EINVAL;
  }
  return assembler_run(a, count, Enervide, Eimag, Ecinetic_out, E_comp, basereel, update_only, d_V, d_G, d_phi,
                       nullptr, phi_cplx, d_vals, d_rhs);
}

int fdmgpu_assembler_run_ri(fdmgpu_assembler *a, int count, const double *Enervide, const double *Eimag,
                            const double *Ecinetic_out, int E_comp, int basereel, int update_only,
                            const double *const *d_V, const double *const *d_phi, const double *const *d_phi2,
                            int phi_cplx, double *const *d_vals, double *const *d_rhs) {
  if (!a || a->a.nri != 2 || !d_phi2 || !d_phi) {
    set_error("fdmgpu_assembler_run_ri: an assembler of the irregular solution and both phiato");
    return FDMGPU_EINVAL;
  }
  for (int k = 0; k < count && k < ASM_MAXP; k++)
    if (!d_phi2[k]) {
      set_error("fdmgpu_assembler_run_ri: null irregular phiato for point " + std::to_string(k));
      return FDMGPU_EINVAL;
    }
  return assembler_run(a, count, Enervide, Eimag, Ecinetic_out, E_comp, basereel, update_only, d_V, nullptr, d_phi,
                       d_phi2, phi_cplx, d_vals, d_rhs);
}

static int assembler_run(fdmgpu_assembler *a, int count, const double *Enervide, const double *Eimag,
                         const double *Ecinetic_out, int E_comp, int basereel, int update_only,
                         const double *const *d_V, const double *const *d_G, const double *const *d_phi,
                         const double *const *d_phi2, int phi_cplx, double *const *d_vals, double *const *d_rhs) {
  if (!a || !Enervide || !Eimag || !Ecinetic_out || !d_V || !d_vals || count < 1 || count > ASM_MAXP) {
    set_error("fdmgpu_assembler_run: invalid argument");
    return FDMGPU_EINVAL;
  }
  if ((a->a.natome > 0) != (d_phi != nullptr)) {
    set_error(a->a.natome > 0 ? "fdmgpu_assembler_run: a cluster with atoms needs phiato (fdmgpu_assembler_run_atoms)"
                              : "fdmgpu_assembler_run_atoms: the assembler holds no atoms");
    return FDMGPU_EINVAL;
  }
  AsmPoints P{};
  for (int k = 0; k < count; k++) {
    if (!d_V[k] || !d_vals[k] || (d_phi && !d_phi[k]) || (d_G && !d_G[k])) {
      set_error("fdmgpu_assembler_run: null device buffer for point " + std::to_string(k));
      return FDMGPU_EINVAL;
    }
    P.G[k] = d_G ? d_G[k] : nullptr;
    P.phi2[k] = d_phi2 ? d_phi2[k] : nullptr;
    P.e[k] = {Enervide[k], Eimag[k], Ecinetic_out[k]};
    P.V[k] = d_V[k];
    P.vals[k] = d_vals[k];
    P.rhs[k] = d_rhs ? d_rhs[k] : nullptr;
    P.phi[k] = d_phi ? d_phi[k] : nullptr;
  }
  FDM_HIP(hipSetDevice(g_lib.device));
  return a->a.run(count, P, E_comp != 0, basereel != 0, update_only != 0, phi_cplx != 0);
}

int fdmgpu_assembler_run(fdmgpu_assembler *a, int count, const double *Enervide, const double *Eimag,
                         const double *Ecinetic_out, int E_comp, int basereel, int update_only,
                         const double *const *d_V, double *const *d_vals, double *const *d_rhs) {
  return fdmgpu_assembler_run_atoms(a, count, Enervide, Eimag, Ecinetic_out, E_comp, basereel, update_only, d_V,
                                    nullptr, 0, d_vals, d_rhs);
}

struct fdmgpu_sphere {
  SphereSolver s;
};
struct fdmgpu_tensor {
  TensorPlan p;
};

int fdmgpu_tensor_create(const fdmgpu_tensor_tables *t, int max_points, void *hip_stream, fdmgpu_tensor **out) {
  if (!t || !out) {
    set_error("fdmgpu_tensor_create: invalid argument");
    return FDMGPU_EINVAL;
  }
  int rc = ensure_init();
  if (rc) return rc;
  std::unique_ptr<fdmgpu_tensor> h(new (std::nothrow) fdmgpu_tensor());
  if (!h) return FDMGPU_ENOMEM;
  rc = h->p.init(t, max_points, hip_stream ? (hipStream_t)hip_stream : g_lib.stream);
  if (rc) return rc;
  *out = h.release();
  return 0;
}

void fdmgpu_tensor_destroy(fdmgpu_tensor *h) { delete h; }

int fdmgpu_tensor_ninitr(const fdmgpu_tensor *h) { return h ? h->p.ninitr : -1; }

int fdmgpu_tensor_run(fdmgpu_tensor *h, int count, const double *const *d_rof, const double *const *d_tau,
                      double *const *d_secdd, double *const *d_secdq, double *const *d_secqq) {
  if (!h || !d_rof || !d_tau || count < 1 || count > TENS_MAXP) {
    set_error("fdmgpu_tensor_run: invalid argument");
    return FDMGPU_EINVAL;
  }
  TensPoints P{};
  for (int k = 0; k < count; k++) {
    P.rof[k] = d_rof[k];
    P.taull[k] = d_tau[k];
    P.dd[k] = d_secdd ? d_secdd[k] : nullptr;
    P.dq[k] = d_secdq ? d_secdq[k] : nullptr;
    P.qq[k] = d_secqq ? d_secqq[k] : nullptr;
  }
  FDM_HIP(hipSetDevice(g_lib.device));
  return h->p.run(count, P);
}

int fdmgpu_sphere_create(const fdmgpu_sphere_tables *t, int max_points, void *hip_stream, fdmgpu_sphere **out) {
  if (!t || !out) {
    set_error("fdmgpu_sphere_create: invalid argument");
    return FDMGPU_EINVAL;
  }
  int rc = ensure_init();
  if (rc) return rc;
  std::unique_ptr<fdmgpu_sphere> h(new (std::nothrow) fdmgpu_sphere());
  if (!h) return FDMGPU_ENOMEM;
  rc = h->s.init(t, max_points, hip_stream ? (hipStream_t)hip_stream : g_lib.stream);
  if (rc) return rc;
  *out = h.release();
  return 0;
}

void fdmgpu_sphere_destroy(fdmgpu_sphere *s) { delete s; }

int fdmgpu_sphere_set_integrals(fdmgpu_sphere *s, int atom, int nbseuil, const double *psii, const double *Eseuil,
                                double Rmtsd, int ip0, int ip_max) {
  if (!s) {
    set_error("fdmgpu_sphere_set_integrals: invalid argument");
    return FDMGPU_EINVAL;
  }
  FDM_HIP(hipSetDevice(g_lib.device));
  return s->s.set_integrals(atom, nbseuil, psii, Eseuil, Rmtsd, ip0, ip_max);
}

int fdmgpu_sphere_run(fdmgpu_sphere *s, int count, const double *Ecinetic, const double *Eimag,
                      const double *Enervide, int complex_out, double *const *d_phi, double *const *d_tau) {
  return fdmgpu_sphere_run_integrals(s, count, Ecinetic, Eimag, Enervide, nullptr, complex_out, d_phi, d_tau, nullptr);
}

static int sphere_run(fdmgpu_sphere *s, int count, const double *Ecinetic, const double *Eimag,
                      const double *Enervide, const double *Energ, int complex_out, double *const *d_phi,
                      double *const *d_tau, double *const *d_rof, double *const *d_phi2);

int fdmgpu_sphere_run_integrals(fdmgpu_sphere *s, int count, const double *Ecinetic, const double *Eimag,
                                const double *Enervide, const double *Energ, int complex_out, double *const *d_phi,
                                double *const *d_tau, double *const *d_rof) {
  return sphere_run(s, count, Ecinetic, Eimag, Enervide, Energ, complex_out, d_phi, d_tau, d_rof, nullptr);
}

int fdmgpu_sphere_run_ri(fdmgpu_sphere *s, int count, const double *Ecinetic, const double *Eimag,
                         const double *Enervide, int complex_out, double *const *d_phi, double *const *d_phi2,
                         double *const *d_tau) {
  if (!d_phi2) {
    set_error("fdmgpu_sphere_run_ri: invalid argument");
    return FDMGPU_EINVAL;
  }
  return sphere_run(s, count, Ecinetic, Eimag, Enervide, nullptr, complex_out, d_phi, d_tau, nullptr, d_phi2);
}

static int sphere_run(fdmgpu_sphere *s, int count, const double *Ecinetic, const double *Eimag,
                      const double *Enervide, const double *Energ, int complex_out, double *const *d_phi,
                      double *const *d_tau, double *const *d_rof, double *const *d_phi2) {
  if (!s || !Ecinetic || !Eimag || !Enervide || !d_phi || count < 1 || count > SPH_MAXP || (d_rof && !Energ)) {
    set_error("fdmgpu_sphere_run: invalid argument");
    return FDMGPU_EINVAL;
  }
  SphPoints P{};
  for (int k = 0; k < count; k++) {
    if (!d_phi[k] && s->s.nphi > 0) {  // atoms without border points have no phiato
      set_error("fdmgpu_sphere_run: null phiato buffer for point " + std::to_string(k));
      return FDMGPU_EINVAL;
    }
    if (!complex_out && (std::fabs(Eimag[k]) > 1e-10 || Ecinetic[k] < 1e-10)) {
      set_error("fdmgpu_sphere_run: complex radial solutions (Eimag != 0 or Ecinetic < 1e-10) need complex_out");
      return FDMGPU_EINVAL;
    }
    P.ecin[k] = Ecinetic[k];
    P.eimag[k] = Eimag[k];
    P.ener[k] = Enervide[k];
    P.energ[k] = Energ ? Energ[k] : 0.0;
    P.phi[k] = d_phi[k];
    P.tau[k] = d_tau ? d_tau[k] : nullptr;
    P.rof[k] = d_rof ? d_rof[k] : nullptr;
    P.phi2[k] = d_phi2 ? d_phi2[k] : nullptr;
  }
  FDM_HIP(hipSetDevice(g_lib.device));
  return s->s.run(count, P, complex_out != 0);
}

int fdmgpu_tau_reduce(const fdmgpu_tau_desc *d, int count, const double *const *d_x, double *const *d_tau,
                      double *const *d_norm, int accumulate, int *d_info, void *hip_stream) {
  if (!d || count < 0 || (count > 0 && !d_x) || d->natome < 0 || (d->natome > 0 && (!d->ianew || !d->nlmsa)) ||
      (d->natome > 0 && count > 0 && !d_tau)) {
    set_error("fdmgpu_tau_reduce: invalid argument");
    return FDMGPU_EINVAL;
  }
  const int n = d->nlmso;
  if (n < 1 || n > TAU_NT || d->nligne < n || (d->basereel && (d->nligneso < 0 || d->nligneso + n > d->nligne))) {
    set_error("fdmgpu_tau_reduce: need 1 <= nlmso <= 256 and the outer-sphere rows inside the system");
    return FDMGPU_EINVAL;
  }
  int na_max = 0;
  for (int a = 0; a < d->natome; a++) {
    if (d->nlmsa[a] < 0 || d->ianew[a] < 0 || d->ianew[a] + d->nlmsa[a] > d->nligne) {
      set_error("fdmgpu_tau_reduce: rows of atom " + std::to_string(a) + " outside the system");
      return FDMGPU_EINVAL;
    }
    na_max = std::max(na_max, d->nlmsa[a]);
  }
  if (tau_lds_bytes(n, na_max) > TAU_MAX_LDS) {
    set_error("fdmgpu_tau_reduce: nlmso^2 + nlmso*(max nlmsa + 1) complex entries exceed the 160 KiB LDS");
    return FDMGPU_EINVAL;
  }
  for (int k = 0; k < count; k++)
    if (!d_x[k] || (d->natome > 0 && !d_tau[k])) {
      set_error("fdmgpu_tau_reduce: null device buffer for point " + std::to_string(k));
      return FDMGPU_EINVAL;
    }
  if (count == 0) return 0;
  int rc = ensure_init();
  if (rc) return rc;
  hipStream_t s = hip_stream ? (hipStream_t)hip_stream : g_lib.stream;
  // every point's norm and info are written by its first atom chunk; with no
  // atoms one launch of one (empty) atom still inverts
  const int natome = d->natome;
  for (int p0 = 0; p0 < count; p0 += TAU_MAXP) {
    const int np = std::min(TAU_MAXP, count - p0);
    int64_t off = 0;
    for (int a0 = 0; a0 < std::max(natome, 1); a0 += TAU_MAXA) {
      const int na = natome ? std::min(TAU_MAXA, natome - a0) : 1;
      TauJob J{};
      for (int k = 0; k < np; k++) {
        J.x[k] = d_x[p0 + k];
        J.tau[k] = natome ? d_tau[p0 + k] : nullptr;
        J.norm[k] = (a0 == 0 && d_norm) ? d_norm[p0 + k] : nullptr;
      }
      J.info = (a0 == 0 && d_info) ? d_info + p0 : nullptr;
      int chunk_max = 0;
      for (int a = 0; a < na; a++) {
        J.row[a] = natome ? d->ianew[a0 + a] : 0;
        J.na[a] = natome ? d->nlmsa[a0 + a] : 0;
        J.off[a] = off;
        off += (int64_t)J.na[a] * J.na[a];
        chunk_max = std::max(chunk_max, J.na[a]);
      }
      rc = tau_launch(J, np, na, chunk_max, n, d->nligneso, d->cal_comp != 0, d->basereel != 0, d->fdm_comp != 0,
                      accumulate != 0, s);
      if (rc) return rc;
    }
  }
  return 0;
}

int fdmgpu_tau_ri(const fdmgpu_tau_ri_desc *d, int count, const double *const *d_xreg, const double *const *d_xirr,
                  double *const *d_tau_ri, int accumulate, void *hip_stream) {
  if (!d || count < 0 || d->natome < 0 || (d->natome > 0 && (!d->ianew_reg || !d->ianew_irr || !d->nlmsa)) ||
      d->n_ri < 1 || d->n_ri > 2 || d->n_ato_ri < 1 || d->n_ato_ri > d->n_ri || d->nlmso < 1 ||
      (count > 0 && d->natome > 0 && (!d_xreg || !d_xirr || !d_tau_ri))) {
    set_error("fdmgpu_tau_ri: invalid argument");
    return FDMGPU_EINVAL;
  }
  for (int a = 0; a < d->natome; a++) {
    const int na = d->nlmsa[a];
    if (na < 0 || d->ianew_reg[a] < 0 || d->ianew_reg[a] + na > d->nligne_reg || d->ianew_irr[a] < 0 ||
        d->ianew_irr[a] + d->n_ato_ri * na > d->nligne_irr) {
      set_error("fdmgpu_tau_ri: rows of atom " + std::to_string(a) + " outside a solution");
      return FDMGPU_EINVAL;
    }
  }
  for (int k = 0; k < count && d->natome > 0; k++)
    if (!d_xreg[k] || !d_xirr[k] || !d_tau_ri[k]) {
      set_error("fdmgpu_tau_ri: null device buffer for point " + std::to_string(k));
      return FDMGPU_EINVAL;
    }
  if (count == 0 || d->natome == 0) return 0;
  int rc = ensure_init();
  if (rc) return rc;
  hipStream_t s = hip_stream ? (hipStream_t)hip_stream : g_lib.stream;
  for (int p0 = 0; p0 < count; p0 += TAU_MAXP) {
    const int np = std::min(TAU_MAXP, count - p0);
    int64_t off = 0;
    for (int a0 = 0; a0 < d->natome; a0 += TAU_MAXA) {
      const int na = std::min(TAU_MAXA, d->natome - a0);
      TauRIJob J{};
      for (int k = 0; k < np; k++) {
        J.xr[k] = d_xreg[p0 + k];
        J.xi[k] = d_xirr[p0 + k];
        J.tau[k] = d_tau_ri[p0 + k];
      }
      for (int a = 0; a < na; a++) {
        J.row_reg[a] = d->ianew_reg[a0 + a];
        J.row_irr[a] = d->ianew_irr[a0 + a];
        J.na[a] = d->nlmsa[a0 + a];
        J.off[a] = off;
        off += (int64_t)d->n_ri * J.na[a] * J.na[a];
      }
      rc = tau_ri_launch(J, np, na, d->nlmso, d->n_ato_ri, d->n_ri, accumulate != 0, s);
      if (rc) return rc;
    }
  }
  return 0;
}

int fdmgpu_tau_symmetrize(int natome, const int *nlmsa, int ylm_comp, int count, double *const *d_tau,
                          void *hip_stream) {
  if (natome < 0 || count < 0 || (natome > 0 && !nlmsa) || (count > 0 && natome > 0 && !d_tau)) {
    set_error("fdmgpu_tau_symmetrize: invalid argument");
    return FDMGPU_EINVAL;
  }
  int na_max = 0;
  for (int a = 0; a < natome; a++) {
    const int L = (int)std::lround(std::sqrt((double)nlmsa[a])) - 1;
    if (nlmsa[a] < 1 || (L + 1) * (L + 1) != nlmsa[a]) {
      set_error("fdmgpu_tau_symmetrize: nlmsa must be (lmax+1)^2 (harmonics in (L, m) order)");
      return FDMGPU_EINVAL;
    }
    na_max = std::max(na_max, nlmsa[a]);
  }
  if ((size_t)na_max * na_max * 16 > TAU_MAX_LDS) {
    set_error("fdmgpu_tau_symmetrize: nlmsa^2 complex entries exceed the 160 KiB LDS");
    return FDMGPU_EINVAL;
  }
  for (int k = 0; k < count && natome > 0; k++)
    if (!d_tau[k]) {
      set_error("fdmgpu_tau_symmetrize: null device buffer for point " + std::to_string(k));
      return FDMGPU_EINVAL;
    }
  if (count == 0 || natome == 0) return 0;
  int rc = ensure_init();
  if (rc) return rc;
  hipStream_t s = hip_stream ? (hipStream_t)hip_stream : g_lib.stream;
  for (int p0 = 0; p0 < count; p0 += TAU_MAXP) {
    const int np = std::min(TAU_MAXP, count - p0);
    int64_t off = 0;
    for (int a0 = 0; a0 < natome; a0 += TAU_MAXA) {
      const int na = std::min(TAU_MAXA, natome - a0);
      TauSymJob J{};
      for (int k = 0; k < np; k++) J.tau[k] = d_tau[p0 + k];
      int chunk_max = 0;
      for (int a = 0; a < na; a++) {
        J.na[a] = nlmsa[a0 + a];
        J.off[a] = off;
        off += (int64_t)J.na[a] * J.na[a];
        chunk_max = std::max(chunk_max, J.na[a]);
      }
      rc = tau_sym_launch(J, np, na, chunk_max, ylm_comp != 0, s);
      if (rc) return rc;
    }
  }
  return 0;
}

}  // extern "C"
