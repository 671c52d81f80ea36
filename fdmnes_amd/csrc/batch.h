// Device-resident solver for groups of energy points / clusters.
#pragma once
#include <map>
#include <memory>
#include <vector>

#include "refine.h"
#include "solver.h"

namespace fdmgpu {

// Workgroups of the first stage of the sigma reduction (partial sums per point).
constexpr int SIG_PARTS = 256;

// One system slot: structure (profile + compressed-row pattern) and the FP64
// buffers its energy points are staged into.
struct Slot {
  std::unique_ptr<Plan> plan;
  int n = 0;
  int64_t nnz = 0;
  int64_t *row_ptr = nullptr;  // device
  int *col = nullptr;          // device
  double *vals = nullptr;      // device, staged matrix values (interleaved when complex)
  double *B = nullptr;         // device, staged RHS, (NB*TB) x nrhs, FP64
  double *Xd = nullptr;        // device, FP64 solution (refinement)
  double *Rr = nullptr;        // device, FP64 residual (refinement)
  double *scr = nullptr;       // device, norm partials
  int *long_rows = nullptr;    // device, rows with more than LONG_ROW entries
  int nlong = 0;
  double factor_macs = 0, solve_macs = 0;
};

// Solver in precision R (double = FP64 factors, float = FP32 factors with
// FP64 refinement), complex when C.
template <class R, bool C>
struct Solver {
  static constexpr int NC = C ? 2 : 1;
  int nrhs = 0;
  hipStream_t stream = nullptr;
  bool own_stream = false;
  std::vector<Slot> slots;
  Group<R, C> grp;
  bool finalized = false;
  std::map<uint32_t, hipGraphExec_t> graphs;  // factorisation graphs by active-slot mask
  std::map<uint32_t, hipGraphExec_t> sweeps_g;  // refinement-sweep graphs by active-slot mask
  double *hnorm = nullptr;                       // pinned host copy of the norm partials
  double *sig_buf = nullptr;                     // device: sigma partials and results
  double *hsig = nullptr;                        // pinned host copy of the sigma results
  bool use_graphs = true;
  // refinement / fallback statistics of the last solve
  std::vector<int> last_sweeps;
  std::vector<double> last_rel;
  // FP64 groups of one slot each for the points whose FP32 refinement fails,
  // allocated on a slot's first fallback and kept (scheduled like grp)
  std::vector<std::unique_ptr<Group<double, C>>> fb;
  // Test hook: host copies, from the last solve, of each solved slot's factor
  // and inverse tiles after the factorisation and, per refinement sweep, of
  // the FP32 correction d and the FP64 x after it (by slot, by sweep).
  bool snapshots = false;
  std::vector<std::vector<char>> snap_factor;
  std::vector<std::vector<std::vector<char>>> snap_d, snap_x;

  ~Solver();
  int init(int nrhs, hipStream_t s);
  int add_system(int n, const int *lb1, const int *lb2, const int64_t *row_ptr, const int *col, int *slot);
  int finalize();
  int set_active(int count, const int *ids);
  // values/b/x: device pointers per listed slot
  int solve(int count, const int *ids, const double *const *vals, const double *const *b, double *const *x,
            int max_sweeps, double tol, int *fail_rows, int *sweeps, double *rel);
  int adjoint(int count, const int *ids, const double *const *src, const double *const *x, double *const *grad);
  int sigma(int count, const int *ids, const double *const *src, const double *const *x, double *out);
  // The same reduction on the solver's stream with the partial sums in part
  // (count * SIG_PARTS doubles) and sigma written to d_out (device, count
  // doubles): no allocation, no synchronisation.
  int sigma_into(int count, const int *ids, const double *const *src, const double *const *x, double *part,
                 double *d_out);
  void drop_graphs();  // after a schedule change

 private:
  int factor_phase(uint32_t mask);
  int sweep_phase(uint32_t mask);
  int fallback_fp64(int slot, int *fail_row, int max_sweeps, double tol, double *rel);
  int snap_factors(int count, const int *ids);
  int snap_sweep(const std::vector<int> &slot_ids);
};

}  // namespace fdmgpu
