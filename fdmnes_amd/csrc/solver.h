// Host-side orchestration of the block-skyline UL solver.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <string>
#include <vector>

#include "kernels.h"

namespace fdmgpu {

void set_error(const std::string &msg);
const char *get_error();

#define FDM_HIP(call)                                                              \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      set_error(std::string(#call) + ": " + hipGetErrorString(e_));                \
      return -2;                                                                   \
    }                                                                              \
  } while (0)

// Block layout of one profile matrix (host copy plus device mirror).
struct Plan {
  int n = 0, NB = 0;
  std::vector<int> J0, J1, I0, Jx;
  std::vector<int64_t> tbase;
  int64_t ntiles = 0;
  int maxlen = 0;  // longest tile range of any solve kind
  // device mirrors
  int *dJ0 = nullptr, *dJ1 = nullptr, *dI0 = nullptr, *dJx = nullptr;
  int64_t *dtbase = nullptr;

  Plan() = default;
  Plan(const Plan &) = delete;
  Plan &operator=(const Plan &) = delete;
  ~Plan() { release(); }

  // lb1/lb2 1-based inclusive; returns 0 or -1 with an error message.
  int build(int n, const int *lb1, const int *lb2);
  int upload();
  void release();
  DevPlan dev() const { return DevPlan{n, NB, dJ0, dJ1, dI0, dJx, dtbase}; }
  // Algorithmic work on the element profile: complex (or real) multiply-adds
  // of the elimination, and stored entries of the factors (per RHS, the
  // multiply-adds of one forward + one backward substitution).
  static double factor_macs(int n, const int *lb1, const int *lb2);
  static double solve_macs(int n, const int *lb1, const int *lb2);
  // multiply-adds the blocked kernels actually perform (tile granularity)
  double tile_update_macs() const;
  double tile_other_macs() const;  // panels (2/3 TB^3) and triangular block products (TB^3/2 per tile)
};

// A group of independent systems factorised and solved by grouped launches
// (one launch per kernel and block step serves every member).  Members are
// aligned so their last block steps coincide.
template <class R, bool C>
struct Group {
  static constexpr int NC = Elem<R, C>::NC;
  struct Member {
    Plan *plan = nullptr;
    R *tiles = nullptr, *dinv = nullptr, *X = nullptr, *part = nullptr;
    int *fail = nullptr;
    int *done = nullptr;  // NB block flags of the persistent substitution
    int nrhs = 0;
  };
  std::vector<Member> m;
  std::vector<char> active;  // members taking part in factor()/solve()
  SysDev<R> *d_sys = nullptr;
  // grid cap of the tile GEMM launches: a persistent grid of that many
  // workgroups, or (at least the launch's block count) one per output block
  int persist_wg = 1 << 24;
  int gk = 8;                // block steps per trailing update (0: one-step tile GEMM)
  int trail_tile = 64;       // output columns per workgroup of the group update: 64 or 128 (two tiles)
  bool persist_solve = true; // one dependency-driven launch per substitution
  int solve_wg = 512;        // workgroups of the persistent substitution
  int *ctl = nullptr;        // device: [0] task counter, [1] wait-timeout flag
  hipStream_t stream = nullptr;
  // Lookahead: at the end of a group the update of the block rows and
  // columns of the next group runs on `stream`, the rest of the trailing
  // matrix on `side`, beside the next group's per-step chain.
  // 0 off; 1 on, the rest on the trailing update's grid; > 1 on, the rest on
  // a persistent grid of that many workgroups (slots left for the chain)
  int lookahead = 0;
  // Overlap inside a group: the left-looking update of block row and column
  // K-1 from steps K+1..Khi runs on `side` beside step K's panel and
  // products; `stream` adds step K's own contribution afterwards
  int overlap = 0;
  hipStream_t side = nullptr;
  hipEvent_t ev_fact = nullptr, ev_rest = nullptr;  // stream -> side, side -> stream
  hipEvent_t ev_step = nullptr, ev_left = nullptr;  // the same for the overlap
  int maxparts = 1;
  // per-kernel timing (optional): cumulative ms of the G_UPDATE launches
  bool time_updates = false;
  double update_ms = 0;
  int64_t update_launches = 0;
  double update_bytes = 0;  // algorithmic bytes of the timed launches (C read + write, A and B read once)
  std::vector<hipEvent_t> ev;

  Group() = default;
  Group(const Group &) = delete;
  Group &operator=(const Group &) = delete;
  ~Group() { release(); }

  // Allocates a member (tiles, inverses, RHS, partial sums).  X may be given
  // (device, (NB*TB) x nrhs, interleaved) or allocated when nullptr.
  int add(Plan *plan, int nrhs, R *X = nullptr);
  int finalize(hipStream_t s);  // uploads descriptors; call after the last add()
  void release();
  size_t factor_bytes() const;
  int zero(int g);  // zero member g's tiles and pivot flag
  // vals: interleaved (re, im) when complex
  int scatter_csr(int g, const int64_t *d_row_ptr, const int *d_col, const double *d_vals);
  int factor();
  int solve(int kind);  // in place on every member's X
  int read_fail(std::vector<int> &rows);  // also reports a substitution wait timeout
  int check_solve();                      // substitution wait timeout only

 private:
  std::vector<char> owns_x;
};

}  // namespace fdmgpu
