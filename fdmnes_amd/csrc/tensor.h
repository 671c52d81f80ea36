// The absorbing atom's Cartesian cross-section tensors on the device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/fdmgpu.h"

namespace fdmgpu {

constexpr int TENS_MAXP = 32;     // energy points per call
constexpr int TENS_MAXINIT = 16;  // ninitr
constexpr int TENS_NOUT = 9 + 27 + 81;

// one product term of tens_ab: Cg * 0.5 (conj(rof1) rof2 T12 - conj(conj(rof2) rof1 T21))
struct TensTerm {
  int lm1, lm2;   // 0-based harmonics of the two final states
  int r1, r2;     // their radial-integral offsets (complex index into the point's rof)
  double cg_re, cg_im;
};

struct TensDev {
  int ncomp, nrep, ninitr, hydrogen, nrof;
  const int *term_ptr;        // [ncomp * ninitr + 1]
  const TensTerm *terms;
  const int *rep_ptr;         // [nrep + 1]
  const int *rep_comp;        // component of each (rep, spherical pair)
  const double *rep_w;        // clm_s * clm_i
  const int *fill;            // [TENS_NOUT][4]: rep (-1: zero), is, isi, conj
  double rot[9];
  int out_dd, out_dq, out_qq;  // which tensors are produced
};

struct TensPoints {
  const double *rof[TENS_MAXP];    // device radial integrals per point
  const double *taull[TENS_MAXP];  // device flat tau per point (complex interleaved)
  double *dd[TENS_MAXP], *dq[TENS_MAXP], *qq[TENS_MAXP];
};

struct TensorPlan {
  int ninitr = 0, max_points = 0, nrof = 0, nlm = 0;
  int64_t toff = 0;  // complex offset of the absorber's block in a point's tau
  int ld = 0;        // its leading dimension (nlmsa of the absorber)
  bool dd = false, dq = false, qq = false;
  hipStream_t stream = nullptr;
  TensDev d{};
  std::vector<void *> owned;
  ~TensorPlan();
  int init(const fdmgpu_tensor_tables *t, int max_points, hipStream_t s);
  int run(int count, const TensPoints &P);
};

}  // namespace fdmgpu
