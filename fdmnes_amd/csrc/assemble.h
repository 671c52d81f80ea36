// Per-energy assembly of a cluster's compressed rows on the device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/fdmgpu.h"

namespace fdmgpu {

constexpr int ASM_MAXP = 16;  // energy points per assembly call

// Energy of one point: Enervide (grid diagonal), Eimag (broadening), and the
// outer-sphere kinetic energy Ecinetic_out (mat.f90:771-792).
struct AsmEnergy {
  double ener, eimag, ecin;
};
struct AsmPoints {
  AsmEnergy e[ASM_MAXP];
  const double *V[ASM_MAXP];  // device, npoint
  double *vals[ASM_MAXP];     // device, nnz (interleaved when complex)
  double *rhs[ASM_MAXP];      // device, (nligne, nlmso) row-major, or null
  const double *phi[ASM_MAXP];  // device, phiato of a cluster with atoms (interleaved when complex)
  const double *phi2[ASM_MAXP];  // device, the irregular atomic solutions' phiato (irregular system)
  const double *G[ASM_MAXP];    // device, gradient of V (npoint x 3, point-major) for relativistic rows
};
// stencil and gradient coefficients of the relativistic terms, 1..12 in
// the reference's neighbour order (clapl[0] unused)
struct RelCoef {
  double clapl[13], cgrad[13];
};

struct Assembler {
  int npoint = 0, nligne = 0, nlmso = 0, lmaxso = 0, nsortf = 0, nb = 0, nrow_out = 0, nlink = 0;
  int64_t nnz = 0;
  double clapl0 = 0, pw = 0;
  int max_points = 0;
  hipStream_t stream = nullptr;
  // device copies of the energy-independent tables
  int64_t *row_ptr = nullptr, *qdiag = nullptr, *oq = nullptr;
  int *orow = nullptr, *optr = nullptr, *ob = nullptr;
  double *base = nullptr, *ow = nullptr, *ylm = nullptr, *brad = nullptr;
  int *lharm = nullptr;                          // l of each outer unknown
  double *wso = nullptr, *dself = nullptr;       // outer weights; diagonal mirror terms (null: none)
  double *fun = nullptr;   // Bessel/Neumann table: [point][nb][lmaxso+1] x {B, N} complex
  double *part = nullptr;  // outer-row partial sums: [point][run][nlmso^2] x {block, RHS} complex
  // muffin-tin atoms (natome = 0: none)
  int natome = 0, nar = 0, nal = 0, nrow_atom = 0, nlmax = 0;
  int nri = 1;  // 2: the irregular cluster solution (two blocks per atom, no outer unknowns)
  int64_t nphi = 0;
  int64_t *aq = nullptr, *alphi = nullptr, *bq = nullptr, *phoff = nullptr, *poff = nullptr, *yoff = nullptr;
  int *aa = nullptr, *aptr = nullptr, *ianew = nullptr, *nlmsa = nullptr, *nbordf = nullptr, *arow_atom = nullptr,
      *arow_lm = nullptr;
  double *alw = nullptr, *poidsa = nullptr, *ylmato = nullptr;
  // scalar-relativistic rows (init_rel): per grid point and direction the
  // position its interior neighbour's term lands in (-1: none), chi rvol(i)
  // and rvol(j); per outer-column row and row-atom pair the grid point, per
  // link its direction and chi rvol(i)
  bool rel = false;
  RelCoef relc{};
  int64_t *rq = nullptr;
  double *rfac = nullptr, *rdiv = nullptr, *ofac = nullptr;
  int *opoint = nullptr, *odir = nullptr, *apoint = nullptr, *adir = nullptr;

  // complex harmonics or characters (init_complex): the imaginary parts of
  // base, of the outside links' weights, of the outer harmonics and of the
  // points' own-image terms; the arithmetic is then complex
  bool cx = false;
  double *base_im = nullptr, *ow_im = nullptr, *ylm_im = nullptr, *dself_im = nullptr;

  ~Assembler();
  // at: null for a cluster without atoms
  int init(const fdmgpu_assembly_tables *t, const fdmgpu_atom_tables *at, int max_points, hipStream_t s);
  int init_atoms(const fdmgpu_atom_tables *at, int64_t natrow, const int64_t *row_ptr_h);
  int init_rel(const fdmgpu_rel_tables *t);
  int init_complex(const double *base_im_h, const double *ow_im_h, const double *ylm_im_h, const double *dself_h,
                   const double *dself_im_h);
  // update_only: vals/rhs already hold an earlier call's output for this
  // structure; only the energy-dependent entries are rewritten.  phi_cplx:
  // the points' phiato are complex (atoms only).  An assembler with
  // relativistic tables needs pts.G for every point.
  int run(int count, const AsmPoints &pts, bool e_comp, bool basereel, bool update_only, bool phi_cplx = false);
};

}  // namespace fdmgpu
