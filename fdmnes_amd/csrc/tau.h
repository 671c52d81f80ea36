// Post-solve reduction of solutions to the atom amplitude products tau.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace fdmgpu {

constexpr int TAU_NT = 256;                  // threads per workgroup
constexpr int TAU_MAXP = 16;                 // energy points per launch
constexpr int TAU_MAXA = 32;                 // atoms per launch
constexpr size_t TAU_MAX_LDS = 160 * 1024;   // LDS of one workgroup on gfx950

// One launch: grid (points, atoms), kernel argument by value.
struct TauJob {
  const double *x[TAU_MAXP];  // device solutions, (nligne, nlmso) row-major (interleaved when complex)
  double *tau[TAU_MAXP];      // device outputs, interleaved complex
  double *norm[TAU_MAXP];     // device, nlmso^2 column-major, or null
  int *info;                  // device, one per point of this launch, or null
  int row[TAU_MAXA], na[TAU_MAXA];
  int64_t off[TAU_MAXA];      // complex offset of each atom's block in tau
};

// Tau_RI of the regular and irregular cluster solutions (complex, interleaved)
struct TauRIJob {
  const double *xr[TAU_MAXP], *xi[TAU_MAXP];  // regular / irregular solutions, (nligne, nlmso) row-major
  double *tau[TAU_MAXP];                      // per point: atoms' (n_ri, na, na) blocks, interleaved
  int row_reg[TAU_MAXA], row_irr[TAU_MAXA], na[TAU_MAXA];
  int64_t off[TAU_MAXA];                      // complex offset of each atom's block
};
// symmetrisation of the atoms' tau blocks in place
struct TauSymJob {
  double *tau[TAU_MAXP];
  int na[TAU_MAXA];
  int64_t off[TAU_MAXA];
};

int tau_ri_launch(const TauRIJob &J, int npts, int natoms, int n, int n_ato_ri, int n_ri, bool accumulate,
                  hipStream_t s);
int tau_sym_launch(const TauSymJob &J, int npts, int natoms, int na_max, bool ylm_comp, hipStream_t s);
size_t tau_lds_bytes(int n, int na_max);
int tau_launch(const TauJob &J, int npts, int natoms, int na_max, int n, int nso, bool cal, bool basereel,
               bool fdm_comp, bool accumulate, hipStream_t s);

}  // namespace fdmgpu
