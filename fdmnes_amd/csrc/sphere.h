// The atoms' radial solutions and phiato on the device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/fdmgpu.h"

namespace fdmgpu {

constexpr int SPH_MAXP = 16;  // energy points per call

// device tables (kernel argument by value)
struct SphTabs {
  int natome, nball, nrmax, rel;
  const double *r, *V, *clm, *clp, *cl0, *cgm, *cgp, *cg0, *f2;  // radial grids and stencils, concatenated
  const int64_t *roff;   // per atom: offset of its grid
  const int *nr, *Z, *lmax, *inr;
  const double *rmtg;
  const int64_t *toff;   // per atom: offset of its Tau entries (lmax + 1 per atom)
  const int *batom, *bint;  // per border point: atom, 1-based upper interpolation point
  const double *rrel;    // per border point: distance to its atom
  const int64_t *yoff;   // per border point: offset of its harmonics (and of its phiato row)
  const double *ylmato;
};
struct SphPoints {
  double ecin[SPH_MAXP], eimag[SPH_MAXP], ener[SPH_MAXP];
  double energ[SPH_MAXP];  // photoelectron energy Energ (the quadrupole factor of the radial integrals)
  double *phi[SPH_MAXP];  // device phiato per point, (nbord, nlmsa) per atom, concatenated
  double *tau[SPH_MAXP];  // device Tau per point, (lmax + 1) complex per atom, or null
  double *rof[SPH_MAXP];  // device radial integrals of the absorbing atom per point, or null
  double *phi2[SPH_MAXP];  // device phiato of the irregular solution per point, or null
};
// the absorbing atom's radial integrals of the golden rule (kernel argument)
struct RofTabs {
  int atom = -1, lmax = 0, nbs = 0, ip0 = 1, ip_max = 1, nlast = 0;
  double eseuil[2] = {0, 0};
  const double *psii = nullptr;  // [nbs][nr of the atom] core functions on the atom's grid
  double rmtsd = 0;
};

struct SphereSolver {
  int natome = 0, nball = 0, nrmax = 0, lmax_all = 0, max_points = 0;
  int64_t nphi = 0, ntau = 0, ustride = 0;
  hipStream_t stream = nullptr;
  struct Dev {
    int natome = 0, nball = 0, nrmax = 0, rel = 0;
    double *r = nullptr, *V = nullptr, *clm = nullptr, *clp = nullptr, *cl0 = nullptr, *cgm = nullptr,
           *cgp = nullptr, *cg0 = nullptr, *f2 = nullptr;
    int64_t *roff = nullptr;
    int *nr = nullptr, *Z = nullptr, *lmax = nullptr, *inr = nullptr;
    double *rmtg = nullptr;
    int64_t *toff = nullptr;
    int *batom = nullptr, *bint = nullptr;
    double *rrel = nullptr;
    int64_t *yoff = nullptr;
    double *ylmato = nullptr;
    operator SphTabs() const {
      return SphTabs{natome, nball, nrmax, rel, r, V, clm, clp, cl0, cgm, cgp, cg0, f2, roff, nr, Z, lmax, inr,
                     rmtg, toff, batom, bint, rrel, yoff, ylmato};
    }
  } d;
  double *u = nullptr;  // renormalised radial solutions: [point][atom][L] x (re[nrmax], im[nrmax])
  double *u2 = nullptr;  // the irregular solutions over r, laid out as u (allocated on first use)
  std::vector<int> h_nr, h_lmax, h_inr;  // host copies of the atoms' sizes and matching points
  std::vector<std::vector<double>> h_r;
  RofTabs rt;                      // set by set_integrals
  double *d_psii = nullptr;

  ~SphereSolver();
  int init(const fdmgpu_sphere_tables *t, int max_points, hipStream_t s);
  int set_integrals(int atom, int nbseuil, const double *psii, const double *Eseuil, double rmtsd, int ip0,
                    int ip_max);
  int64_t nrof() const { return rt.atom < 0 ? 0 : (int64_t)rt.nbs * (rt.ip_max - rt.ip0 + 1) * (rt.lmax + 1); }
  int run(int count, const SphPoints &P, bool cplx);
};

}  // namespace fdmgpu
