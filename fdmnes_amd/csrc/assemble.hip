// Per-energy assembly of a cluster's compressed rows on the device: the
// energy-dependent entries of calcMatRow (mat.f90:1142-1632), written into
// the value array the solver reads.  The tables
// of everything that does not change with the energy are uploaded once; a
// call then computes, for up to ASM_MAXP points at a time,
//   1. the outer-sphere functions fnorm j_l(k r), fnorm n_l(k r) at every
//      outer point (phiso, mat.f90:771-792, :975-994; recurrence of
//      sphere.f90:7226-7248),
//   2. every entry's energy-independent value (copy of the base table;
//      skipped when the caller's buffers already hold an earlier point's
//      values for this structure, which differ only in the entries below),
//   3. the grid diagonals V - E - clapl(0) (- i Eimag)   (mat.f90:1147-1150),
//   4. the outer columns of grid rows with neighbours outside the sphere and
//      their incoming-wave RHS                         (mat.f90:1312-1398),
//   5. the outer rows' outer-column block and RHS       (mat.f90:1401-1511),
// and for a cluster with muffin-tin atoms, from the caller's radial
// solutions phiato at the atoms' border points,
//   6. the grid rows' couplings to an atom's expansion  (mat.f90:1233-1309),
//   7. the atom rows' own blocks                         (mat.f90:1583-1632),
// and with scalar-relativistic rows (mat.f90:1155-1156, :1171-1185) the
// diagonal's -a2s4 (V - E)^2 and every stencil coefficient of a grid row,
// -clapl(v) + bder cgrad(v) dV/dx_k, in the interior entries (8.), the outer
// columns (4.) and the atom couplings (6.).
// Multiply-adds are not fused (contraction off) and the grid rows' sums run
// in the reference's term order; the outer rows' sums over the boundary
// points are taken in runs of 256 points, then over the runs, in order.  The entries agree with a
// host assembly of the same rows to the last bits of the Bessel/Neumann
// values and of those sums.
#include <algorithm>
#include <vector>
#include <cmath>
#include <string>

#include "assemble.h"
#include "solver.h"

#pragma clang fp contract(off)

namespace fdmgpu {
namespace {

struct cz {
  double re, im;
};
__device__ inline cz add(cz a, cz b) { return {a.re + b.re, a.im + b.im}; }
__device__ inline cz sub(cz a, cz b) { return {a.re - b.re, a.im - b.im}; }
__device__ inline cz mul(cz a, cz b) { return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
__device__ inline cz scale(cz a, double s) { return {a.re * s, a.im * s}; }
// a / b by Smith's algorithm (R. L. Smith, CACM 5(8):435, 1962)
__device__ inline cz cdiv(cz a, cz b) {
  if (fabs(b.re) < fabs(b.im)) {
    const double r = b.re / b.im, d = b.re * r + b.im;
    return {(a.re * r + a.im) / d, (a.im * r - a.re) / d};
  }
  const double r = b.im / b.re, d = b.im * r + b.re;
  return {(a.im * r + a.re) / d, (a.im - a.re * r) / d};
}
__device__ inline cz csin(cz z) { return {sin(z.re) * cosh(z.im), cos(z.re) * sinh(z.im)}; }
__device__ inline cz ccos(cz z) { return {cos(z.re) * cosh(z.im), -(sin(z.re) * sinh(z.im))}; }
// principal square root
__device__ inline cz csqrt(cz z) {
  if (z.re == 0 && z.im == 0) return {0, z.im};
  const double t = sqrt((fabs(z.re) + hypot(z.re, z.im)) * 0.5);
  if (z.re >= 0) return {t, z.im / (2 * t)};
  return {fabs(z.im) / (2 * t), copysign(t, z.im)};
}

constexpr double ASM_A2S4 = 0.25 * (0.0072973525664 * 0.0072973525664);  // alpha^2 / 4 (mat.f90:1123)

// bder of grid point i (mat.f90:1156)
__device__ inline double rel_bder(const AsmPoints &P, int g, int i) {
  const double vme = P.V[g][i] - P.e[g].ener;
  return -ASM_A2S4 / (1 - ASM_A2S4 * vme);
}
// stencil coefficient of direction d (0..11) of grid point i with the
// relativistic term, before the character and rvol factors (mat.f90:1164-1185)
__device__ inline double rel_coef(const RelCoef &rc, double bder, const double *Gi, int d) {
  return -rc.clapl[d + 1] + bder * rc.cgrad[d + 1] * Gi[(d >> 1) % 3];
}

__device__ inline int lof(int lm) {  // l of the harmonic index l*l + l + m
  int l = 0;
  while ((l + 1) * (l + 1) <= lm) l++;
  return l;
}

// 1. fun[((g*nb + b)*(L+1) + l)*2 + {0, 1}] = fnorm {j_l, n_l}(k brad[b])
__global__ void k_asm_fun(int nb, int L, const double *__restrict__ brad, AsmPoints P, int e_comp,
                          cz *__restrict__ fun) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x, g = blockIdx.y;
  if (b >= nb) return;
  const AsmEnergy E = P.e[g];
  cz *f = fun + ((int64_t)g * nb + b) * (L + 1) * 2;
  cz fn, z;
  if (e_comp) {
    cz k = csqrt({E.ecin, E.eimag});
    if (hypot(k.re, k.im) < 1e-10) k = {1e-10, 0};
    fn = csqrt({k.re / 3.1415926535897932384626433832, k.im / 3.1415926535897932384626433832});
    z = scale(k, brad[b]);
  } else {
    const double k = sqrt(E.ecin);
    fn = {sqrt(k / 3.1415926535897932384626433832), 0};
    z = {k * brad[b], 0};
  }
  cz bl2, nl2, bl1, nl1;
  if (e_comp) {
    nl1 = cdiv(mul({-fn.re, -fn.im}, ccos(z)), z);
    bl1 = cdiv(mul(fn, csin(z)), z);
  } else {  // real arithmetic, as the host's real branch
    nl1 = {-fn.re * cos(z.re) / z.re, 0};
    bl1 = {fn.re * sin(z.re) / z.re, 0};
  }
  f[0] = bl1;
  f[1] = nl1;
  if (L < 1) return;
  if (e_comp) {
    bl2 = add(cdiv(bl1, z), nl1);
    nl2 = sub(cdiv(nl1, z), bl1);
  } else {
    bl2 = {bl1.re / z.re + nl1.re, 0};
    nl2 = {nl1.re / z.re - bl1.re, 0};
  }
  f[2] = bl2;
  f[3] = nl2;
  for (int l = 2; l <= L; l++) {
    cz bn, nn;
    const double c = 2 * l - 1;
    if (e_comp) {
      nn = sub(cdiv(scale(nl2, c), z), nl1);
      bn = sub(cdiv(scale(bl2, c), z), bl1);
    } else {
      nn = {c * nl2.re / z.re - nl1.re, 0};
      bn = {c * bl2.re / z.re - bl1.re, 0};
    }
    f[2 * l] = bn;
    f[2 * l + 1] = nn;
    bl1 = bl2;
    nl1 = nl2;
    bl2 = bn;
    nl2 = nn;
  }
}

// outer-column factor of outer point b (calcMatRow, mat.f90:1330-1346): the
// Neumann function in the real basis, j + i n in the complex one
__device__ inline cz cfac(const cz *f, int l, bool basereel) {
  const cz B = f[2 * l], N = f[2 * l + 1];
  if (basereel) return N;
  return {B.re - N.im, B.im + N.re};
}

// 2. values from the base table; RHS cleared
template <bool CAL>
__global__ void k_asm_base(int64_t nnz, int64_t nrhs_el, const double *__restrict__ base,
                           const double *__restrict__ base_im, AsmPoints P) {
  const int g = blockIdx.y;
  double *v = P.vals[g], *r = P.rhs[g];
  const int64_t n = nnz > nrhs_el ? nnz : nrhs_el;
  for (int64_t q = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
    if (q < nnz) {
      if constexpr (CAL) {
        v[2 * q] = base[q];
        v[2 * q + 1] = base_im ? base_im[q] : 0.0;
      } else {
        v[q] = base[q];
      }
    }
    if (r && q < nrhs_el) {
      if constexpr (CAL) {
        r[2 * q] = 0;
        r[2 * q + 1] = 0;
      } else {
        r[q] = 0;
      }
    }
  }
}

// 3. grid diagonals
template <bool CAL>
__global__ void k_asm_diag(int npoint, const int64_t *__restrict__ qdiag, double clapl0,
                           const double *__restrict__ dself, const double *__restrict__ dself_im, AsmPoints P,
                           int e_comp, int rel) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, g = blockIdx.y;
  if (i >= npoint) return;
  const AsmEnergy E = P.e[g];
  const int64_t q = qdiag[i];
  if (q < 0) return;  // a point inside an atom has no row
  double d = P.V[g][i] - E.ener - clapl0;
  if (rel) {  // mat.f90:1155; the point's own images follow in k_asm_rel
    const double vme = P.V[g][i] - E.ener;
    d = d - ASM_A2S4 * (vme * vme);
  } else if (dself) {  // mirror images of the point itself, added in axis order (mat.f90:1225-1231)
    d = d + dself[3 * i];
    d = d + dself[3 * i + 1];
    d = d + dself[3 * i + 2];
  }
  if constexpr (CAL) {
    double di = e_comp ? -E.eimag : 0.0;
    if (dself_im) {  // complex characters: the images' terms are complex
      di = di + dself_im[3 * i];
      di = di + dself_im[3 * i + 1];
      di = di + dself_im[3 * i + 2];
    }
    P.vals[g][2 * q] = d;
    P.vals[g][2 * q + 1] = di;
  } else {
    P.vals[g][q] = d;
  }
}

// 4. outer columns of the grid rows touching the outside, and their RHS:
// one thread per (row, harmonic), outside neighbours in stencil order
// (relativistic rows: the weight of each link recomputed from its
// direction, chi rvol(i) and the row's point)
template <bool CAL, bool REL>
__global__ void k_asm_outcols(int nrow_out, int nl, int nb, int L, const int *__restrict__ orow,
                              const int64_t *__restrict__ oq, const int *__restrict__ optr,
                              const int *__restrict__ ob, const double *__restrict__ ow,
                              const double *__restrict__ ylm, const int *__restrict__ lharm,
                              const cz *__restrict__ fun, AsmPoints P, int basereel,
                              const int *__restrict__ opoint, const int *__restrict__ odir,
                              const double *__restrict__ ofac, RelCoef rc, const double *__restrict__ ow_im,
                              const double *__restrict__ ylm_im, int irr) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x, g = blockIdx.y;
  if (t >= nrow_out * nl) return;
  const int r = t / nl, lm = t % nl, l = lharm[lm];
  const cz *fg = fun + (int64_t)g * nb * (L + 1) * 2;
  double bder = 0;
  const double *Gi = nullptr;
  if constexpr (REL) {
    const int i = opoint[r];
    bder = rel_bder(P, g, i);
    Gi = P.G[g] + 3 * (int64_t)i;
  }
  cz row = {0, 0}, rrow = {0, 0};
  for (int k = optr[r]; k < optr[r + 1]; k++) {
    const int b = ob[k];
    double w;
    if constexpr (REL)
      w = rel_coef(rc, bder, Gi, odir[k]) * ofac[k];
    else
      w = ow[k];
    const double y = ylm[(int64_t)b * nl + lm];
    const cz *f = fg + (int64_t)b * (L + 1) * 2;
    if (CAL && irr) {  // the irregular solution: no outer unknowns, incoming wave -i j + n (mat.f90:1318-1398)
      const cz c = {f[2 * l].im + f[2 * l + 1].re, -f[2 * l].re + f[2 * l + 1].im};
      rrow = sub(rrow, scale(scale(c, y), w));
      continue;
    }
    if (CAL && ylm_im) {  // complex harmonics (Yc) and weights chi -clapl(v) rvol(i) (mat.f90:1328-1377)
      const cz yc = {y, ylm_im[(int64_t)b * nl + lm]}, wc = {w, ow_im ? ow_im[k] : 0.0};
      row = add(row, mul(mul(cfac(f, l, basereel), yc), wc));
      rrow = sub(rrow, mul(mul(f[2 * l], yc), wc));
      continue;
    }
    row = add(row, scale(scale(cfac(f, l, basereel), y), w));
    rrow = sub(rrow, scale(scale(f[2 * l], y), w));
  }
  const int64_t q = oq[r] + lm, qr = (int64_t)orow[r] * nl + lm;
  double *v = P.vals[g], *rh = P.rhs[g];
  if constexpr (CAL) {
    if (!irr) {
      v[2 * q] = row.re;
      v[2 * q + 1] = row.im;
    }
    if (rh) {
      rh[2 * qr] = rrow.re;
      rh[2 * qr + 1] = rrow.im;
    }
  } else {
    v[q] = row.re;
    if (rh) rh[qr] = rrow.re;
  }
}

// 5. outer rows: block (lm, lp) = sum_b w_b Y_lm(b) Y_lp(b) cfac(b, l(lp)) and
// RHS (lm, lf) = -sum_b w_b Y_lm(b) Y_lf(b) B(b, l(lf)), over the nsortf inner
// outer points.  First pass: one workgroup per (point, run of ASM_CB points
// b, 256 (lm, lp) pairs), each thread summing its pair over the run in
// order; second pass: one thread per (point, pair) adds the runs' partial
// sums in run order.  The result does not depend on timing.
constexpr int ASM_CB = 256;
__global__ __launch_bounds__(256) void k_asm_outrows_part(int nl, int nb, int L, int nsortf,
                                                           const double *__restrict__ wso,
                                                           const double *__restrict__ ylm,
                                                           const int *__restrict__ lharm,
                                                           const cz *__restrict__ fun, int basereel,
                                                           cz *__restrict__ part,
                                                           const double *__restrict__ ylm_im) {
  const int run = blockIdx.x, nrun = gridDim.x, g = blockIdx.z;
  const int pr = blockIdx.y * blockDim.x + threadIdx.x;
  if (pr >= nl * nl) return;
  const int lm = pr / nl, lp = pr % nl, l = lharm[lp];
  const cz *fg = fun + (int64_t)g * nb * (L + 1) * 2;
  cz s = {0, 0}, sr = {0, 0};
  const int b1 = min(nsortf, (run + 1) * ASM_CB);
  if (ylm_im) {  // complex harmonics: w conj(Yc_lm) Yc_lp (mat.f90:1425-1497)
    for (int b = run * ASM_CB; b < b1; b++) {
      const cz y1 = {ylm[(int64_t)b * nl + lm], -ylm_im[(int64_t)b * nl + lm]};
      const cz y2 = {ylm[(int64_t)b * nl + lp], ylm_im[(int64_t)b * nl + lp]};
      const cz cb = scale(mul(y1, y2), wso[b]), cr = mul(scale(y1, wso[b]), y2);
      const cz *f = fg + (int64_t)b * (L + 1) * 2;
      s = add(s, mul(cb, cfac(f, l, basereel)));
      sr = add(sr, mul(cr, f[2 * l]));
    }
  } else {
#pragma unroll 8
  for (int b = run * ASM_CB; b < b1; b++) {
    const double c = wso[b] * ylm[(int64_t)b * nl + lm] * ylm[(int64_t)b * nl + lp];
    const cz *f = fg + (int64_t)b * (L + 1) * 2;
    s = add(s, scale(cfac(f, l, basereel), c));
    sr = add(sr, scale(f[2 * l], c));
  }
  }
  cz *o = part + (((int64_t)g * nrun + run) * nl * nl + pr) * 2;
  o[0] = s;
  o[1] = sr;
}
template <bool CAL>
__global__ void k_asm_outrows_sum(int nso, int nl, int nsortf, int nrun, const int64_t *__restrict__ row_ptr,
                                  const cz *__restrict__ part, AsmPoints P) {
  const int pr = blockIdx.x * blockDim.x + threadIdx.x, g = blockIdx.y;
  if (pr >= nl * nl) return;
  const int lm = pr / nl, lp = pr % nl;
  cz t = {0, 0}, tr = {0, 0};
  for (int run = 0; run < nrun; run++) {
    const cz *o = part + (((int64_t)g * nrun + run) * nl * nl + pr) * 2;
    t = add(t, o[0]);
    tr = add(tr, o[1]);
  }
  const int ii = nso + lm;
  const int64_t q = row_ptr[ii] + nsortf + lp, qr = (int64_t)ii * nl + lp;
  double *v = P.vals[g], *rh = P.rhs[g];
  if constexpr (CAL) {
    v[2 * q] = t.re;
    v[2 * q + 1] = t.im;
    if (rh) {
      rh[2 * qr] = -tr.re;
      rh[2 * qr + 1] = -tr.im;
    }
  } else {
    v[q] = t.re;
    if (rh) rh[qr] = -tr.re;
  }
}

// 6. a grid row's couplings to the expansion of an atom its stencil reaches:
// the sum over its neighbours inside the atom, in stencil order, of
// -clapl(v) phiato(neighbour, Lm); one thread per (row-atom pair, Lm)
// (relativistic rows: each link's weight recomputed from its direction and
// the row's point; atoms are built in P1, where chi rvol(i) = 1)
template <bool CAL, bool PHIC, bool REL>
__global__ void k_asm_atomcols(int nar, int nlmax, int nri, const int64_t *__restrict__ aq, const int *__restrict__ aa,
                               const int *__restrict__ aptr, const int64_t *__restrict__ alphi,
                               const double *__restrict__ alw, const int *__restrict__ nlmsa, AsmPoints P,
                               const int *__restrict__ apoint, const int *__restrict__ adir, RelCoef rc) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x, g = blockIdx.y;
  if (t >= nar * nlmax) return;
  const int r = t / nlmax, lmb = t % nlmax, na = nlmsa[aa[r]];
  if (lmb >= nri * na) return;
  // the irregular solution's second column block takes the irregular phiato (mat.f90:1266-1307)
  const int ri = lmb >= na, lm = lmb - ri * na;
  const double *phi = ri ? P.phi2[g] : P.phi[g];
  double bder = 0;
  const double *Gi = nullptr;
  if constexpr (REL) {
    const int i = apoint[r];
    bder = rel_bder(P, g, i);
    Gi = P.G[g] + 3 * (int64_t)i;
  }
  double sr = 0, si = 0;
  for (int k = aptr[r]; k < aptr[r + 1]; k++) {
    double w;
    if constexpr (REL)
      w = rel_coef(rc, bder, Gi, adir[k]);
    else
      w = alw[k];
    const int64_t o = alphi[k] + lm;
    if constexpr (PHIC) {
      sr = sr + w * phi[2 * o];
      si = si + w * phi[2 * o + 1];
    } else {
      sr = sr + w * phi[o];
    }
  }
  const int64_t q = aq[r] + lmb;
  double *v = P.vals[g];
  if constexpr (CAL) {
    v[2 * q] = sr;
    v[2 * q + 1] = si;
  } else {
    v[q] = sr;
  }
}

// 7. atom row (a, Lm), column Lm' of its own block: the sum over the atom's
// outside border points, in border order, of poidsa Ylmato(Lm) phiato(Lm');
// one thread per (atom row, Lm')
template <bool CAL, bool PHIC>
__global__ void k_asm_atomrows(int nrow_atom, int nlmax, int nri, const int *__restrict__ arow_atom,
                               const int *__restrict__ arow_lm, const int64_t *__restrict__ bq,
                               const int *__restrict__ nlmsa, const int *__restrict__ nbordf,
                               const int64_t *__restrict__ phoff, const int64_t *__restrict__ poff,
                               const int64_t *__restrict__ yoff, const double *__restrict__ poidsa,
                               const double *__restrict__ ylmato, AsmPoints P) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x, g = blockIdx.y;
  if (t >= nrow_atom * nlmax) return;
  const int k = t / nlmax, lpb = t % nlmax, a = arow_atom[k], lm = arow_lm[k], na = nlmsa[a];
  if (lpb >= nri * na) return;
  // irregular solution: column block rj from its phiato over its half of the border (mat.f90:1580-1630)
  const int rj = lpb >= na, lp = lpb - rj * na, nh = nbordf[a] / 2;
  const int b0 = (nri == 2 && rj) ? nh : 0, b1 = (nri == 2 && !rj) ? nh : nbordf[a];
  const double *phi = rj ? P.phi2[g] : P.phi[g], *pw = poidsa + poff[a], *Y = ylmato + yoff[a];
  double sr = 0, si = 0;
  for (int b = b0; b < b1; b++) {
    const double y = Y[(int64_t)b * na + lm];
    const int64_t o = phoff[a] + (int64_t)b * na + lp;
    if constexpr (!CAL) {
      sr = sr + pw[b] * y * phi[o];
    } else if constexpr (PHIC) {
      sr = sr + pw[b] * (phi[2 * o] * y);
      si = si + pw[b] * (phi[2 * o + 1] * y);
    } else {
      sr = sr + pw[b] * (phi[o] * y);
    }
  }
  const int64_t q = bq[k] + lpb;
  double *v = P.vals[g];
  if constexpr (CAL) {
    v[2 * q] = sr;
    v[2 * q + 1] = si;
  } else {
    v[q] = sr;
  }
}

// 8. relativistic grid rows: the entries towards interior neighbours (and
// the point's own images, on the diagonal k_asm_diag wrote) recomputed with
// the relativistic coefficient, neighbours in the reference's order
// (mat.f90:1164-1232); one thread per grid point, which alone writes its row
template <bool CAL>
__global__ void k_asm_rel(int npoint, const int64_t *__restrict__ qdiag, const int64_t *__restrict__ rq,
                          const double *__restrict__ rfac, const double *__restrict__ rdiv, RelCoef rc,
                          AsmPoints P) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x, g = blockIdx.y;
  if (i >= npoint) return;
  const int64_t qd = qdiag[i];
  if (qd < 0) return;  // inside an atom: no row
  double *v = P.vals[g];
  const int64_t *q = rq + 12 * (int64_t)i;
  for (int d = 0; d < 12; d++) {
    if (q[d] < 0 || q[d] == qd) continue;
    if constexpr (CAL) {
      v[2 * q[d]] = 0;
      v[2 * q[d] + 1] = 0;
    } else {
      v[q[d]] = 0;
    }
  }
  const double bder = rel_bder(P, g, i);
  const double *Gi = P.G[g] + 3 * (int64_t)i;
  for (int d = 0; d < 12; d++) {
    if (q[d] < 0) continue;
    const double t = (rel_coef(rc, bder, Gi, d) * rfac[12 * (int64_t)i + d]) / rdiv[12 * (int64_t)i + d];
    if constexpr (CAL)
      v[2 * q[d]] = v[2 * q[d]] + t;
    else
      v[q[d]] = v[q[d]] + t;
  }
}

template <class T>
int upload(T **d, const T *h, int64_t n) {
  FDM_HIP(hipMalloc(d, sizeof(T) * (n > 0 ? n : 1)));
  if (n > 0) FDM_HIP(hipMemcpy(*d, h, sizeof(T) * n, hipMemcpyHostToDevice));
  return 0;
}

}  // namespace

Assembler::~Assembler() {
  for (void *p : {(void *)row_ptr, (void *)qdiag, (void *)oq, (void *)orow, (void *)optr, (void *)ob, (void *)base,
                  (void *)ow, (void *)ylm, (void *)brad, (void *)lharm, (void *)wso, (void *)dself, (void *)fun,
                  (void *)part, (void *)aq, (void *)alphi, (void *)bq, (void *)phoff, (void *)poff, (void *)yoff,
                  (void *)aa, (void *)aptr, (void *)ianew, (void *)nlmsa, (void *)nbordf, (void *)arow_atom,
                  (void *)arow_lm, (void *)alw, (void *)poidsa, (void *)ylmato, (void *)rq, (void *)rfac,
                  (void *)rdiv, (void *)ofac, (void *)opoint, (void *)odir, (void *)apoint, (void *)adir,
                  (void *)base_im, (void *)ow_im, (void *)ylm_im, (void *)dself_im})
    if (p) (void)hipFree(p);
}

int Assembler::init(const fdmgpu_assembly_tables *t, const fdmgpu_atom_tables *at, int maxp, hipStream_t s) {
  const int nall = (t->lmaxso + 1) * (t->lmaxso + 1);  // a representation of mmm has a subset (lharm)
  int64_t natrow = 0;  // rows of the atoms' expansions
  const int nrr = (at && at->nri == 2) ? 2 : 1;
  const int nout_rows = nrr == 2 ? 0 : t->nlmso;  // outer rows (none in the irregular solution)
  if (at) {
    if (at->natome < 1 || at->nar < 0 || at->nal < at->nar || at->nphi < 1 || !at->ianew || !at->nlmsa ||
        !at->nbordf || !at->phoff || !at->bq || !at->poidsa || !at->ylmato ||
        (at->nar > 0 && (!at->aq || !at->aa || !at->aptr || !at->alphi || !at->alw))) {
      set_error("fdmgpu_assembler_create_atoms: inconsistent atom table sizes or null table");
      return FDMGPU_EINVAL;
    }
    for (int a = 0; a < at->natome; a++) {
      if (at->nlmsa[a] < 1 || at->nlmsa[a] > 169 || at->nbordf[a] < 1 || at->ianew[a] < 0 ||
          at->ianew[a] + nrr * at->nlmsa[a] > t->nligne - nout_rows || at->phoff[a] < 0 ||
          at->phoff[a] + (int64_t)at->nbordf[a] * at->nlmsa[a] > at->nphi) {
        set_error("fdmgpu_assembler_create_atoms: atom " + std::to_string(a) + " outside the rows or phiato");
        return FDMGPU_EINVAL;
      }
      natrow += nrr * at->nlmsa[a];
    }
  }
  const int64_t npr = t->nligne - nout_rows - natrow;  // interstitial grid rows
  if (t->npoint <= 0 || (at ? (npr < 1 || npr > t->npoint) : t->nligne != t->npoint + t->nlmso) || t->lmaxso < 0 ||
      t->lmaxso > 40 ||
      (t->lharm ? (t->nlmso < 1 || t->nlmso > nall) : t->nlmso != nall) || t->nsortf <= 0 || t->nb < t->nsortf ||
      t->nrow_out < 0 ||
      t->nlink < t->nrow_out || t->nnz <= 0 || maxp < 1 || maxp > ASM_MAXP) {
    set_error("fdmgpu_assembler_create: inconsistent table sizes");
    return FDMGPU_EINVAL;
  }
  if (!t->row_ptr || !t->base || !t->qdiag || !t->ylm || !t->brad ||
      (t->nrow_out > 0 && (!t->orow || !t->oq || !t->optr || !t->ob || !t->ow))) {
    set_error("fdmgpu_assembler_create: null table");
    return FDMGPU_EINVAL;
  }
  // every position written per energy must lie in the value array
  if (t->row_ptr[0] != 0 || t->row_ptr[t->nligne] != t->nnz) {
    set_error("fdmgpu_assembler_create: row_ptr does not span nnz");
    return FDMGPU_EINVAL;
  }
  const int nri_ = (at && at->nri == 2) ? 2 : 1;       // the irregular cluster solution: no outer rows
  const int nso = nri_ == 2 ? t->nligne : t->nligne - t->nlmso;  // first outer row
  for (int i = 0; i < t->npoint; i++)
    if (at ? (t->qdiag[i] < -1 || t->qdiag[i] >= t->nnz)
           : (t->qdiag[i] < t->row_ptr[i] || t->qdiag[i] >= t->row_ptr[i + 1])) {
      set_error("fdmgpu_assembler_create: diagonal position outside row " + std::to_string(i + 1));
      return FDMGPU_EINVAL;
    }
  for (int r = 0; r < t->nrow_out; r++) {
    const int i = t->orow[r];
    const bool cols_ok = nri_ == 2 ? t->oq[r] == -1
                                   : (t->oq[r] >= t->row_ptr[i] && t->oq[r] + t->nlmso <= t->row_ptr[i + 1]);
    if (i < 0 || i >= nso || !cols_ok ||
        t->optr[r + 1] < t->optr[r] || t->optr[r + 1] > t->nlink) {
      set_error("fdmgpu_assembler_create: outer-column table inconsistent at entry " + std::to_string(r));
      return FDMGPU_EINVAL;
    }
  }
  if (t->lharm)
    for (int k = 0; k < t->nlmso; k++)
      if (t->lharm[k] < 0 || t->lharm[k] > t->lmaxso) {
        set_error("fdmgpu_assembler_create: lharm out of range");
        return FDMGPU_EINVAL;
      }
  if (t->nrow_out > 0 && (t->optr[0] != 0 || t->optr[t->nrow_out] != t->nlink)) {
    set_error("fdmgpu_assembler_create: optr does not span nlink");
    return FDMGPU_EINVAL;
  }
  for (int k = 0; k < t->nlink; k++)
    if (t->ob[k] < 0 || t->ob[k] >= t->nb) {
      set_error("fdmgpu_assembler_create: outer point index out of range");
      return FDMGPU_EINVAL;
    }
  for (int lm = 0; lm < t->nlmso && nri_ == 1; lm++)
    if (t->row_ptr[nso + lm + 1] - t->row_ptr[nso + lm] != t->nsortf + t->nlmso) {
      set_error("fdmgpu_assembler_create: outer row " + std::to_string(lm) + " is not nsortf + nlmso long");
      return FDMGPU_EINVAL;
    }
  npoint = t->npoint;
  nligne = t->nligne;
  nlmso = t->nlmso;
  lmaxso = t->lmaxso;
  nsortf = t->nsortf;
  nb = t->nb;
  nrow_out = t->nrow_out;
  nlink = t->nlink;
  nnz = t->nnz;
  clapl0 = t->clapl0;
  pw = t->pw;
  max_points = maxp;
  stream = s;
  if (upload(&row_ptr, t->row_ptr, nligne + 1) || upload(&base, t->base, nnz) || upload(&qdiag, t->qdiag, npoint) ||
      upload(&orow, t->orow, nrow_out) || upload(&oq, t->oq, nrow_out) || upload(&optr, t->optr, nrow_out + 1) ||
      upload(&ob, t->ob, nlink) || upload(&ow, t->ow, nlink) || upload(&ylm, t->ylm, (int64_t)nb * nlmso) ||
      upload(&brad, t->brad, nb))
    return FDMGPU_EHIP;
  // P1 defaults: l of l*l + l + m, the uniform weight pw, no mirror terms
  std::vector<int> hl(nlmso);
  std::vector<double> hw(nsortf, pw);
  for (int k = 0; k < nlmso; k++) {
    int l = 0;
    while ((l + 1) * (l + 1) <= k) l++;
    hl[k] = t->lharm ? t->lharm[k] : l;
  }
  if (t->wso) std::copy(t->wso, t->wso + nsortf, hw.begin());
  if (upload(&lharm, hl.data(), nlmso) || upload(&wso, hw.data(), nsortf) ||
      (t->dself && upload(&dself, t->dself, 3 * (int64_t)npoint)))
    return FDMGPU_EHIP;
  FDM_HIP(hipMalloc(&fun, sizeof(cz) * (size_t)maxp * nb * (lmaxso + 1) * 2));
  const size_t nrun = (nsortf + ASM_CB - 1) / ASM_CB;
  FDM_HIP(hipMalloc(&part, sizeof(cz) * 2 * (size_t)maxp * nrun * nlmso * nlmso));
  return at ? init_atoms(at, natrow, t->row_ptr) : 0;
}

int Assembler::init_atoms(const fdmgpu_atom_tables *at, int64_t natrow, const int64_t *row_ptr_h) {
  // every position and phiato offset the kernels use lies in its array
  if (at->nar > 0 && (at->aptr[0] != 0 || at->aptr[at->nar] != at->nal)) {
    set_error("fdmgpu_assembler_create_atoms: aptr does not span nal");
    return FDMGPU_EINVAL;
  }
  const int nr2 = at->nri == 2 ? 2 : 1;
  for (int r = 0; r < at->nar; r++) {
    const int a = at->aa[r];
    if (a < 0 || a >= at->natome || at->aq[r] < 0 || at->aq[r] + nr2 * at->nlmsa[a] > nnz ||
        at->aptr[r + 1] < at->aptr[r]) {
      set_error("fdmgpu_assembler_create_atoms: row-atom pair " + std::to_string(r) + " inconsistent");
      return FDMGPU_EINVAL;
    }
    for (int k = at->aptr[r]; k < at->aptr[r + 1]; k++)
      if (at->alphi[k] < 0 || at->alphi[k] + at->nlmsa[a] > at->nphi) {
        set_error("fdmgpu_assembler_create_atoms: phiato offset out of range");
        return FDMGPU_EINVAL;
      }
  }
  std::vector<int> ra, rl;
  std::vector<int64_t> po(at->natome), yo(at->natome);
  int64_t pb = 0, yb = 0;
  for (int a = 0; a < at->natome; a++) {
    po[a] = pb;
    yo[a] = yb;
    pb += at->nbordf[a];
    yb += (int64_t)at->nbordf[a] * at->nlmsa[a];
    nlmax = std::max(nlmax, nr2 * at->nlmsa[a]);
    for (int lm = 0; lm < nr2 * at->nlmsa[a]; lm++) {  // both row blocks of the irregular solution
      const int row = at->ianew[a] + lm;
      if ((int64_t)ra.size() >= natrow) {
        set_error("fdmgpu_assembler_create_atoms: more atom rows than the sizes report");
        return FDMGPU_EINVAL;
      }
      const int64_t q = at->bq[ra.size()];
      if (q < row_ptr_h[row] || q + nr2 * at->nlmsa[a] > row_ptr_h[row + 1]) {
        set_error("fdmgpu_assembler_create_atoms: block of atom row " + std::to_string(row + 1) + " outside the row");
        return FDMGPU_EINVAL;
      }
      ra.push_back(a);
      rl.push_back(lm % at->nlmsa[a]);  // the harmonic of the row (its Ylmato column)
    }
  }
  if ((int64_t)ra.size() != natrow) {
    set_error("fdmgpu_assembler_create_atoms: atom rows differ from the sizes");
    return FDMGPU_EINVAL;
  }
  nri = nr2;
  natome = at->natome;
  nar = at->nar;
  nal = at->nal;
  nphi = at->nphi;
  nrow_atom = (int)natrow;
  if (upload(&aq, at->aq, nar) || upload(&aa, at->aa, nar) || upload(&aptr, at->aptr, nar + 1) ||
      upload(&alphi, at->alphi, nal) || upload(&alw, at->alw, nal) || upload(&ianew, at->ianew, natome) ||
      upload(&nlmsa, at->nlmsa, natome) || upload(&nbordf, at->nbordf, natome) ||
      upload(&phoff, at->phoff, natome) || upload(&bq, at->bq, natrow) || upload(&poidsa, at->poidsa, pb) ||
      upload(&ylmato, at->ylmato, yb) || upload(&poff, po.data(), natome) || upload(&yoff, yo.data(), natome) ||
      upload(&arow_atom, ra.data(), natrow) || upload(&arow_lm, rl.data(), natrow))
    return FDMGPU_EHIP;
  return 0;
}

int Assembler::init_rel(const fdmgpu_rel_tables *t) {
  if (rel) {
    set_error("fdmgpu_assembler_set_relativistic: the assembler already holds relativistic tables");
    return FDMGPU_EINVAL;
  }
  if (t->npoint != npoint || t->nrow_out != nrow_out || t->nlink != nlink || t->nar != nar || t->nal != nal ||
      !t->rq || !t->rfac || !t->rdiv || (nrow_out > 0 && (!t->opoint || !t->odir || !t->ofac)) ||
      (nar > 0 && (!t->apoint || !t->adir))) {
    set_error("fdmgpu_assembler_set_relativistic: table sizes differ from the assembler's, or a null table");
    return FDMGPU_EINVAL;
  }
  // every position and index the kernels use lies in its array; a target
  // lies in its point's row, positions are written only by their point
  std::vector<int64_t> rp(nligne + 1), qd(npoint);
  FDM_HIP(hipMemcpy(rp.data(), row_ptr, sizeof(int64_t) * (nligne + 1), hipMemcpyDeviceToHost));
  FDM_HIP(hipMemcpy(qd.data(), qdiag, sizeof(int64_t) * npoint, hipMemcpyDeviceToHost));
  for (int i = 0; i < npoint; i++)
    for (int d = 0; d < 12; d++) {
      const int64_t q = t->rq[12 * (int64_t)i + d];
      if (q < 0) continue;
      const int64_t qq = qd[i];
      int row = -1;  // the point's row: the one holding its diagonal
      if (qq >= 0) row = (int)(std::upper_bound(rp.begin(), rp.end(), qq) - rp.begin()) - 1;
      if (qq < 0 || q < rp[row] || q >= rp[row + 1] || !(t->rdiv[12 * (int64_t)i + d] > 0)) {
        set_error("fdmgpu_assembler_set_relativistic: entry of point " + std::to_string(i) + " outside its row");
        return FDMGPU_EINVAL;
      }
    }
  for (int r = 0; r < nrow_out; r++)
    if (t->opoint[r] < 0 || t->opoint[r] >= npoint) {
      set_error("fdmgpu_assembler_set_relativistic: opoint out of range");
      return FDMGPU_EINVAL;
    }
  for (int k = 0; k < nlink; k++)
    if (t->odir[k] < 0 || t->odir[k] > 11) {
      set_error("fdmgpu_assembler_set_relativistic: odir out of range");
      return FDMGPU_EINVAL;
    }
  for (int r = 0; r < nar; r++)
    if (t->apoint[r] < 0 || t->apoint[r] >= npoint) {
      set_error("fdmgpu_assembler_set_relativistic: apoint out of range");
      return FDMGPU_EINVAL;
    }
  for (int k = 0; k < nal; k++)
    if (t->adir[k] < 0 || t->adir[k] > 11) {
      set_error("fdmgpu_assembler_set_relativistic: adir out of range");
      return FDMGPU_EINVAL;
    }
  const int64_t n12 = 12 * (int64_t)npoint;
  if (upload(&rq, t->rq, n12) || upload(&rfac, t->rfac, n12) || upload(&rdiv, t->rdiv, n12) ||
      upload(&opoint, t->opoint, nrow_out) || upload(&odir, t->odir, nlink) || upload(&ofac, t->ofac, nlink) ||
      upload(&apoint, t->apoint, nar) || upload(&adir, t->adir, nal))
    return FDMGPU_EHIP;
  for (int v = 0; v < 13; v++) {
    relc.clapl[v] = t->clapl[v];
    relc.cgrad[v] = t->cgrad[v];
  }
  rel = true;
  return 0;
}

int Assembler::init_complex(const double *bi, const double *wi, const double *yi, const double *ds,
                            const double *dsi) {
  if (cx || rel || natome > 0) {
    set_error("fdmgpu_assembler_set_complex: already complex, relativistic or with atoms");
    return FDMGPU_EINVAL;
  }
  if (!bi || !yi || !ds || !dsi || (nlink > 0 && !wi)) {
    set_error("fdmgpu_assembler_set_complex: null table");
    return FDMGPU_EINVAL;
  }
  if (dself) {
    FDM_HIP(hipFree(dself));
    dself = nullptr;
  }
  if (upload(&base_im, bi, nnz) || upload(&ow_im, wi, nlink) || upload(&ylm_im, yi, (int64_t)nb * nlmso) ||
      upload(&dself, ds, 3 * (int64_t)npoint) || upload(&dself_im, dsi, 3 * (int64_t)npoint))
    return FDMGPU_EHIP;
  cx = true;
  return 0;
}

int Assembler::run(int count, const AsmPoints &P, bool e_comp, bool basereel, bool update_only, bool phi_cplx) {
  if (count < 1 || count > max_points) {
    set_error("fdmgpu_assembler_run: count must be in 1.." + std::to_string(max_points));
    return FDMGPU_EINVAL;
  }
  for (int k = 0; k < count; k++)
    if (rel != (P.G[k] != nullptr)) {
      set_error(rel ? "fdmgpu_assembler_run: relativistic rows need the potential's gradient (fdmgpu_assembler_run_rel)"
                    : "fdmgpu_assembler_run_rel: the assembler holds no relativistic tables");
      return FDMGPU_EINVAL;
    }
  const bool cal = e_comp || !basereel || cx;  // complex harmonics or characters make it complex (mat.f90:97-104)
  if (nri == 2) {
    if (!e_comp || basereel) {
      set_error("fdmgpu_assembler_run_ri: the irregular solution needs complex energy and the complex basis");
      return FDMGPU_EINVAL;
    }
    for (int k = 0; k < count; k++)
      if (!P.phi2[k]) {
        set_error("fdmgpu_assembler_run_ri: the irregular solution needs the irregular phiato");
        return FDMGPU_EINVAL;
      }
  }
  if (phi_cplx && !cal) {
    set_error("fdmgpu_assembler_run_atoms: complex phiato needs complex arithmetic (E_comp or !basereel)");
    return FDMGPU_EINVAL;
  }
  const int nso = nligne - nlmso;
  cz *F = (cz *)fun;
  const int NT = 256;
  hipLaunchKernelGGL(k_asm_fun, dim3((nb + NT - 1) / NT, count), dim3(NT), 0, stream, nb, lmaxso, brad, P, (int)e_comp,
                     F);
  const int64_t nrhs_el = (int64_t)nligne * nlmso;
  const int64_t n = nnz > nrhs_el ? nnz : nrhs_el;
  const int gb = (int)std::min<int64_t>((n + NT - 1) / NT, 2048);
  const int gd = (npoint + NT - 1) / NT, go = (nrow_out * nlmso + NT - 1) / NT;
  const int nrun = (nsortf + ASM_CB - 1) / ASM_CB, npr = (nlmso * nlmso + NT - 1) / NT;
  const dim3 gp(nrun, npr, count), gs(npr, count);
  cz *Q = (cz *)part;
  if (cal) {
    if (!update_only)
      hipLaunchKernelGGL(k_asm_base<true>, dim3(gb, count), dim3(NT), 0, stream, nnz, nrhs_el, base, base_im, P);
    hipLaunchKernelGGL(k_asm_diag<true>, dim3(gd, count), dim3(NT), 0, stream, npoint, qdiag, clapl0, dself, dself_im,
                       P, (int)e_comp, (int)rel);
    if (rel)
      hipLaunchKernelGGL(k_asm_rel<true>, dim3(gd, count), dim3(NT), 0, stream, npoint, qdiag, rq, rfac, rdiv, relc, P);
    auto oc = rel ? k_asm_outcols<true, true> : k_asm_outcols<true, false>;
    if (nrow_out > 0)
      hipLaunchKernelGGL(oc, dim3(go, count), dim3(NT), 0, stream, nrow_out, nlmso, nb, lmaxso, orow, oq, optr, ob, ow,
                         ylm, lharm, F, P, (int)basereel, opoint, odir, ofac, relc, ow_im, ylm_im, (int)(nri == 2));
    if (nri == 1) {  // the irregular solution has no outer rows
      hipLaunchKernelGGL(k_asm_outrows_part, gp, dim3(256), 0, stream, nlmso, nb, lmaxso, nsortf, wso, ylm, lharm,
                         F, (int)basereel, Q, ylm_im);
      hipLaunchKernelGGL(k_asm_outrows_sum<true>, gs, dim3(256), 0, stream, nso, nlmso, nsortf, nrun, row_ptr, Q, P);
    }
  } else {
    if (!update_only)
      hipLaunchKernelGGL(k_asm_base<false>, dim3(gb, count), dim3(NT), 0, stream, nnz, nrhs_el, base,
                         (const double *)nullptr, P);
    hipLaunchKernelGGL(k_asm_diag<false>, dim3(gd, count), dim3(NT), 0, stream, npoint, qdiag, clapl0, dself,
                       (const double *)nullptr, P, 0, (int)rel);
    if (rel)
      hipLaunchKernelGGL(k_asm_rel<false>, dim3(gd, count), dim3(NT), 0, stream, npoint, qdiag, rq, rfac, rdiv, relc,
                         P);
    auto oc = rel ? k_asm_outcols<false, true> : k_asm_outcols<false, false>;
    if (nrow_out > 0)
      hipLaunchKernelGGL(oc, dim3(go, count), dim3(NT), 0, stream, nrow_out, nlmso, nb, lmaxso, orow, oq, optr, ob, ow,
                         ylm, lharm, F, P, 1, opoint, odir, ofac, relc, (const double *)nullptr,
                         (const double *)nullptr, 0);
    hipLaunchKernelGGL(k_asm_outrows_part, gp, dim3(256), 0, stream, nlmso, nb, lmaxso, nsortf, wso, ylm, lharm,
                       F, 1, Q, (const double *)nullptr);
    hipLaunchKernelGGL(k_asm_outrows_sum<false>, gs, dim3(256), 0, stream, nso, nlmso, nsortf, nrun, row_ptr, Q, P);
  }
  if (natome > 0) {
    const dim3 gc((nar * nlmax + NT - 1) / NT, count), gr((nrow_atom * nlmax + NT - 1) / NT, count);
    auto cols = rel ? (cal ? (phi_cplx ? k_asm_atomcols<true, true, true> : k_asm_atomcols<true, false, true>)
                           : k_asm_atomcols<false, false, true>)
                    : (cal ? (phi_cplx ? k_asm_atomcols<true, true, false> : k_asm_atomcols<true, false, false>)
                           : k_asm_atomcols<false, false, false>);
    auto rows = cal ? (phi_cplx ? k_asm_atomrows<true, true> : k_asm_atomrows<true, false>)
                    : k_asm_atomrows<false, false>;
    if (nar > 0)
      hipLaunchKernelGGL(cols, gc, dim3(NT), 0, stream, nar, nlmax, nri, aq, aa, aptr, alphi, alw, nlmsa, P, apoint,
                         adir, relc);
    hipLaunchKernelGGL(rows, gr, dim3(NT), 0, stream, nrow_atom, nlmax, nri, arow_atom, arow_lm, bq, nlmsa, nbordf, phoff,
                       poff, yoff, poidsa, ylmato, P);
  }
  FDM_HIP(hipGetLastError());
  return 0;
}

}  // namespace fdmgpu
